"""The C-ABI library loads without a GPU and exports every symbol include/dfmi.h declares."""
import re
from pathlib import Path

import pytest

from deep_frame_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "dfmi.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(dfmi_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.library()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert set(names) == set(_native.SIGNATURES)


def test_default_settings_mirror_reference_multigrid_defaults():
    s = _native.default_settings()
    assert (s.coarsest_dofs, s.max_levels, s.chebyshev_degree, s.power_iterations, s.max_iterations) == (80000, 8, 3, 20, 2000)
    assert (s.chebyshev_ratio, s.growth, s.tolerance, s.dead_ratio) == (20.0, 2.0, 1e-8, 1e-12)
    assert s.retries == 2  # MULTIGRID["retries"], topology_multigrid.py:9
    assert s.fine_smoother == 0 and s.precision == 1


def test_invalid_arguments_raise_without_touching_a_gpu():
    import ctypes
    bad = (ctypes.c_int32 * 3)(0, 4, 4)
    with pytest.raises(RuntimeError, match="dfmi_create failed"):
        _native.call("dfmi_create", bad, (ctypes.c_double * 3)(1, 1, 1), (ctypes.c_double * 576)(),
                     None, ctypes.byref(ctypes.c_void_p()))
