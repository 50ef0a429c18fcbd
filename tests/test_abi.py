"""The C-ABI library loads without a GPU and exports every symbol include/dfmi.h declares."""
import re
from pathlib import Path

import pytest

from deep_frame_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "dfmi.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(dfmi_[a-z_0-9]+)\s*\(", text)))


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
    assert s.coarse_krylov == 0 and s.block_krylov == 0  # the reference's V-cycle and per-column PCG
    # MULTIGRID smoother "chebyshev", sweeps 2, damping 1.0, cycle "V", coarse "galerkin" (topology_multigrid.py:8)
    assert (s.smoother, s.sweeps, s.damping, s.cycle, s.coarse) == (0, 2, 1.0, 0, 0)


def test_invalid_arguments_raise_without_touching_a_gpu():
    import ctypes
    bad = (ctypes.c_int32 * 3)(0, 4, 4)
    with pytest.raises(RuntimeError, match="dfmi_create failed"):
        _native.call("dfmi_create", bad, (ctypes.c_double * 3)(1, 1, 1), (ctypes.c_double * 576)(),
                     None, ctypes.byref(ctypes.c_void_p()))


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of dfmi_settings and dfmi_report have the C compiler's field offsets and sizes."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dfmi.h"', "int main(void) {"]
    for struct, mirror in (("dfmi_settings", _native.Settings), ("dfmi_report", _native.Report)):
        lines.append(f'  printf("{struct} sizeof %zu\\n", sizeof({struct}));')
        for name, _ in mirror._fields_:
            lines.append(f'  printf("{struct} {name} %zu\\n", offsetof({struct}, {name}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for struct, mirror in (("dfmi_settings", _native.Settings), ("dfmi_report", _native.Report)):
        assert got[(struct, "sizeof")] == __import__("ctypes").sizeof(mirror)
        for name, _ in mirror._fields_:
            assert got[(struct, name)] == getattr(mirror, name).offset, (struct, name)
