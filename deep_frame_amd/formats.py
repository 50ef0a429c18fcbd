"""On-disk artefacts of an optimisation stage, in the reference's formats (NumPy ``.npz``, no pickles).

* ``design.npz``: ``design`` = the flat FP64 design vector (tools/formulation_study.py:1256);
* ``density_half.npz``: ``density`` = the intermediate physical field, float32, in the half-grid shape
  (:1257); readers accept the float64 files older runs wrote;
* ``checkpoint.npz``: ``design``, ``level`` (the last history level) and ``iteration`` (history length),
  written after every optimiser iteration and read back on resume (:1220-1223, :1245-1247);
* ``best_beta{beta:g}.npz``: ``design`` with ``iteration``, ``level``, ``beta``, ``f0``, ``mass_g``,
  ``max_violation`` of the best feasible design per beta (:1225-1228);
* ``design_sha256``: SHA-256 of the contiguous float64 bytes of a design (:1332-1333).
All loads use ``allow_pickle=False``.
"""
import hashlib
from pathlib import Path

import numpy as np

BEST_KEYS = ("iteration", "level", "beta", "f0", "mass_g", "max_violation")


def save_design(path, design):
    np.savez_compressed(path, design=np.asarray(design, dtype=np.float64).ravel())


def load_design(path):
    with np.load(path, allow_pickle=False) as data:
        return np.asarray(data["design"], dtype=np.float64)


def save_density_half(path, physical, shape):
    np.savez_compressed(path, density=np.asarray(physical).reshape(tuple(shape)).astype(np.float32))


def load_density_half(path):
    with np.load(path, allow_pickle=False) as data:
        return np.asarray(data["density"])


def checkpoint(path):
    """Saver ``save(x, history, levels)`` as the optimiser calls it (tools/formulation_study.py:1220-1223)."""
    def save(x, history, levels):
        np.savez_compressed(path, design=x, level=history[-1]["level"], iteration=len(history))
    return save


def load_checkpoint(path):
    """(design, level, iteration) of a checkpoint (tools/formulation_study.py:1245-1247)."""
    with np.load(path, allow_pickle=False) as data:
        return np.asarray(data["design"], dtype=np.float64), int(data["level"]), int(data["iteration"])


def keep_best(out):
    """Saver ``save(x, best)`` writing ``best_beta{beta:g}.npz`` (tools/formulation_study.py:1225-1228)."""
    def save(x, best):
        np.savez_compressed(Path(out) / f"best_beta{best['beta']:g}.npz", design=x, **{key: best[key] for key in BEST_KEYS})
    return save


def design_sha256(design):
    return hashlib.sha256(np.ascontiguousarray(design, dtype=np.float64).tobytes()).hexdigest()
