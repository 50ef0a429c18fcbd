"""On-disk artefacts (deep_frame_amd/formats.py) in the reference's formats (tools/formulation_study.py:
1220-1228, 1245-1263, 1332-1333): keys, dtypes and shapes, round trips, and the real half-frame density
export the reference ships (exports/handoff/neural_r3_v05/density_half.npz, copied as a fixture)."""
import hashlib
from pathlib import Path

import numpy as np

from deep_frame_amd import formats

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_design_and_density_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    design = rng.uniform(0, 1, 6 * 4 * 5)
    formats.save_design(tmp_path / "design.npz", design.reshape(6, 4, 5))
    with np.load(tmp_path / "design.npz", allow_pickle=False) as data:
        assert data.files == ["design"] and data["design"].dtype == np.float64 and data["design"].shape == (120,)
    assert np.array_equal(formats.load_design(tmp_path / "design.npz"), design)
    formats.save_density_half(tmp_path / "density_half.npz", design, (6, 4, 5))
    with np.load(tmp_path / "density_half.npz", allow_pickle=False) as data:
        assert data.files == ["density"] and data["density"].dtype == np.float32 and data["density"].shape == (6, 4, 5)
    assert np.array_equal(formats.load_density_half(tmp_path / "density_half.npz"),
                          design.reshape(6, 4, 5).astype(np.float32))


def test_checkpoint_and_best(tmp_path):
    design = np.linspace(0, 1, 17)
    history = [{"level": 0}, {"level": 1}, {"level": 1}]
    formats.checkpoint(tmp_path / "checkpoint.npz")(design, history, None)
    with np.load(tmp_path / "checkpoint.npz", allow_pickle=False) as data:
        assert sorted(data.files) == ["design", "iteration", "level"]
    x, level, iteration = formats.load_checkpoint(tmp_path / "checkpoint.npz")
    assert np.array_equal(x, design) and level == 1 and iteration == 3
    best = {"iteration": 7, "level": 2, "beta": 4.0, "f0": 0.31, "mass_g": 12.5, "max_violation": -0.01}
    formats.keep_best(tmp_path)(design, best)
    with np.load(tmp_path / "best_beta4.npz", allow_pickle=False) as data:
        assert sorted(data.files) == sorted(("design",) + formats.BEST_KEYS)
        assert float(data["mass_g"]) == 12.5 and int(data["iteration"]) == 7


def test_design_sha256_is_of_float64_bytes():
    design = np.arange(5, dtype=np.float32)
    assert formats.design_sha256(design) == hashlib.sha256(np.arange(5, dtype=np.float64).tobytes()).hexdigest()


def test_reference_half_density_export_loads():
    density = formats.load_density_half(GOLDEN / "density_half_neural_r3_v05.npz")
    assert density.shape == (51, 96, 24) and density.dtype == np.float64
    assert density.min() >= 0 and density.max() <= 1 and 0.02 < density.mean() < 0.05
