"""Writes the golden fixtures of this directory from data files of the reference.

Reads (as JSON/text data, nothing is imported or executed):
  docs/validation/formulation_covariance_cantilever.json   dense 3x3 flexibility of the
      32 x 12 x 12 orthotropic cantilever (construction tools/formulation_study.py:405-428)
  deep_frame/config.py:710-716 PRINT_MATERIAL orthotropic constants (copied as numbers below)
  docs/topology_optimization.md "Numerische Pruefungen": Euler-Bernoulli tip deflection
  exports/handoff/neural_r3_v05/density_half.npz   an optimised half-frame density (51 x 96 x 24),
      loaded with allow_pickle=False and rewritten unchanged (key "density", float64)
  exports/handoff/dgx_075mm/run3_r3_v05/density_half.npz   the optimised 0.75 mm half-frame density
      (91 x 170 x 48), same treatment
Run from the repository root with the reference mounted at /root/reference.
"""
import json
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def main():
    cov = json.loads((REF / "docs/validation/formulation_covariance_cantilever.json").read_text())
    fixture = {
        "source": "docs/validation/formulation_covariance_cantilever.json",
        "shape_full": [32, 12, 12], "origin_mm": [0.0, -6.0, 0.0], "density_seed": 11,
        "density": "default_rng(11).uniform(0.3, 1.0, (32, 6, 12)) mirrored about y: concat(d[:, ::-1], d)",
        "simp": {"penal": 3.0, "eps": 1e-6},
        "orthotropic": {"e_xy_mpa": 4430.0, "e_z_mpa": 2170.0, "nu_xy": 0.30, "nu_xz": 0.30,
                        "g_xy_mpa": 1703.8, "g_z_mpa": 834.6},
        "loads": [["tip", "Fz"], ["mid", "Fz"], ["tip", "Fy"]],
        "interfaces": {"tip": {"box": [[32, -6, 0], [32, 6, 12]], "reference_mm": [32, 0, 6]},
                       "mid": {"box": [[15, -6, 12], [17, 6, 12]], "reference_mm": [16, 0, 12]}},
        "sigma": cov["sigma"],
        "dense_flexibility": cov["dense_flexibility"],
        "dense_full_model": cov["dense_full_model"],
    }
    (OUT / "covariance_cantilever.json").write_text(json.dumps(fixture, indent=1))
    text = (REF / "docs/topology_optimization.md").read_text(encoding="utf-8")
    assert "0,0730389 mm" in text and "867,609 Hz" in text
    beam = {"source": "docs/topology_optimization.md (Numerische Pruefungen)", "shape": [24, 4, 4],
            "spacing_mm": [2.0, 1.5, 1.5], "young_mpa": 4430.0, "poisson": 0.3, "force_n": -1.0,
            "tip_deflection_mm": 0.0730389, "euler_bernoulli_mm": 0.0770504}
    (OUT / "euler_bernoulli.json").write_text(json.dumps(beam, indent=1))
    with np.load(REF / "exports/handoff/neural_r3_v05/density_half.npz", allow_pickle=False) as data:
        density = np.asarray(data["density"])
    assert density.shape == (51, 96, 24) and density.dtype == np.float64
    np.savez_compressed(OUT / "density_half_neural_r3_v05.npz", density=density)
    with np.load(REF / "exports/handoff/dgx_075mm/run3_r3_v05/density_half.npz", allow_pickle=False) as data:
        density = np.asarray(data["density"])
    assert density.shape == (91, 170, 48) and density.dtype == np.float64
    np.savez_compressed(OUT / "density_half_075mm_run3.npz", density=density)


if __name__ == "__main__":
    main()
