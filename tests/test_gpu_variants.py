"""Batched uneven design variants: every variant converges to rel-res 1e-8 and its compliance and
sensitivity match FP64 (the library's FP64 solve for all variants, the oracle's SuperLU for one)."""
import numpy as np
import pytest

from deep_frame_amd import element, problems
from oracle import hex8, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("streams", [1, 3])
def test_variant_batch_matches_fp64(streams, torch_gpu):
    from deep_frame_amd import variants
    vs = variants.generate(count=3, seed=5, nxy=(16, 28), nz=(6, 10), rhs=(6, 12))
    print("variants", vs, "streams", streams, flush=True)
    out = variants.run(vs, check=True, streams=streams)
    for rec in out["variants"]:
        assert rec["converged"] and rec["max_rel_res"] < 1e-8
        assert rec["compliance_rel_err"] < 1e-6
        assert rec["sensitivity_rel_err"] < 1e-5
    assert out["dof_rhs_per_s"] > 0
    # one variant against the CPU oracle
    v = vs[0]
    prob = problems.frame(v["shape"], seed=v["seed"], rhs=v["rhs"])
    mod, dmod = problems.moduli(prob["density"])
    pts, conn, dofs = hex8.grid(v["shape"])
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    u, _ = system.solve(dofs, conn, ke, mod, prob["fixed"], prob["forces"], len(pts))
    comp, sens = system.compliance_sensitivity(dofs, ke, u, prob["forces"], dmod)
    inp = variants.build(v)
    c, s, reps = variants.solve_one(inp, variants.MIXED, ke)
    assert np.allclose(c.cpu().numpy(), comp, rtol=1e-6, atol=0)
    ref = sens.sum(axis=1)
    assert np.linalg.norm(s.cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
