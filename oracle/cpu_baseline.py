"""The reference's CPU solve path, timed (test and benchmark infrastructure only).

Restates ``HexElasticity(linear_solver="cpu_superlu")`` (deep_frame/topology_optimization.py:638-650 assembly,
:651-734 solve, compliance and sensitivity, :758-759 the SuperLU call) on the synthetic workloads: CSR K(rho)
reduced to the free DOFs, ``splu(K_ff, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True})``, the
k right-hand sides, compliance f.u and the sensitivity -E'(rho) sum_k u_e^T Ke u_e.

Run as ``python -m oracle.cpu_baseline <workload> <threads>`` it prints one JSON line with the wall time split
into assembly, factorisation, the k-RHS solve and compliance + sensitivity, the peak RSS and the largest
relative residual.  The benchmark starts it in a child process so that it can hold it to a wall-clock and a
memory cap.  Workloads: ``cantilever`` (48 x 24 x 24, 1 RHS), ``frame:NX,NY,NZ`` (frame family, 42 RHS).
"""
import json
import os
import sys
import time

import numpy as np


def peak_rss_gb():
    """High-water resident set of this process (VmHWM: reset by exec, unlike ru_maxrss which a forked child
    inherits from its parent)."""
    with open("/proc/self/status") as status:
        for line in status:
            if line.startswith("VmHWM:"):
                return int(line.split()[1]) / 2 ** 20
    return None


def workload(name):
    from deep_frame_amd import problems
    if name == "cantilever":
        return problems.cantilever(seed=0)
    shape = tuple(int(v) for v in name.split(":")[1].split(","))
    return problems.frame(shape, seed=0, rhs=42)


def run(name, threads):
    from scipy.sparse.linalg import splu
    from threadpoolctl import threadpool_limits

    from deep_frame_amd import element, problems
    from oracle import hex8, system

    prob = workload(name)
    shape = prob["shape"]
    ke = element.stiffness((1, 1, 1), element.isotropic(0.3))
    mod, dmod = problems.moduli(prob["density"])
    pts, conn, dofs = hex8.grid(shape)
    ndof = 3 * len(pts)
    forces = np.asarray(prob["forces"], float).reshape(ndof, -1)
    times = {}
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        k = hex8.assemble(dofs, ke, mod, ndof)
        free = system.free_dofs(conn, mod, prob["fixed"], len(pts))
        kff = k[free][:, free].tocsc()
        times["assembly_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        lu = splu(kff, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True})
        times["factor_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        u = np.zeros_like(forces)
        u[free] = lu.solve(forces[free])
        times["solve_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        system.compliance_sensitivity(dofs, ke, u, forces, dmod)
        times["compliance_sensitivity_s"] = time.perf_counter() - t0
    rel = np.linalg.norm(kff @ u[free] - forces[free], axis=0) / np.linalg.norm(forces[free], axis=0)
    return {"workload": name, "shape": list(shape), "dof": ndof, "rhs": forces.shape[1], "threads": threads,
            "seconds": sum(times.values()), **times, "max_rel_res": float(rel.max()),
            "peak_rss_gb": peak_rss_gb()}


if __name__ == "__main__":
    threads = int(sys.argv[2])
    for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[var] = str(threads)
    print(json.dumps(run(sys.argv[1], threads)), flush=True)
