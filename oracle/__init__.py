"""CPU oracle for the SIMP inner loop (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  Every function restates, in plain FP64 NumPy/SciPy, the algorithm
of the reference file:line it cites, and is pinned by the checks of
``tests/test_oracle.py`` (golden vectors, known answers, closed forms, finite differences).

Parity status per module:
  hex8      pinned (rigid modes, patch energy, Euler-Bernoulli KAT, dense
            flexibility golden vector of the covariance cantilever)
  system    pinned (same golden vectors; SuperLU residual <= 1e-12)
  gmg       pinned (R = P^T, Galerkin = R A P by brute-force dense products; fcg: two steps
            equal the Galerkin projection onto span{Vg, VAVg}, many steps the exact level solve)
  filters   pinned (brute-force neighbour sums, adjoint identity, FD)
  oc        pinned (bisection invariants, volume target, move limits)
  mma       parity unpinned against mmapy 0.3.1 (absent); pinned against the
            KKT conditions of Svanberg's subproblem and a second, independent
            solver of the same subproblem
  loads     pinned (sum f = F, sum r x f = M exactly; min-norm property)
  covariance pinned (golden load_mean / load_worst of the covariance cantilever;
            central differences of the mean and KS-worst gradients)
  modal     pinned (documented first frequency 867.609 Hz of the clamped beam;
            central differences of the KS aggregate on both mass branches)
  floating  pinned (3-2-1 support: equals the pinned solve of the unadjusted forces;
            line fixture: equilibrium on the free DOFs and orthogonality to the kernel)
"""
