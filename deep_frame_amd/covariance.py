"""Interface load-covariance measures on device-resident unit-wrench solutions.

Mirrors ``deep_frame/topology_problem.py`` ``InterfaceCovariance`` (:573-684), the direct consumer of
the 42 unit-wrench solves of the frame (7 interfaces x 6 DOF):

  * the load directions D from the covariance Sigma (:583-591): eigen-decomposition, descending, rank
    by a 1e-10 relative cut, D = V sqrt(Lambda);
  * ``matrix`` (:639-646): the flexibility of the load cases, factor D^T (W^T U) D symmetrised, where W
    holds the unit-wrench forces and U their solutions; W^T U is one device pass (``dfmi_gram``);
  * ``measure`` (:657-684): load_mean = tr F and load_worst = lambda_max(F) with the KS aggregate, and,
    when their limits are set, their gradients: the case compliance sensitivities summed over the cases
    (mean) and sum_i w_i energy(U D v_i) over the KS weights above ks_cutoff (worst).  The combined fields
    U M are one device pass (``dfmi_combine``) and their element energies the library's
    compliance/sensitivity kernel.

The 42 x 42 eigenproblem and the KS weights are host arithmetic on the small matrix, as in the
reference.  Interface bodies (``body_term``) and calibration tables are not part of this build.
"""
import numpy as np

from deep_frame_amd import _native

COVARIANCE = {"ks": 50.0, "ks_cutoff": 1e-9}


def load_directions(sigma):
    """(D (n, rank), rank) with D D^T = Sigma on its range (topology_problem.py:583-591)."""
    sigma = np.asarray(sigma, dtype=float)
    values, vectors = np.linalg.eigh((sigma + sigma.T) / 2)
    values, vectors = values[::-1], vectors[:, ::-1]
    if values[-1] < -1e-9 * values[0]:
        raise ValueError("Load covariance is not positive semidefinite")
    rank = int(np.sum(values > 1e-10 * values[0]))
    return vectors[:, :rank] * np.sqrt(values[:rank]), rank


class InterfaceCovariance:
    def __init__(self, sigma, factor=1.0, settings=None):
        import torch
        self.torch = torch
        self.settings = {**COVARIANCE, **(settings or {})}
        self.sigma = np.asarray(sigma, dtype=float)
        self.directions, self.rank = load_directions(self.sigma)
        self.factor = float(factor)
        self._work = None

    def unit_flexibility(self, u, f):
        """W^T U for k unit-wrench forces f and solutions u, both (k, ndof) FP64 device tensors -> (k, k) host."""
        torch = self.torch
        if u.shape != f.shape or u.shape[0] != self.sigma.shape[0]:
            raise ValueError("u and f must both be (k, ndof) with k = the covariance size")
        u, f = u.contiguous(), f.contiguous()
        k, n = u.shape
        blocks = 512
        if self._work is None or self._work.numel() < blocks * k * k:
            self._work = torch.empty(blocks * k * k, dtype=torch.float64, device=u.device)
        out = torch.empty((k, k), dtype=torch.float64, device=u.device)
        _native.call("dfmi_gram", _native.ptr(f), _native.ptr(u), n, k, k, _native.ptr(out), _native.ptr(self._work),
                     blocks, _native.stream_handle(torch, u.device))
        return out.cpu().numpy()

    def matrix(self, u, f):
        """Case flexibility factor D^T (W^T U) D, symmetrised (topology_problem.py:639-646)."""
        flex = self.factor * self.directions.T @ self.unit_flexibility(u, f) @ self.directions
        return (flex + flex.T) / 2

    def fields(self, u, mat):
        """U M for solutions u (k, ndof) and M (k, m) -> (m, ndof) device tensor."""
        torch = self.torch
        mat = np.ascontiguousarray(mat, dtype=float)
        m_dev = torch.as_tensor(mat, device=u.device)
        y = torch.empty((mat.shape[1], u.shape[1]), dtype=torch.float64, device=u.device)
        _native.call("dfmi_combine", _native.ptr(u.contiguous()), u.shape[1], u.shape[0], _native.ptr(m_dev),
                     mat.shape[1], _native.ptr(y), 0, _native.stream_handle(torch, u.device))
        return y

    def energy(self, mg, u, mat, dmoduli):
        """-factor E' sum_m (U M_m)_e^T Ke (U M_m)_e per element (topology_problem.py:647-652), device tensor."""
        y = self.fields(u, mat)
        _, sens = mg.compliance_sensitivity_device(y, y, dmoduli, self.factor)
        return sens

    def measure(self, mg, u, f, dmoduli, limits=None):
        """Rows load_mean and load_worst as topology_problem.py:657-684 (values, limits, g and gradients)."""
        limits = {"mean_n_mm": None, "worst_n_mm": None, **(limits or {})}
        s = self.settings["ks"]
        flex = self.matrix(u, f)
        mean = float(np.trace(flex))
        values, vectors = np.linalg.eigh(flex)
        values, vectors = values[::-1], vectors[:, ::-1]
        worst = float(values[0])
        scale = limits["worst_n_mm"] or worst
        weights = np.exp(s * (values - worst) / scale)
        ks = worst + scale * np.log(weights.sum()) / s
        weights /= weights.sum()
        info = {"eigenvalues_n_mm": values[:6].tolist(), "ks_n_mm": float(ks), "ks": s, "rank": self.rank}
        rows = [{"name": "load_mean", "value": mean, "limit": limits["mean_n_mm"], "unit": "N mm", "sense": "<=",
                 "g": None, "gradient": None},
                {"name": "load_worst", "value": worst, "limit": limits["worst_n_mm"], "unit": "N mm", "sense": "<=",
                 "g": None, "gradient": None, "info": info}]
        if limits["mean_n_mm"]:
            gradient = self.energy(mg, u, self.directions, dmoduli)
            rows[0].update(g=mean / limits["mean_n_mm"] - 1, gradient=gradient / limits["mean_n_mm"])
        if limits["worst_n_mm"]:
            keep = np.flatnonzero(weights > self.settings["ks_cutoff"])
            mat = self.directions @ (vectors[:, keep] * np.sqrt(weights[keep]))
            gradient = self.energy(mg, u, mat, dmoduli)
            rows[1].update(g=float(ks / limits["worst_n_mm"] - 1), gradient=gradient / limits["worst_n_mm"])
        return rows
