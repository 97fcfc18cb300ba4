"""Error measures shared by the parity tests.

cell       max |dM_ij| / M11 of the same cell.  |M_ij| <= M11 for any sum of
           pure Mueller matrices, so this is a relative error per element.
row        max |dM_ij| / max over phi of M11 in the same theta row.
grid_l2    sqrt(sum w (dM)^2 / sum w M^2) over all cells and 16 elements, w =
           2*pi*dcos of the cell's theta row: the norm of the reference's own
           CPU-vs-GPU release gate (tests/run_cuda_release_gate.sh:57-87), here
           applied before the azimuth average.
gate_l2    the same norm on azimuth-averaged rows [theta, 2*pi*dcos, M11..M44]
           -- exactly the reference's gate.
convergence_errors
           per-element errors of azimuth-averaged rows against a converged
           reference: M11 as the 2*pi*dcos-weighted relative L2, every other
           element as the 2*pi*dcos-weighted RMS of P_ij/P11 differences.
plot_element_errors
           the per-element measure of the reference's orientation-convergence
           plot (scripts/plot_po_orientation_sampling_convergence.py:79-84):
           unweighted RMS over theta rows of M11/M11_ref - 1 and of the
           P_ij/P11 differences.
"""
import numpy as np


def cell(a, b):
    s = np.abs(b[..., :1])
    return float(np.max(np.abs(a - b) / np.where(s > 0, s, 1.0)))


def row(a, b):
    r = np.abs(b[..., 0]).max(axis=0)[None, :, None]
    return float(np.max(np.abs(a - b) / np.where(r > 0, r, 1.0)))


def two_pi_dcos(theta):
    th = np.asarray(theta, dtype=np.float64)
    mid = 0.5 * (th[1:] + th[:-1])
    lo = np.r_[th[0], mid]
    hi = np.r_[mid, th[-1]]
    return 2.0 * np.sin(0.5 * (lo + hi)) * np.sin(0.5 * (hi - lo)) * 2.0 * np.pi


def grid_l2(a, b, theta):
    w = two_pi_dcos(theta)[None, :, None]
    return float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b ** 2)))


def gate_l2(rows_a, rows_b):
    w = rows_b[:, 1:2]
    d = rows_a[:, 2:] - rows_b[:, 2:]
    return float(np.sqrt(np.sum(w * d * d) / np.sum(w * rows_b[:, 2:] ** 2)))


def p11_rel(a, b):
    return float(np.max(np.abs(a[..., 0] - b[..., 0]) / np.abs(b[..., 0])))


def summary(a, b, theta):
    q = np.abs(a - b) / np.maximum(np.abs(b[..., :1]), 1e-300)
    return ("cell %.2e row %.2e grid_l2 %.2e (cell quantiles 50/99/99.9%%: %.1e %.1e %.1e)"
            % (cell(a, b), row(a, b), grid_l2(a, b, theta), *np.quantile(q, [0.5, 0.99, 0.999])))


def convergence_errors(rows, ref):
    """16 errors (fractions) of rows [n_theta, 18] against ref [n_theta, 18]."""
    w = ref[:, 1] / np.sum(ref[:, 1])
    m, r = rows[:, 2:], ref[:, 2:]
    e = np.empty(16)
    e[0] = np.sqrt(np.sum(w * (m[:, 0] - r[:, 0]) ** 2) / np.sum(w * r[:, 0] ** 2))
    d = m[:, 1:] / m[:, :1] - r[:, 1:] / r[:, :1]
    e[1:] = np.sqrt(np.sum(w[:, None] * d * d, axis=0))
    return e


def plot_element_errors(rows, ref):
    """16 errors (fractions) in the form of the reference's convergence plot."""
    m, r = rows[:, 2:], ref[:, 2:]
    e = np.sqrt(np.mean((m / m[:, :1] - r / r[:, :1]) ** 2, axis=0))
    e[0] = np.sqrt(np.mean((m[:, 0] / r[:, 0] - 1.0) ** 2))
    return e
