"""Synthetic beam batches for throughput runs and size-independent checks.

Beams are resampled from a set of real prepared beams (for example a fixture
written from the reference's tracer): each synthetic beam is a copy of a real
one, rigidly rotated (every 3-D vector of the record -- aperture axes, beam
direction, centre, polarisation vectors, normal -- by the same random
rotation) and uniformly scaled (aperture vertices and centre by s, area by
s^2).  The aperture polygons, vertex/edge statistics and Jones matrices stay
those of the real beam set, so the per-(beam, direction) work matches a traced
particle's; only the scattering physics of the batch as a whole is synthetic.
"""
import numpy as np

from . import abi


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def resample(source, beams_per_orient, scale=1.0, seed=0, weights=None):
    """Batch with len(beams_per_orient) orientations built from `source`
    (a BeamBatch).  Orientation o holds beams_per_orient[o] beams."""
    rng = np.random.default_rng(seed)
    counts = np.asarray(beams_per_orient, dtype=np.int64)
    n = int(counts.sum())
    pick = rng.integers(0, source.n_beams, n)
    R = _random_rotations(rng, n)
    bs = source.beam_scalars[pick].copy()
    for col in (abi.B_HOR, abi.B_VER, abi.B_DIR, abi.B_CEN, abi.B_NT, abi.B_NP, abi.B_NXDT,
                abi.B_NXDP, abi.B_NORMAL):
        bs[:, col:col + 3] = np.einsum("nij,nj->ni", R, bs[:, col:col + 3])
    bs[:, abi.B_CEN:abi.B_CEN + 3] *= scale
    bs[:, abi.B_AREA] *= scale * scale
    nv = source.beam_nv[pick]
    vo = source.vert_offsets
    starts = vo[pick]
    rel = np.arange(int(nv.sum())) - np.repeat(np.cumsum(nv) - nv, nv)
    verts = np.repeat(starts, nv) + rel
    new_vo = np.concatenate([[0], np.cumsum(nv)])
    if weights is None:
        weights = np.full(len(counts), 1.0 / max(len(counts), 1))
    return abi.BeamBatch(np.concatenate([[0], np.cumsum(counts)]), weights, bs, nv,
                         source.beam_external[pick], new_vo, source.vx[verts] * scale,
                         source.vy[verts] * scale, source.slope_yx[verts],
                         source.slope_xy[verts], source.edge_flags[verts])


def concat(batches, weights=None):
    """One batch holding the orientations of `batches` in order (weights
    their own unless given)."""
    oo, vo = [np.zeros(1, np.int64)], [np.zeros(1, np.int64)]
    nb = nvert = 0
    for b in batches:
        oo.append(b.orient_offsets[1:] + nb)
        vo.append(b.vert_offsets[1:] + nvert)
        nb += b.n_beams
        nvert += int(b.vert_offsets[-1])
    cat = lambda k: np.concatenate([getattr(b, k) for b in batches])  # noqa: E731
    w = cat("weights") if weights is None else weights
    return abi.BeamBatch(np.concatenate(oo), w, cat("beam_scalars"), cat("beam_nv"),
                         cat("beam_external"), np.concatenate(vo), cat("vx"), cat("vy"),
                         cat("slope_yx"), cat("slope_xy"), cat("edge_flags"))


def config2_grid():
    """theta 0..10 deg at 0.05 deg and 10.5..180 deg at 0.5 deg (541 rows) x
    96 phi -- the non-uniform grid of the orientation-average workload."""
    th = np.radians(np.r_[np.round(np.arange(201) * 0.05, 10), 10.5 + 0.5 * np.arange(340)])
    return abi.Grid.build(th, 96)


def uniform_grid(n_zen, n_az, theta0_deg=0.0, theta1_deg=180.0):
    th = np.radians(theta0_deg) + np.arange(n_zen + 1) * (
        np.radians(theta1_deg - theta0_deg) / n_zen)
    return abi.Grid.build(th, n_az)


def take_beams(batch, orient, n):
    """One-orientation batch holding the first n beams of orientation `orient`
    (weight 1)."""
    b0 = int(batch.orient_offsets[orient])
    b1 = min(int(batch.orient_offsets[orient + 1]), b0 + int(n))
    beams = np.arange(b0, b1)
    nv = batch.beam_nv[beams]
    vo = batch.vert_offsets
    verts = np.arange(vo[b0], vo[b1])
    return abi.BeamBatch([0, len(beams)], [1.0], batch.beam_scalars[beams], nv,
                         batch.beam_external[beams], vo[b0:b1 + 1] - vo[b0],
                         batch.vx[verts], batch.vy[verts], batch.slope_yx[verts],
                         batch.slope_xy[verts], batch.edge_flags[verts])
