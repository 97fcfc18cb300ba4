"""Particle files for the reference CLI runs of the tests and fixtures."""
import numpy as np


def write_column_aggregate(path, n_columns=8, seed=7, d_range=(5.0, 30.0), ld_range=(1.0, 4.0)):
    """Writes an aggregate of hexagonal columns in the reference's particle
    file format (README.md "strict text format": concave flag, '1 8' =
    aggregate of 8-facet components, symmetry 180 360, blank-line separated
    facets wound with outward normals).  Column i: diameter D ~ U(d_range)
    across corners, length D * U(ld_range), axis along x, y or z with a
    random spin about it, centre placed at random with overlapping bounding
    spheres rejected.  Axes stay on the coordinate axes because the loader
    snaps coordinates to 2^-40 of the particle's extent
    (src/particle/Particle.cpp:796-806) and then checks each facet's
    planarity to 2^-40 of the facet's own size (:116-160): a hexagon at a
    constant axis coordinate and a side rectangle parallel to the axis stay
    exactly planar under that snapping, a tilted facet does not.  Returns
    the bounding radius of every column."""
    rng = np.random.default_rng(seed)
    cols = []
    while len(cols) < n_columns:
        d = rng.uniform(*d_range)
        length = d * rng.uniform(*ld_range)
        r_bound = np.hypot(0.5 * d, 0.5 * length)
        axis = int(rng.integers(3))
        spin = rng.uniform(0.0, np.pi / 3)
        spread = 0.5 * (40.0 + 10.0 * len(cols))
        for _ in range(1000):
            c = rng.uniform(-spread, spread, 3)
            if all(np.linalg.norm(c - c2) > r_bound + r2 + 0.5 for c2, r2, _ in cols):
                break
        else:
            continue
        ang = spin + np.arange(6) * np.pi / 3
        ring = np.stack([0.5 * d * np.cos(ang), 0.5 * d * np.sin(ang)], 1)
        top = np.c_[ring, np.full(6, 0.5 * length)]
        bot = np.c_[ring, np.full(6, -0.5 * length)]
        facets = [top, bot[::-1]] + [np.array([bot[k], bot[(k + 1) % 6], top[(k + 1) % 6], top[k]])
                                     for k in range(6)]
        perm = [(1, 2, 0), (2, 0, 1), (0, 1, 2)][axis]  # cyclic, so the winding stays outward
        cols.append((c, r_bound, [f[:, perm] + c for f in facets]))
    with open(path, "w") as f:
        f.write("1\n1 8\n180 360\n")
        for _, _, facets in cols:
            for fc in facets:
                f.write("\n" + "".join("%.17g %.17g %.17g\n" % tuple(v) for v in fc))
    return [r for _, r, _ in cols]
