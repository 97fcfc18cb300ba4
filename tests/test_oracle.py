"""Pins the FP64 CPU oracle: reference fixtures first, then closed forms.

The fixtures under tests/golden hold the reference's own prepared beams and
the Mueller grids its CPU routines (and its unmodified CLI) produced for them.
"""
import numpy as np
import pytest

import fourier
from conftest import load_fixture


def rel_to_m11(a, b):
    scale = np.abs(b[..., :1])
    return float(np.max(np.abs(a - b) / np.where(scale > 0, scale, 1.0)))


def _setup(pkg, fx):
    abi = pkg.abi
    return (abi.Grid.from_fixture(fx), abi.BeamBatch.from_fixture(fx),
            abi.consts_from_fixture(fx))


# The reference CPU path evaluates every phase with a polynomial whose cosine
# is off by up to 6.3e-9 (HandlerPO_fast.h:116-146); coherent sums of hundreds
# of beams amplify that in cells where the sum cancels.  With the same
# polynomial the oracle reproduces the fixtures to rounding (TIGHT); with libm
# it agrees to the reference's own sincos error (LOOSE, relative to the local
# M11; the largest-M11 cells agree far better, see the weighted L2 check).
TIGHT = 1e-12
LOOSE = 2e-5


def weighted_l2(a, b):
    """sqrt(sum (a-b)^2 / sum b^2) over cells and elements."""
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("name", ["config1_fixed", "fixed_absorbing"])
def test_fixed_orientation_matches_reference(pkg, oracle, fixtures, name):
    fx = fixtures[name]
    grid, batch, consts = _setup(pkg, fx)
    rows = fx.get("kept_theta_rows")
    with oracle.ref_sincos():
        m, m_ns, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    ma, _, _ = oracle.diffract(grid, batch, consts)
    if rows is not None:
        m, m_ns, ma = m[:, rows], m_ns[:, rows], ma[:, rows]
    # against the reference CPU routine on the same prepared beams
    assert rel_to_m11(m, fx["mueller_full"]) < TIGHT
    assert rel_to_m11(m_ns, fx["mueller_noshadow"]) < TIGHT
    # against the unmodified reference binary's fixed-orientation .dat
    assert rel_to_m11(m, fx["mueller_fixed_dat"]) < TIGHT
    assert rel_to_m11(ma, fx["mueller_full"]) < LOOSE
    assert weighted_l2(ma, fx["mueller_full"]) < 2e-7


@pytest.mark.parametrize("name", ["sobol_average", "config5_near_forward"])
def test_sobol_average_matches_reference(pkg, oracle, fixtures, name):
    """Orientation averages: the small-particle average on the non-uniform
    grid, and the large-particle near-forward case (D=200 um, L=400 um,
    m=1.3116+1e-3i, 12 reflections, theta 0-5 deg)."""
    fx = fixtures[name]
    grid, batch, consts = _setup(pkg, fx)
    with oracle.ref_sincos():
        m, m_ns, per = oracle.diffract(grid, batch, consts, noshadow=True, per_orient=True)
    ma, _, _ = oracle.diffract(grid, batch, consts)
    assert rel_to_m11(ma, fx["mueller_full"]) < LOOSE
    assert weighted_l2(ma, fx["mueller_full"]) < 2e-7
    assert rel_to_m11(m, fx["mueller_full"]) < TIGHT
    assert rel_to_m11(m_ns, fx["mueller_noshadow"]) < TIGHT
    np.testing.assert_allclose(per.sum(axis=0), m, rtol=0, atol=1e-12 * np.abs(m).max())
    # azimuth average, poles and 2*pi*dcos vs the reference CLI's .dat
    for mm, ref in ((m, fx["ref_total_dat"]), (m_ns, fx["ref_total_dat_noshadow"])):
        avg = oracle.azimuth_average(grid, mm)
        np.testing.assert_allclose(np.degrees(avg[:, 0]), ref[:, 0], rtol=1e-14, atol=1e-12)
        np.testing.assert_allclose(avg[:, 1], ref[:, 1], rtol=1e-12, atol=0)
        assert rel_to_m11(avg[:, 2:], ref[:, 2:]) < TIGHT


def test_shadow_beam_is_the_only_difference(pkg, oracle, fixtures):
    """No-shadow output drops exactly the external beams (HandlerPO.cpp:2780-2783)."""
    fx = fixtures["sobol_average"]
    grid, batch, consts = _setup(pkg, fx)
    keep = np.where(batch.beam_external == 0)[0]
    assert len(keep) < batch.n_beams
    j_full, j_ns = oracle.orientation_jones(grid, batch, consts, 0, noshadow=True)
    # rebuild orientation 0 without its external beams
    o0 = batch.subset([0])
    internal = np.where(o0.beam_external == 0)[0]
    vo = o0.vert_offsets
    verts = np.concatenate([np.arange(vo[b], vo[b + 1]) for b in internal])
    nv = o0.beam_nv[internal]
    b2 = pkg.abi.BeamBatch([0, len(internal)], o0.weights, o0.beam_scalars[internal], nv,
                           o0.beam_external[internal], np.concatenate([[0], np.cumsum(nv)]),
                           o0.vx[verts], o0.vy[verts], o0.slope_yx[verts], o0.slope_xy[verts],
                           o0.edge_flags[verts])
    j2, _ = oracle.orientation_jones(grid, b2, consts, 0)
    np.testing.assert_allclose(j2, j_ns, rtol=0, atol=1e-13 * np.abs(j_full).max())
    assert np.abs(j_full - j_ns).max() > 1e-3 * np.abs(j_full).max()


# ---------------------------------------------------------------- closed forms
def _poly_edges(x, y, tol=1e-12):
    n = len(x)
    syx = np.zeros(n)
    sxy = np.zeros(n)
    fl = np.zeros(n, np.uint8)
    for i in range(n):
        j = (i + 1) % n
        dx, dy = x[j] - x[i], y[j] - y[i]
        if abs(dx) >= tol:
            fl[i] |= 1
            syx[i] = dy / dx
        if abs(dy) >= tol:
            fl[i] |= 2
            sxy[i] = dx / dy
    return syx, sxy, fl


def _signed_area(x, y):
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


def _rect_transform(k, A, B, x0, x1, y0, y1):
    def f(a, lo, hi):
        if abs(a) < 1e-14:
            return hi - lo
        return (np.exp(1j * k * a * hi) - np.exp(1j * k * a * lo)) / (1j * k * a)
    return f(A, x0, x1) * f(B, y0, y1)


def _quad_transform(k, A, B, x, y, n=400):
    """Brute-force 2-D quadrature of exp(ik(Ax+By)) over a convex polygon: fan
    triangulation from vertex 0, Gauss-Legendre on each triangle (Duffy map)."""
    g, w = np.polynomial.legendre.leggauss(n)
    u = 0.5 * (g + 1)
    wu = 0.5 * w
    total = 0j
    for i in range(1, len(x) - 1):
        p0 = np.array([x[0], y[0]])
        p1 = np.array([x[i], y[i]])
        p2 = np.array([x[i + 1], y[i + 1]])
        jac = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1])
        s, t = np.meshgrid(u, u, indexing="ij")
        ws = np.outer(wu, wu) * s  # Duffy: (s, s*t), jacobian s
        px = p0[0] + s * (p1[0] - p0[0]) + s * t * (p2[0] - p1[0])
        py = p0[1] + s * (p1[1] - p0[1]) + s * t * (p2[1] - p1[1])
        total += jac * np.sum(ws * np.exp(1j * k * (A * px + B * py)))
    return total


def test_rectangle_aperture_closed_form(pkg, oracle):
    consts = pkg.abi.make_consts(0.532)
    k = consts.wave_index
    lam = 0.532
    x = np.array([-3.0, 4.0, 4.0, -3.0])
    y = np.array([-1.5, -1.5, 2.5, 2.5])
    syx, sxy, fl = _poly_edges(x, y)
    rng = np.random.default_rng(5)
    for _ in range(200):
        A, B = rng.uniform(-0.8, 0.8, 2)
        got = oracle.aperture(x, y, syx, sxy, fl, A, B, consts, 28.0)
        want = -1j / lam * _rect_transform(k, A, B, -3, 4, -1.5, 2.5)
        assert abs(got - want) <= 1e-11 * 28.0 / lam, (A, B, got, want)


def test_random_polygon_against_quadrature(pkg, oracle):
    consts = pkg.abi.make_consts(0.532)
    k = consts.wave_index
    lam = 0.532
    rng = np.random.default_rng(11)
    for trial in range(6):
        from scipy.spatial import ConvexHull
        pts = rng.uniform(-3, 3, (rng.integers(3, 12), 2)) + rng.normal(size=2)
        hull = ConvexHull(pts).vertices  # counter-clockwise in 2-D
        x, y = pts[hull, 0], pts[hull, 1]
        syx, sxy, fl = _poly_edges(x, y)
        area = _signed_area(x, y)
        assert area > 0
        for _ in range(4):
            A, B = rng.uniform(-0.15, 0.15, 2)
            got = oracle.aperture(x, y, syx, sxy, fl, A, B, consts, area)
            want = -1j / lam * _quad_transform(k, A, B, x, y)
            assert abs(got - want) <= 1e-9 * area / lam, (trial, A, B, got, want)


def test_triangle_closed_form_and_forward_limit(pkg, oracle):
    consts = pkg.abi.make_consts(1.0)
    k = consts.wave_index
    x = np.array([0.0, 2.0, 0.0])
    y = np.array([0.0, 0.0, 1.0])
    syx, sxy, fl = _poly_edges(x, y)
    # forward direction: F = -(i/lambda) * area exactly (small-phase branch)
    f0 = oracle.aperture(x, y, syx, sxy, fl, 0.0, 0.0, consts, 1.0)
    assert f0 == pytest.approx(-1j * 1.0, abs=1e-15)

    def tri(A, B):  # closed form over {x>=0, y>=0, x/2+y<=1}
        a, b = k * A, k * B
        # integral_0^1 dy integral_0^{2(1-y)} e^{i(a x + b y)} dx
        # = integral_0^1 e^{iby} (e^{2ia(1-y)} - 1)/(ia) dy
        def g(c):  # integral_0^1 e^{icy} dy
            return 1.0 if abs(c) < 1e-14 else (np.exp(1j * c) - 1) / (1j * c)
        return (np.exp(2j * a) * g(b - 2 * a) - g(b)) / (1j * a)
    rng = np.random.default_rng(3)
    for _ in range(100):
        A, B = rng.uniform(-1, 1, 2)
        got = oracle.aperture(x, y, syx, sxy, fl, A, B, consts, 1.0)
        assert abs(got - (-1j) * tri(A, B)) < 1e-11, (A, B)
    # both sides of the small-phase threshold (max|phase| = 1e-3) against the
    # closed form: the moment expansion drops terms of order phase^3/6 and the
    # edge formula loses ~eps/phase^2 to cancellation, both ~3e-10 here.
    d = 1e-3 / (k * 2.0)
    for A in (d * 0.999, d * 1.001, -d * 0.5):
        got = oracle.aperture(x, y, syx, sxy, fl, A, 0.3 * A, consts, 1.0)
        assert abs(got - (-1j) * tri(A, 0.3 * A)) < 5e-10, A


def test_incoherent_sum_is_the_sum_of_single_beam_runs(pkg, oracle, fixtures):
    """The oracle's incoherent sum (sum_o w_o sum_b Mueller(d_b)) against the
    pinned coherent path run on one beam at a time: a batch of one beam has
    no interference, so the two must agree to rounding."""
    abi = pkg.abi
    fx = fixtures["sobol_average"]
    consts = abi.consts_from_fixture(fx)
    src = abi.BeamBatch.from_fixture(fx).subset([0, 3])
    keep = [0, 1, 2, 5, 8, 13, 21] + [src.orient_offsets[1] + i for i in (0, 2, 7, 11)]
    vo = src.vert_offsets
    counts = [sum(1 for b in keep if b < src.orient_offsets[1]),
              sum(1 for b in keep if b >= src.orient_offsets[1])]
    verts = np.concatenate([np.arange(vo[b], vo[b + 1]) for b in keep])
    nv = src.beam_nv[keep]
    batch = abi.BeamBatch(np.r_[0, np.cumsum(counts)], src.weights, src.beam_scalars[keep], nv,
                          src.beam_external[keep], np.r_[0, np.cumsum(nv)], src.vx[verts],
                          src.vy[verts], src.slope_yx[verts], src.slope_xy[verts],
                          src.edge_flags[verts])
    grid = abi.Grid.build(np.linspace(0, np.pi, 37), 8)
    inc = oracle.incoherent(grid, batch, consts)
    want = np.zeros_like(inc)
    for i in range(batch.n_beams):
        o = 0 if i < counts[0] else 1
        one = batch.subset([o])
        j = i - batch.orient_offsets[o]
        ov = one.vert_offsets
        b1 = abi.BeamBatch([0, 1], [one.weights[0]], one.beam_scalars[j:j + 1], one.beam_nv[j:j + 1],
                           one.beam_external[j:j + 1], [0, one.beam_nv[j]],
                           one.vx[ov[j]:ov[j + 1]], one.vy[ov[j]:ov[j + 1]],
                           one.slope_yx[ov[j]:ov[j + 1]], one.slope_xy[ov[j]:ov[j + 1]],
                           one.edge_flags[ov[j]:ov[j + 1]])
        want += oracle.diffract(grid, b1, consts)[0]
    np.testing.assert_allclose(inc, want, rtol=0, atol=1e-13 * np.abs(want[..., 0]).max())
    coh = oracle.diffract(grid, batch, consts)[0]
    assert np.abs(coh[..., 0] - inc[..., 0]).max() > 1e-3 * inc[..., 0].max()  # beams interfere
    assert np.all(inc[..., 0] >= 0)


def dirichlet_upsampling_matrix(n_low, n_full):
    """The upsampling rule as a real matrix in closed form (Dirichlet
    kernel): K(u) = (1/n_low) sum_k cos(2 pi f_k u) over the frequencies the
    rule keeps."""
    q = np.arange(n_full)[:, None]
    p = np.arange(n_low)[None, :]
    a = 2 * np.pi * (q / n_full - p / n_low)
    m = n_low // 2
    zero = q * n_low == p * n_full
    h = np.where(zero, 1.0, np.sin(0.5 * a))
    if n_low % 2:
        k = np.sin((m + 0.5) * a) / h / n_low
    else:
        k = (np.sin((m - 0.5) * a) / h + np.cos(m * a)) / n_low
    return np.where(zero, 1.0, k)


def test_theta_upsampling_reproduces_cosine_series(pkg):
    """The theta upsampling's even extension: a cosine series of degree
    < n_theta_low / 2 sampled on the coarse uniform rows is reproduced on
    the fine rows."""
    rng = np.random.default_rng(8)
    for nt_low, nt_full in ((16, 32), (45, 90), (60, 180)):
        c = rng.normal(size=nt_low // 2)
        def f(t):
            return sum(c[j] * np.cos(j * t) for j in range(len(c)))
        low = np.tile(f(np.linspace(0, np.pi, nt_low + 1))[None, :, None], (2, 1, 16))
        up = fourier.upsample_theta(low, nt_full)
        assert up.shape == (2, nt_full + 1, 16)
        np.testing.assert_allclose(up[1, :, 5], f(np.linspace(0, np.pi, nt_full + 1)),
                                   rtol=0, atol=1e-12 * np.abs(c).sum())


def test_phi_upsampling_forms(pkg):
    """The checker of the device upsampler (tests/fourier.py, the reference's
    FFT padding) pinned two ways: it reproduces a trigonometric polynomial
    of degree < n_low/2 sampled on the coarse grid exactly, and it equals
    the closed-form (Dirichlet) matrix of the same rule on arbitrary data,
    odd and even n_low."""
    rng = np.random.default_rng(4)
    for n_low, n_full in ((16, 32), (48, 96), (45, 90), (80, 240), (21, 50)):
        x = rng.normal(size=(n_low, 3, 16))
        fft_form = fourier.upsample_phi(x, n_full)
        mat_form = np.einsum("qp,p...->q...", dirichlet_upsampling_matrix(n_low, n_full), x)
        np.testing.assert_allclose(mat_form, fft_form, rtol=0, atol=1e-12 * np.abs(x).max())
        deg = (n_low - 1) // 2
        c = rng.normal(size=(deg + 1, 2))
        def trig(phi):
            return sum(c[j, 0] * np.cos(j * phi) + c[j, 1] * np.sin(j * phi) for j in range(deg + 1))
        lo = trig(2 * np.pi * np.arange(n_low) / n_low)[:, None, None] * np.ones((1, 1, 16))
        want = trig(2 * np.pi * np.arange(n_full) / n_full)
        np.testing.assert_allclose(fourier.upsample_phi(lo, n_full)[:, 0, 0], want,
                                   rtol=0, atol=1e-12 * np.abs(c).sum())


def least_width(x, y):
    """Least width of a vertex set as the beam packing kernel computes it
    (pack_beams_kernel, po_kernels.hip): over all vertex pairs, the spread of
    the vertices across the line through the pair."""
    w = np.inf
    for i in range(len(x)):
        for j in range(i + 1, len(x)):
            ex, ey = x[j] - x[i], y[j] - y[i]
            n = np.hypot(ex, ey)
            if n > 0:
                c = ex * (y - y[i]) - ey * (x - x[i])
                w = min(w, (max(c.max(), 0.0) - min(c.min(), 0.0)) / n)
    return w


def test_small_phase_width_bound():
    """The device kernels test the vertex phases one by one only where
    max(|A|, |B|) <= 2.04e-3 / (k w), w the least width of the vertex set.
    Pinned two ways: w against a brute-force sweep of directions, and the
    implication (every vertex phase within 1e-3 rad => below the bound) on
    random directions about the bound, for convex, non-convex and sliver
    polygons."""
    rng = np.random.default_rng(11)
    k = 2 * np.pi / 0.532
    ang = np.linspace(0, np.pi, 200001)
    u = np.stack([np.cos(ang), np.sin(ang)])
    polys = [(np.array([0., 2, 0]), np.array([0., 0, 1])),
             (np.array([0., 1, 1, 0]), np.array([0., 0, 1e-4, 1e-4])),   # sliver
             (np.array([0., 4, 4, 2, 0]), np.array([0., 0, 3, 0.5, 3]))]  # non-convex
    for _ in range(20):
        n = rng.integers(3, 12)
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(0.2, 1.0, n) * rng.uniform(0.1, 30)
        polys.append((r * np.cos(t) + rng.normal(), 0.3 * r * np.sin(t)))
    for x, y in polys:
        proj = u[0][:, None] * x + u[1][:, None] * y
        brute = (proj.max(1) - proj.min(1)).min()
        w = least_width(x, y)
        # no direction is narrower than w, and the sweep (step dt) comes
        # within diameter * dt of it
        diam = np.hypot(x[:, None] - x, y[:, None] - y).max()
        assert w <= brute * (1 + 1e-12) and brute - w <= diam * (ang[1] - ang[0]), (w, brute)
        bound = 2.04e-3 / (k * w)
        s = rng.uniform(0.0, 3.0, 20000) * bound * np.sqrt(2)
        a = rng.uniform(0, 2 * np.pi, 20000)
        A, B = s * np.cos(a), s * np.sin(a)
        ph = np.abs(k * (A[:, None] * x + B[:, None] * y)).max(1)
        small = ph <= 1e-3
        assert small.any() and (~small).any()
        assert np.all(np.maximum(np.abs(A), np.abs(B))[small] <= bound)


def test_mueller_identities(oracle):
    rng = np.random.default_rng(1)
    G = np.diag([1.0, -1.0, -1.0, -1.0])
    for _ in range(50):
        j8 = rng.normal(size=8)
        m = oracle.mueller_from_jones(j8).reshape(4, 4)
        jc = j8[0::2] + 1j * j8[1::2]
        assert m[0, 0] == pytest.approx(0.5 * np.sum(np.abs(jc) ** 2), rel=1e-14)
        det = abs(jc[0] * jc[3] - jc[1] * jc[2]) ** 2
        np.testing.assert_allclose(m @ G @ m.T, det * G, atol=1e-12 * m[0, 0] ** 2)
        assert m[0, 0] ** 2 >= np.sum(m[1:, 0] ** 2) * (1 - 1e-12)  # DOP <= 1
    np.testing.assert_allclose(oracle.mueller_from_jones(np.array([1., 0, 0, 0, 0, 0, 1, 0])).reshape(4, 4),
                               np.eye(4), atol=1e-15)
    pol = oracle.mueller_from_jones(np.array([1., 0, 0, 0, 0, 0, 0, 0])).reshape(4, 4)
    np.testing.assert_allclose(pol, 0.5 * np.array([[1, 1, 0, 0], [1, 1, 0, 0], [0, 0, 0, 0],
                                                    [0, 0, 0, 0]]), atol=1e-15)
    qwp = oracle.mueller_from_jones(np.array([1., 0, 0, 0, 0, 0, 0, 1])).reshape(4, 4)
    assert abs(qwp[2, 3]) == pytest.approx(1.0) and qwp[3, 2] == pytest.approx(-qwp[2, 3])
    assert abs(qwp[2, 2]) < 1e-15 and abs(qwp[3, 3]) < 1e-15


def test_solid_angle_weights_sum_to_4pi(pkg, oracle):
    for theta in (np.linspace(0, np.pi, 721),
                  np.radians(np.r_[np.arange(0, 10.001, 0.05), np.arange(10.5, 180.001, 0.5)])):
        grid = pkg.abi.Grid.build(theta, 4)
        avg = oracle.azimuth_average(grid, np.zeros((4, len(theta), 16)))
        assert avg[:, 1].sum() == pytest.approx(4 * np.pi, rel=1e-13)


def test_convergence_error_measures():
    """A uniform M11 scale of (1 + d) is an M11 error d and leaves every
    ratio exact; adding e * M11 to one element is an error e in that element
    alone, in both the weighted and the plot form."""
    import metrics
    rng = np.random.default_rng(3)
    theta = np.radians(np.r_[np.arange(0, 10.001, 0.05), np.arange(10.5, 180.001, 0.5)])
    ref = np.zeros((len(theta), 18))
    ref[:, 0], ref[:, 1] = np.degrees(theta), metrics.two_pi_dcos(theta)
    ref[:, 2] = rng.uniform(1, 1e4, len(theta))
    ref[:, 3:] = rng.uniform(-0.5, 0.5, (len(theta), 15)) * ref[:, 2:3]
    for f in (metrics.convergence_errors, metrics.plot_element_errors):
        np.testing.assert_array_equal(f(ref, ref), np.zeros(16))
        a = ref.copy()
        a[:, 2:] *= 1.003
        e = f(a, ref)
        assert e[0] == pytest.approx(0.003, rel=1e-9) and np.all(e[1:] < 1e-15)
        a = ref.copy()
        a[:, 2 + 6] += 0.02 * a[:, 2]
        e = f(a, ref)
        assert e[6] == pytest.approx(0.02, rel=1e-9)
        assert np.all(np.delete(e, 6) < 1e-15)


def test_azimuth_average_undoes_frame_rotation(pkg, oracle):
    """If M(phi_p) = M0 . L(-phi_p)^-1, the lab-frame average returns M0."""
    theta = np.linspace(0.1, 3.0, 5)
    n_az = 12
    grid = pkg.abi.Grid.build(theta, n_az)
    rng = np.random.default_rng(2)
    m0 = rng.normal(size=(len(theta), 4, 4))
    m = np.zeros((n_az, len(theta), 16))
    for p in range(n_az):
        a = 2 * (-p * grid.az_step)
        L = np.eye(4)
        L[1, 1] = L[2, 2] = np.cos(a)
        L[1, 2], L[2, 1] = np.sin(a), -np.sin(a)
        for t in range(len(theta)):
            m[p, t] = (m0[t] @ np.linalg.inv(L)).ravel()
    avg = oracle.azimuth_average(grid, m)
    np.testing.assert_allclose(avg[:, 2:].reshape(-1, 4, 4), m0, atol=1e-13)


# ---- orientation sets --------------------------------------------------------

def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _ry(b):
    c, s = np.cos(b), np.sin(b)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def test_sobol_matches_reference_points(oracle):
    """The oracle's sequential restatement of Sobol2D reproduces the points the
    reference's own Sobol2D (compiled from src/math/Sobol.cpp) wrote, bit for
    bit, for five seeds x 4096 points."""
    fx = load_fixture("sobol_points")
    n = fx["x"].shape[1]
    for k, seed in enumerate(fx["seeds"]):
        x, y = oracle.sobol(n, int(seed))
        np.testing.assert_array_equal(x, fx["x"][k])
        np.testing.assert_array_equal(y, fx["y"][k])
    # Owen scrambling keeps the (0, m, 2)-net property: every elementary
    # interval of area 1/256 holds exactly one point of an aligned block of
    # 256 sequence points.  Output k is sequence point k+1 (the first point
    # is emitted after one Gray-code step), so outputs 255..510 are points
    # 256..511.
    x, y = fx["x"][0, 255:511], fx["y"][0, 255:511]
    for a in range(9):
        ix = np.floor(x * 2 ** a).astype(int)
        iy = np.floor(y * 2 ** (8 - a)).astype(int)
        assert len(set(zip(ix, iy))) == 256


def test_orientation_sets_geometry(oracle):
    n = 512
    for kind in (0, 1):
        rows = oracle.orientation_set(kind, n, seed=42, beta_sym=np.pi / 2, gamma_sym=np.pi / 3)
        assert np.all(rows[:, 2] == 1.0 / n)
        assert np.all((rows[:, 0] >= 0) & (rows[:, 0] <= np.pi / 2 + 1e-15))
        assert np.all((rows[:, 1] >= 0) & (rows[:, 1] < np.pi / 3))
        for r in rows[::37]:
            R = r[7:16].reshape(3, 3)
            np.testing.assert_allclose(R, _ry(r[0]) @ _rz(r[1]), atol=1e-15)
        # uniform solid-angle measure: cos(beta) is uniform on [0, 1]
        assert abs(np.mean(np.cos(rows[:, 0])) - 0.5) < 2e-3
    ham = oracle.orientation_set(1, 8, beta_sym=np.pi, gamma_sym=2 * np.pi)
    np.testing.assert_allclose(ham[:, 1] / (2 * np.pi),
                               [0, .5, .25, .75, .125, .625, .375, .875], atol=1e-16)
    np.testing.assert_allclose(np.cos(ham[:, 0]), 1 - 2 * (np.arange(8) + .5) / 8, atol=1e-15)
    sh = oracle.orientation_set(2, 9)
    q = sh[:, 3:7]
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-15)
    # u3 = radical inverse base 3: 0, 1/3, 2/3, 1/9, 4/9, 7/9, 2/9, 5/9, 8/9
    u3 = np.arctan2(q[:, 2], q[:, 3]) % (2 * np.pi) / (2 * np.pi)
    np.testing.assert_allclose(u3[1:], np.array([1, 2, 1 / 3, 4 / 3, 7 / 3, 2 / 3, 5 / 3, 8 / 3])
                               / 3, atol=1e-14)
    for r in sh:
        R = r[7:16].reshape(3, 3)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14


# ---- aperture FFT (oracle side) -----------------------------------------------

def _rect_G(k, A, B, x0, x1, y0, y1):
    return _rect_transform(k, A, B, x0, x1, y0, y1)


def test_aperture_fft_oracle_against_closed_forms(pkg, oracle):
    """The FFT restatement reproduces the closed-form rectangle transform on
    frequency bins (only the raster's pixel error remains) and between bins
    (Lanczos interpolation at 4x oversampling), and a random convex polygon's
    exact transform from the edge-sum oracle."""
    k = 2 * np.pi / 0.532
    x = np.array([-3.0, 4.0, 4.0, -3.0]) + 0.7
    y = np.array([-1.5, -1.5, 2.5, 2.5]) - 0.3
    area = 28.0
    n, pad = 1024, 4.0
    dx, dy = pad * 7 / n, pad * 4 / n
    m = np.array([(0, 0), (1, 0), (3, 2), (-5, 1), (7, -3), (-20, 11)])
    ab_bin = np.stack([-2 * np.pi * m[:, 0] / (n * dx) / k, -2 * np.pi * m[:, 1] / (n * dy) / k], 1)
    got, _ = oracle.aperture_fft(x, y, n, pad, ab_bin, k)
    want = np.array([_rect_G(k, A, B, x.min(), x.max(), y.min(), y.max()) for A, B in ab_bin])
    assert np.max(np.abs(got - want)) / area < 5e-5
    rng = np.random.default_rng(2)
    ab = rng.uniform(-0.01, 0.01, (24, 2))
    got, _ = oracle.aperture_fft(x, y, n, pad, ab, k)
    want = np.array([_rect_G(k, A, B, x.min(), x.max(), y.min(), y.max()) for A, B in ab])
    assert np.max(np.abs(got - want)) / area < 5e-3
    # random convex polygon vs the edge-sum integral (oracle.aperture = -i/lambda I)
    from scipy.spatial import ConvexHull
    pts = rng.uniform(-4, 4, (9, 2))
    hull = ConvexHull(pts).vertices
    px, py = pts[hull, 0], pts[hull, 1]
    syx, sxy, fl = _poly_edges(px, py)
    parea = _signed_area(px, py)
    consts = pkg.abi.make_consts(0.532)
    got, _ = oracle.aperture_fft(px, py, 1024, 8.0, ab, k)
    want = np.array([oracle.aperture(px, py, syx, sxy, fl, A, B, consts, parea)
                     for A, B in ab]) / (-1j / 0.532)
    assert np.max(np.abs(got - want)) / parea < 2e-3


# ---- control-variate MLP (oracle side) -------------------------------------------

def test_mlp_oracle_gradient_against_finite_differences():
    """The FP64 MLP restatement's analytic gradient equals central finite
    differences of its loss (every parameter of a small network)."""
    from oracle import mlp_oracle as mo
    shapes = mo.layer_shapes(5, 6, 2, 4)
    npar = sum(i * o + o for i, o in shapes)
    rng = np.random.default_rng(4)
    p = rng.normal(scale=0.5, size=npar)
    x = rng.normal(size=(7, 5))
    t = rng.normal(size=(7, 4))
    loss, g = mo.loss_and_grad(p, shapes, x, t)
    h = 1e-6
    fd = np.empty(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (mo.loss_and_grad(p + e, shapes, x, t)[0]
                 - mo.loss_and_grad(p - e, shapes, x, t)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-8 * max(1.0, np.max(np.abs(fd)))
    # GELU: Phi(z) + z phi(z) against a finite difference; GELU(0) = 0
    z = np.linspace(-4, 4, 41)
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + 1e-6) - mo.gelu(z - 1e-6)) / 2e-6,
                               atol=1e-8)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    # BF16 rounding: exact values survive, ties go to even
    np.testing.assert_array_equal(mo.bf16(np.array([1.0, -3.0, 0.5])), [1.0, -3.0, 0.5])
    assert mo.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0          # tie -> even (down)
    assert mo.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6  # tie -> even (up)
