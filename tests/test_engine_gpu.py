"""Parity of the HIP engine with the FP64 oracle, and through it with the
reference, on the reference's own prepared beams, through the C ABI.

Bars (tests/metrics.py defines the measures):
  FP64 engine vs oracle:    grid_l2 1e-10 (the FP64 tolerance of the path),
                            row 2e-8, cell 2e-7
  FP64 engine vs reference: grid_l2 / gate_l2 1e-6 (the reference CPU path
                            evaluates phases with a polynomial whose cosine is
                            off by up to 6.3e-9 and jumps by 1.3e-8 at odd
                            multiples of pi/2 -- it is the reference that
                            carries this error, the oracle with libm sincos
                            does not)
  mixed engine vs oracle:   P11 grid_l2 and P11 gate_l2 1e-4 (the FP32
                            tolerance of the path, in the reference's own gate norm);
                            per-row and per-cell P11 errors are reported.
The per-cell bound is looser than the aggregate ones because a coherent sum of
hundreds of beams can cancel to a local M11 1e4 times below its row, and near
an edge with |C| just above eps1 = 2.2e-7 the edge quotient amplifies the
rounding of its phase arguments (|phase| ~ 1e2 rad) by up to 1/eps1 -- for the
oracle as much as for the engine.
"""
import numpy as np
import pytest

import fourier
import metrics as M
from conftest import load_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.engine.Context(0)
    yield c
    c.close()


def _setup(pkg, fx):
    abi = pkg.abi
    return abi.Grid.from_fixture(fx), abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx)


def _run(ctx, grid, batch, consts, flags=0, noshadow=False, target=None):
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_target_blocks(target or 65536)
    return ctx.mueller(flags, noshadow)


def check_fp64(a, b, theta, label):
    print(label + ": " + M.summary(a, b, theta))
    assert M.grid_l2(a, b, theta) < 1e-10, label
    assert M.row(a, b) < 2e-8, label
    assert M.cell(a, b) < 2e-7, label


def test_device_is_gfx950(pkg):
    info = pkg.engine.device_info()
    print("device:", info)
    assert info["arch"].startswith("gfx950")


@pytest.mark.parametrize("name", ["config1_fixed", "fixed_absorbing"])
def test_fixed_orientation_fp64(pkg, oracle, ctx, name):
    fx = load_fixture(name)
    grid, batch, consts = _setup(pkg, fx)
    m, mn = _run(ctx, grid, batch, consts, noshadow=True)
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    check_fp64(m, o, grid.theta, name + " FP64 vs oracle")
    check_fp64(mn, on, grid.theta, name + " FP64 no-shadow vs oracle")
    # against the reference's own outputs on these beams
    rows = fx.get("kept_theta_rows")
    mr = m[:, rows] if rows is not None else m
    th = grid.theta[rows] if rows is not None else grid.theta
    for key in ("mueller_full", "mueller_fixed_dat"):
        e = M.grid_l2(mr, fx[key], th)
        print("%s FP64 vs reference %s: grid_l2 %.2e cell %.2e" % (name, key, e,
                                                                    M.cell(mr, fx[key])))
        assert e < 1e-6


@pytest.mark.parametrize("mode", [1, 2])
def test_orientation_average_both_work_modes(pkg, oracle, ctx, mode):
    """Whole-orientation items (mode 1) and beam segments (mode 2) must both
    equal the oracle, and the azimuth-averaged rows must pass the reference's
    gate against the reference CLI's own .dat."""
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_work_mode(mode)
    m, mn = _run(ctx, grid, batch, consts, noshadow=True)
    ctx.set_work_mode(0)
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    check_fp64(m, o, grid.theta, "average mode=%d vs oracle" % mode)
    check_fp64(mn, on, grid.theta, "average no-shadow mode=%d vs oracle" % mode)
    for mm, key in ((m, "ref_total_dat"), (mn, "ref_total_dat_noshadow")):
        avg = oracle.azimuth_average(grid, mm)
        g = M.gate_l2(avg, fx[key])
        print("average mode=%d %s gate_l2 vs reference .dat: %.2e" % (mode, key, g))
        assert g < 1e-6


def test_azimuth_average_on_device(pkg, oracle, ctx):
    """The device azimuth average (poles included) equals the oracle's
    restatement of HandlerPOTotal::WriteMatricesToFile on the same grid, and
    its rows pass the reference gate against the reference CLI's .dat."""
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    m, mn = _run(ctx, grid, batch, consts, noshadow=True)
    for mm, ns, key in ((m, False, "ref_total_dat"), (mn, True, "ref_total_dat_noshadow")):
        rows = ctx.azimuth_average(noshadow=ns)
        want = oracle.azimuth_average(grid, mm)
        np.testing.assert_allclose(rows[:, 0], fx[key][:, 0], rtol=1e-15, atol=0)
        np.testing.assert_allclose(rows[:, 1], want[:, 1], rtol=1e-13, atol=0)
        scale = np.abs(want[:, 2:3])
        err = float(np.max(np.abs(rows[:, 2:] - want[:, 2:]) / scale))
        g = M.gate_l2(rows, fx[key])
        print("%s: device vs oracle azimuth average max|d|/M11 %.2e; gate_l2 vs reference "
              ".dat %.2e" % (key, err, g))
        assert err < 1e-13
        assert g < 1e-6
        # the pole rows keep the pole form's zero pattern
        for t in (0, grid.n_zen):
            z = rows[t, 2:].reshape(4, 4)
            assert z[0, 1] == z[0, 2] == z[1, 0] == z[2, 0] == z[1, 2] == z[2, 1] == 0.0


@pytest.mark.parametrize("mode", [1, 2])
def test_per_orientation_grids(pkg, oracle, ctx, mode):
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_work_mode(mode)
    m, _ = ctx.mueller(pkg.abi.PER_ORIENT)
    per = ctx.per_orient()
    ctx.set_work_mode(0)
    _, _, oper = oracle.diffract(grid, batch, consts, per_orient=True)
    for o in range(batch.n_orient):
        check_fp64(per[o], oper[o], grid.theta, "orientation %d" % o)
    np.testing.assert_allclose(per.sum(axis=0), m, rtol=1e-12, atol=1e-15 * np.abs(m).max())


def test_mixed_precision_p11(pkg, oracle, ctx):
    for name in ("config1_fixed", "fixed_absorbing", "sobol_average"):
        fx = load_fixture(name)
        grid, batch, consts = _setup(pkg, fx)
        m, _ = _run(ctx, grid, batch, consts, flags=pkg.abi.PREC_MIXED)
        o, _, _ = oracle.diffract(grid, batch, consts)
        a_m, a_o = oracle.azimuth_average(grid, m), oracle.azimuth_average(grid, o)
        l2_11 = M.grid_l2(m[..., :1], o[..., :1], grid.theta)
        avg11 = float(np.max(np.abs(a_m[:, 2] - a_o[:, 2]) / np.abs(a_o[:, 2])))
        print("%s mixed vs oracle: P11 grid_l2 %.2e, azimuth-avg P11 max rel %.2e, "
              "per-cell P11 max rel %.2e, all elements: %s"
              % (name, l2_11, avg11, M.p11_rel(m, o), M.summary(m, o, grid.theta)))
        g11 = M.gate_l2(a_m[:, :3], a_o[:, :3])
        print("   azimuth-averaged P11 gate_l2 %.2e" % g11)
        assert l2_11 < 1e-4
        assert g11 < 1e-4
        # the rows users read: every azimuth-averaged P11 within 1e-4
        assert avg11 < 1e-4


@pytest.mark.parametrize("mode", [1, 2])
def test_mfma_contraction_equals_vector_sum(pkg, oracle, ctx, mode):
    """The mixed kernel's MFMA contraction (default) and its vector-ALU form
    (MBS_NO_MFMA) compute the same FP32 sums in a different order: they agree
    to FP32 rounding, full and no-shadow, per orientation, in both work modes."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_work_mode(mode)
    out = {}
    for name, extra in (("mfma", 0), ("valu", abi.NO_MFMA)):
        m, mn = ctx.mueller(abi.PREC_MIXED | abi.PER_ORIENT | extra, noshadow=True)
        out[name] = (m, mn, ctx.per_orient())
    ctx.set_work_mode(0)
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    for k, (a, b) in enumerate(zip(out["mfma"], out["valu"])):
        e = M.grid_l2(a, b, grid.theta) if k < 2 else max(
            M.grid_l2(a[i], b[i], grid.theta) for i in range(len(a)))
        print("mode %d part %d: MFMA vs VALU grid_l2 %.2e" % (mode, k, e))
        assert e < 2e-5
    for name in ("mfma", "valu"):
        e = M.grid_l2(out[name][0][..., :1], o[..., :1], grid.theta)
        en = M.grid_l2(out[name][1][..., :1], on[..., :1], grid.theta)
        print("%s P11 vs oracle: full %.2e no-shadow %.2e" % (name, e, en))
        assert e < 1e-4 and en < 1e-4


def test_accumulates_and_is_deterministic(pkg, ctx):
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.diffract(0)
    a = np.zeros(grid.cells * 16)
    ctx.accumulate(a)
    b = a.copy()
    ctx.diffract(0)
    ctx.accumulate(b)          # += semantics of HandleOrientationsToLocalGpu
    np.testing.assert_array_equal(b, 2 * a)
    ctx.diffract(0)
    c = np.zeros_like(a)
    ctx.accumulate(c)
    np.testing.assert_array_equal(c, a)  # bit-reproducible


def test_bad_beams_rejected(pkg, ctx):
    fx = load_fixture("fixed_absorbing")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    for bad_nv in (2, 32):
        nv = batch.beam_nv.copy()
        nv[0] = bad_nv
        vo = np.concatenate([[0], np.cumsum(nv)])
        n = int(vo[-1])
        bb = pkg.abi.BeamBatch(batch.orient_offsets, batch.weights, batch.beam_scalars, nv,
                               batch.beam_external, vo, np.zeros(n), np.zeros(n), np.zeros(n),
                               np.zeros(n), np.zeros(n, np.uint8))
        with pytest.raises(pkg.engine.EngineError) as ei:
            ctx.upload(bb, consts)
        assert ei.value.status == pkg.abi.MBS_ERR_BAD_BEAM
    with pytest.raises(pkg.engine.EngineError) as ei:
        ctx.diffract(0x100)
    assert ei.value.status in (pkg.abi.MBS_ERR_INVALID, pkg.abi.MBS_ERR_NO_BEAMS)


def test_empty_batches(pkg, ctx):
    """No orientations, and orientations without beams: zero result, no error."""
    fx = load_fixture("fixed_absorbing")
    grid, _, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    z = np.zeros(0)
    for oo, w in (([0], []), ([0, 0, 0], [0.5, 0.5])):
        bb = pkg.abi.BeamBatch(oo, w, np.zeros((0, 36)), np.zeros(0, np.int32),
                               np.zeros(0, np.int32), [0], z, z, z, z, np.zeros(0, np.uint8))
        ctx.upload(bb, consts)
        m, _ = ctx.mueller(0)
        assert not m.any()


def test_handler_mirror_semantics(pkg, oracle):
    """HandlerPOGpu: accumulate into Arr2D-shaped grids, no-shadow only when
    requested, start/count selection, False for a batch it cannot take."""
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    h = pkg.handler.HandlerPOGpu(grid, consts)
    Mg = np.zeros((grid.n_az + 1, grid.n_zen + 1, 16))
    Mns = np.zeros_like(Mg)
    assert h.HandleOrientationsToLocalGpu(batch, Mg, Mns, 2, 3)
    assert not Mns.any()                      # full-only by default
    o, _, _ = oracle.diffract(grid, batch.subset([2, 3, 4]), consts)
    check_fp64(Mg[:-1], o, grid.theta, "handler mirror")
    assert not Mg[-1].any()
    h.SetFullOnly(False)
    assert h.HandleOrientationsToLocalGpu(batch, Mg, Mns, 5, 100)   # count clipped
    assert Mns.any()
    assert h.SelectGpuOrientationBatchSize(batch, 0, 8) >= 1
    h.isCoh = False
    assert not h.HandleOrientationsToLocalGpu(batch, Mg, Mns)
    h.close()


def test_theta_probe_and_incoherent_sum(pkg, oracle, ctx):
    """HandlerPO::DiffractThetasGpu (the reference's theta probe) through the
    mirror and the engine's MBS_INCOHERENT sum, against the oracle's
    incoherent sum on the probe grid as the reference builds it
    (GpuDiffraction.cu:4845-4876: theta at a pole moved 1e-6 rad inside,
    vf = (-sin phi, cos phi, 0)) and the probe's scaling: 4 x the sum over
    azimuths / (nAz + 1) (theta_m11_kernel, :1909-1957; the factor 4 is the
    rotation's missing 1/2 squared)."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    theta = np.r_[0.0, np.radians([0.5, 1, 2, 5, 10, 30, 60, 90, 120, 150, 175, 179.5]), np.pi]
    th = theta.copy()
    th[th <= 1e-14] = 1e-6
    th[th >= np.pi - 1e-14] = np.pi - 1e-6
    phi = np.arange(grid.n_az) * grid.az_step
    vf = np.zeros((grid.n_az, len(th), 3))
    vf[:, :, 0] = -np.sin(phi)[:, None]
    vf[:, :, 1] = np.cos(phi)[:, None]
    probe = abi.Grid(th, grid.az_step, vf)
    inc = oracle.incoherent(probe, batch, consts)
    # the engine's incoherent Mueller grid
    m, _ = _run(ctx, probe, batch, consts, flags=abi.INCOHERENT)
    check_fp64(m, inc, th, "incoherent sum FP64 vs oracle")
    for bad in (abi.PREC_MIXED, abi.NOSHADOW, abi.PER_ORIENT):
        with pytest.raises(pkg.engine.EngineError):
            ctx.diffract(abi.INCOHERENT | bad)
    # the probe through the handler mirror
    h = pkg.handler.HandlerPOGpu(grid, consts)
    ok, m11 = h.DiffractThetasGpu(batch, theta)
    assert ok
    want = 4.0 * inc[..., 0].sum(axis=0) / (grid.n_az + 1)
    print("theta probe vs oracle: max rel %.2e" % np.max(np.abs(m11 - want) / want))
    np.testing.assert_allclose(m11, want, rtol=1e-10, atol=0)
    # the handler's own grid is back in place for the coherent path
    Mg = np.zeros((grid.n_az, grid.n_zen + 1, 16))
    assert h.HandleOrientationsToLocalGpu(batch, Mg, None)
    o, _, _ = oracle.diffract(grid, batch, consts)
    check_fp64(Mg, o, grid.theta, "coherent after the probe")
    # nothing to do: untouched; a beam it cannot take: False
    assert h.DiffractThetasGpu(batch, theta[:0]) == (True, None)
    bad = abi.BeamBatch(batch.orient_offsets, batch.weights, batch.beam_scalars,
                        np.where(np.arange(batch.n_beams) == 3, 40, batch.beam_nv),
                        batch.beam_external, batch.vert_offsets, batch.vx, batch.vy,
                        batch.slope_yx, batch.slope_xy, batch.edge_flags)
    assert h.DiffractThetasGpu(bad, theta) == (False, None)
    h.close()


def test_device_rescale_of_a_beam_set(pkg, oracle, ctx):
    """mbs_rescale_beams (the multi-size path without a re-pack per scale)
    against the host rule of the drop-in's pack (the reference's
    GpuDiffraction.cu:273-358, :4324): vertices and centres times s, areas
    times s^2, J_phased = sign a(s) J exp(i k path s), a(s) the mean of
    exp(c_abs p s) over the absorption paths (1 for p <= DBL_EPSILON).
    The host-scaled batch goes through the oracle and through a plain upload."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, full, consts = _setup(pkg, fx)
    batch = full.subset([0, 1, 2])
    nb = batch.n_beams
    rng = np.random.default_rng(17)
    raw = rng.normal(size=(nb, 8))
    path = rng.uniform(1.0, 50.0, nb)
    sign = np.where(rng.random(nb) < 0.4, -1.0, 1.0)
    n_abs = rng.integers(0, 4, nb)
    aoff = np.r_[0, np.cumsum(n_abs)]
    apath = rng.uniform(0.0, 20.0, aoff[-1])
    apath[::5] = 0.0
    c_abs, kj = -0.02, consts.wave_index

    def host_scaled(sc):
        bs = batch.beam_scalars.copy()
        bs[:, abi.B_CEN:abi.B_CEN + 3] *= sc
        bs[:, abi.B_AREA] *= sc * sc
        amp = sign.copy()
        for b in range(nb):
            p = apath[aoff[b]:aoff[b + 1]]
            if len(p):
                amp[b] *= np.mean(np.where(p > np.finfo(float).eps, np.exp(c_abs * p * sc), 1.0))
        rot = np.exp(1j * kj * (path * sc))
        j = (raw[:, 0::2] + 1j * raw[:, 1::2]) * (amp * rot)[:, None]
        bs[:, abi.B_JP:abi.B_JP + 8:2] = j.real
        bs[:, abi.B_JP + 1:abi.B_JP + 8:2] = j.imag
        return abi.BeamBatch(batch.orient_offsets, batch.weights, bs, batch.beam_nv,
                             batch.beam_external, batch.vert_offsets, batch.vx * sc,
                             batch.vy * sc, batch.slope_yx, batch.slope_xy, batch.edge_flags)

    ctx.set_grid(grid)
    ctx.set_target_blocks(65536)
    for sc in (0.5, 1.7, 3.0):
        hb = host_scaled(sc)
        ctx.upload(hb, consts)
        m_host, _ = ctx.mueller(0)
        ctx.upload(batch, consts)
        ctx.upload_scaling(raw, path, sign, aoff, apath, c_abs, kj)
        ctx.rescale(sc)
        m_dev, _ = ctx.mueller(0)
        o, _, _ = oracle.diffract(grid, hb, consts)
        check_fp64(m_dev, o, grid.theta, "device rescale s=%g vs oracle" % sc)
        e = M.grid_l2(m_dev, m_host, grid.theta)
        print("device rescale s=%g vs host-scaled upload: grid_l2 %.2e" % (sc, e))
        assert e < 1e-12
    ctx.rescale(1.0)  # back to the batch as uploaded
    m1, _ = ctx.mueller(0)
    o1, _, _ = oracle.diffract(grid, batch, consts)
    check_fp64(m1, o1, grid.theta, "rescale 1 restores the batch")
    with pytest.raises(pkg.engine.EngineError):
        ctx.rescale(-1.0)
    ctx.upload(batch, consts)  # a new upload drops the scaling data
    with pytest.raises(pkg.engine.EngineError):
        ctx.rescale(2.0)


def test_phi_fft_interpolation_path(pkg, oracle):
    """HandleOrientationsToLocalGpuFftPhi through the mirror: the engine on
    the coarse azimuth grid (nLow = nFull/2 here) and the phi upsampling,
    against the oracle on the same coarse grid upsampled the same way; the
    interpolation error against the direct full grid is printed."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    _, batch, consts = _setup(pkg, fx)
    theta = np.radians(np.r_[0.0, np.linspace(0.5, 179.5, 30), 180.0])
    grid = abi.Grid.build(theta, 64)
    h = pkg.handler.HandlerPOGpu(grid, consts)
    h.SetFullOnly(False)
    Mg = np.zeros((grid.n_az + 1, grid.n_zen + 1, 16))
    Mns = np.zeros_like(Mg)
    assert h.HandleOrientationsToLocalGpuFftPhi(batch, Mg, Mns)
    low = abi.Grid.build(theta, 32)
    o, on, _ = oracle.diffract(low, batch, consts, noshadow=True)
    want = fourier.upsample_phi(o, grid.n_az)
    want_ns = fourier.upsample_phi(on, grid.n_az)
    check_fp64(Mg[:-1], want, theta, "phi-FFT path vs oracle on the coarse grid, upsampled")
    check_fp64(Mns[:-1], want_ns, theta, "phi-FFT path no-shadow")
    direct, _, _ = oracle.diffract(grid, batch, consts)
    print("phi-FFT interpolation vs direct 64-azimuth grid: %s"
          % M.summary(Mg[:-1], direct, theta))
    # the direct path below the size limit, and the options it does not provide
    small = abi.Grid.build(theta, 24)
    h2 = pkg.handler.HandlerPOGpu(small, consts)
    M2 = np.zeros((24, grid.n_zen + 1, 16))
    assert h2.HandleOrientationsToLocalGpuFftPhi(batch, M2, None)
    check_fp64(M2, oracle.diffract(small, batch, consts)[0], theta, "phi-FFT below 32 azimuths")
    h.isCoh = False
    assert not h.HandleOrientationsToLocalGpuFftPhi(batch, Mg, Mns)
    h.close()
    h2.close()


def test_synthetic_large_batch_against_oracle(pkg, oracle, ctx):
    """A resampled batch at the orientation-average particle size (D=50 um) on
    a reduced grid (the oracle is too slow for the full one)."""
    syn = pkg.synthetic
    fx = load_fixture("sobol_average")
    src = pkg.abi.BeamBatch.from_fixture(fx)
    consts = pkg.abi.consts_from_fixture(fx)
    batch = syn.resample(src, [600, 900], scale=5.0, seed=3)
    grid = syn.uniform_grid(90, 12)
    m, _ = _run(ctx, grid, batch, consts)
    o, _, _ = oracle.diffract(grid, batch, consts)
    check_fp64(m, o, grid.theta, "synthetic D=50 FP64 vs oracle")


def _aimed_at_poles(abi, src, seed):
    """One orientation holding every beam of `src`, each rigidly rotated so
    that it leaves along theta = 0 (even beams) or theta = pi (odd beams),
    then by a random angle about that axis."""
    rng = np.random.default_rng(seed)
    bs = src.beam_scalars.copy()
    for b in range(src.n_beams):
        t = np.array([0.0, 0.0, -1.0 if b % 2 == 0 else 1.0])
        d = bs[b, abi.B_DIR:abi.B_DIR + 3] / np.linalg.norm(bs[b, abi.B_DIR:abi.B_DIR + 3])
        v, c = np.cross(d, t), float(d @ t)
        K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        R = np.eye(3) + K + K @ K / (1 + c) if c > -0.5 else np.diag([1.0, -1.0, -1.0]) @ (
            np.eye(3) - K + K @ K / (1 - c))   # d -> -t first, then pi about x
        a = rng.uniform(0, 2 * np.pi)
        Kt = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
        R = (np.cos(a) * np.eye(3) + np.sin(a) * Kt + (1 - np.cos(a)) * np.outer(t, t)) @ R
        for col in (abi.B_HOR, abi.B_VER, abi.B_DIR, abi.B_CEN, abi.B_NT, abi.B_NP, abi.B_NXDT,
                    abi.B_NXDP, abi.B_NORMAL):
            bs[b, col:col + 3] = R @ bs[b, col:col + 3]
    return abi.BeamBatch([0, src.n_beams], [1.0], bs, src.beam_nv, src.beam_external,
                         src.vert_offsets, src.vx, src.vy, src.slope_yx, src.slope_xy,
                         src.edge_flags)


def test_small_phase_cells_near_the_poles(pkg, oracle, ctx):
    """Every beam aimed at a pole, theta rows log-spaced from 1e-9 to 1e-2 rad
    off it: each beam meets cells on both sides of the small-phase switch
    (every vertex phase within 1e-3 rad) and near-singular edges everywhere
    near its pole, and the kernels' per-beam bound that skips the vertex test
    must leave every such decision to it.

    Bars.  FP64: grid L2 1e-10 and per cell 1e-6 of M11.  Rows within ~1e-4
    rad of a beam's direction are ill-conditioned: the phases there are
    differences of O(k size) numbers, and a 1-ulp change of the beam
    directions moves the oracle's own result by up to 4e-8 of M11 in them; the
    engine, which folds the direction phase into the vertex phases as the
    reference's CUDA path does, rounds differently (2.9e-7 measured).
    Mixed: the P11 grid L2 bar of the path (1e-4), and every azimuth-averaged
    P11 row within 1e-4.  The FP32 edge sum cancels to ~phase^2 of its terms
    (error ~1e-7/phase^2), so the mixed kernels take the fifth-order moment
    form up to 0.25 rad; the per-cell P11 error is printed."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    consts = abi.consts_from_fixture(fx)
    batch = _aimed_at_poles(abi, abi.BeamBatch.from_fixture(fx).subset([0, 1]), seed=5)
    off = np.logspace(-9, -2, 36)
    grid = abi.Grid.build(np.r_[0.0, off, np.pi - off[::-1], np.pi], 8)
    # the decision on the host: phases k (A x + B y) of every beam and cell
    th, ph = np.meshgrid(grid.theta, np.arange(grid.n_az) * grid.az_step, indexing="ij")
    d = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)], -1)
    n_small = n_large = 0
    for b in range(batch.n_beams):
        s = batch.beam_scalars[b]
        hor, ver = s[abi.B_HOR:abi.B_HOR + 3], s[abi.B_VER:abi.B_VER + 3]
        bd = s[abi.B_DIR:abi.B_DIR + 3]
        A, B = bd @ hor - d @ hor, bd @ ver - d @ ver
        v = slice(batch.vert_offsets[b], batch.vert_offsets[b + 1])
        p = consts.wave_index * (A[..., None] * batch.vx[v] + B[..., None] * batch.vy[v])
        small = np.abs(p).max(-1) <= 1e-3
        n_small += int(small.any())
        n_large += int((~small).any())
    print("beams with small-phase cells %d, with others %d, of %d"
          % (n_small, n_large, batch.n_beams))
    assert n_small == batch.n_beams and n_large == batch.n_beams
    m, mn = _run(ctx, grid, batch, consts, noshadow=True)
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    for a, b, label in ((m, o, "FP64"), (mn, on, "FP64 no-shadow")):
        err = np.abs(a - b) / np.abs(b[..., :1])
        p, t, el = np.unravel_index(np.argmax(err), err.shape)
        print("beams at the poles %s vs oracle: %s; largest at theta %.6g rad, element %d"
              % (label, M.summary(a, b, grid.theta), grid.theta[t], el))
        assert M.grid_l2(a, b, grid.theta) < 1e-10
        assert M.cell(a, b) < 1e-6
    mm, _ = _run(ctx, grid, batch, consts, flags=abi.PREC_MIXED)
    l2 = M.grid_l2(mm[..., :1], o[..., :1], grid.theta)
    e = np.abs(mm[..., 0] - o[..., 0]) / np.abs(o[..., 0])
    a_m, a_o = oracle.azimuth_average(grid, mm), oracle.azimuth_average(grid, o)
    avg11 = float(np.max(np.abs(a_m[:, 2] - a_o[:, 2]) / np.abs(a_o[:, 2])))
    print("beams at the poles mixed vs oracle: P11 grid_l2 %.2e; azimuth-avg P11 max rel "
          "%.2e; per-cell P11 max rel %.2e; per row: %s"
          % (l2, avg11, e.max(), np.array2string(e.max(axis=0), precision=1)))
    assert l2 < 1e-4
    assert avg11 < 1e-4


def _quat_matrix(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("kind", [0, 1, 2], ids=["sobol", "hammersley", "shoemake"])
def test_orientation_sets_on_device(pkg, oracle, ctx, kind):
    """Device orientation sets equal the oracle's (Sobol bits exactly: gamma is
    gamma_sym * v with identical v), and the reference's own Sobol2D points."""
    for n, seed, bs, gs in ((4096, 1, np.pi / 2, np.pi / 3), (1000, 42, np.pi, 2 * np.pi)):
        rows = ctx.orientation_set(kind, n, seed=seed, beta_sym=bs, gamma_sym=gs)
        want = oracle.orientation_set(kind, n, seed=seed, beta_sym=bs, gamma_sym=gs)
        np.testing.assert_array_equal(rows[:, 2], want[:, 2])
        if kind == 2:
            np.testing.assert_allclose(rows[:, 3:7], want[:, 3:7], rtol=0, atol=5e-16)
        else:
            np.testing.assert_array_equal(rows[:, 1], want[:, 1])
            np.testing.assert_allclose(rows[:, 0], want[:, 0], rtol=4e-16, atol=0)
            assert np.all(rows[:, 6] >= 0)
        np.testing.assert_allclose(rows[:, 7:], want[:, 7:], rtol=0, atol=1e-15)
        for r in rows[::97]:
            np.testing.assert_allclose(_quat_matrix(r[3:7]), r[7:].reshape(3, 3), atol=2e-15)
    if kind == 0:
        fx = load_fixture("sobol_points")
        rows = ctx.orientation_set(0, 4096, seed=int(fx["seeds"][0]), beta_sym=np.pi,
                                   gamma_sym=2 * np.pi)
        np.testing.assert_allclose(rows[:, 1] / (2 * np.pi), fx["y"][0], rtol=2e-16, atol=0)
    with pytest.raises(pkg.engine.EngineError):
        ctx.orientation_set(kind, 0)
    with pytest.raises(pkg.engine.EngineError):
        ctx.orientation_set(7, 10)


@pytest.mark.parametrize("n", [64, 256, 1024, 2048])
def test_aperture_fft_on_device(pkg, oracle, ctx, n):
    """Device aperture FFT (FP32) of real beam apertures against the oracle's
    FP64 FFT of the identical raster: spectra to FP32 accuracy, sampled
    values to 1e-5 of the aperture area; and, through the oracle, against the
    exact edge-sum transform.  n = 2048 runs a full 64-beam batch (2 GiB)."""
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    k = consts.wave_index
    rng = np.random.default_rng(n)
    nb = 64 if n == 2048 else 8
    beams = rng.choice(batch.n_beams, nb, replace=False)
    ab = rng.uniform(-0.02, 0.02, (32, 2))
    ctx.set_timing(True)
    got = ctx.aperture_fft(beams, n, pad=4.0, ab=ab)
    t = ctx.aperture_fft_timing()
    ctx.set_timing(False)
    print("n=%d batch=%d: rows %.3f ms, cols %.3f ms, sample %.3f ms"
          % (n, nb, t["rows_ms"], t["cols_ms"], t["sample_ms"]))
    vo = batch.vert_offsets
    for i in (0, nb - 1):
        b = beams[i]
        vx, vy = batch.vx[vo[b]:vo[b + 1]], batch.vy[vo[b]:vo[b + 1]]
        want, X = oracle.aperture_fft(vx, vy, n, 4.0, ab, k)
        S = ctx.aperture_fft_spectrum(i, n)
        rel = np.linalg.norm(S - X) / np.linalg.norm(X)
        area = abs(X[0, 0]) * (oracle.aperture_raster(vx, vy, n, 4.0)[1][2]
                               * oracle.aperture_raster(vx, vy, n, 4.0)[1][3])
        err = np.max(np.abs(got[i] - want)) / area
        print("beam %d: spectrum rel L2 %.2e, sampled max |d|/area %.2e" % (b, rel, err))
        assert rel < 1e-5
        assert err < 1e-5
    with pytest.raises(pkg.engine.EngineError):
        ctx.aperture_fft(beams[:1], 1000)
    with pytest.raises(pkg.engine.EngineError):
        ctx.aperture_fft(beams[:1], 4096)
    with pytest.raises(pkg.engine.EngineError):
        ctx.aperture_fft(np.zeros(65, np.int64), 256)
    with pytest.raises(pkg.engine.EngineError):
        ctx.aperture_fft([batch.n_beams], 256)


@pytest.mark.parametrize("n,pad", [(256, 1.0), (256, 1.5), (256, 3.0), (256, 6.0), (512, 8.0)])
def test_aperture_fft_row_window(pkg, oracle, ctx, n, pad):
    """Both passes transform only the window of rows (columns) that can carry
    the polygon: n / pad rounded up to a power of two, at least n / 8, as
    n / L transforms of length L.  Padding 1, 1.5, 3 and 6 at n = 256 give
    L = 256, 256, 128 and 64 (one, one, two and four residues), padding 8 at
    n = 512 gives L = 64 (eight); every spectrum against the oracle's FP64
    FFT of the identical raster, and the sampled values to 1e-5 of the
    aperture area."""
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    k = consts.wave_index
    rng = np.random.default_rng(int(pad * 10) + n)
    beams = rng.choice(batch.n_beams, 6, replace=False)
    ab = rng.uniform(-0.02, 0.02, (16, 2))
    got = ctx.aperture_fft(beams, n, pad=pad, ab=ab)
    vo = batch.vert_offsets
    worst_s = worst_v = 0.0
    for i, b in enumerate(beams):
        vx, vy = batch.vx[vo[b]:vo[b + 1]], batch.vy[vo[b]:vo[b + 1]]
        want, X = oracle.aperture_fft(vx, vy, n, pad, ab, k)
        S = ctx.aperture_fft_spectrum(i, n)
        _, (_, _, dx, dy) = oracle.aperture_raster(vx, vy, n, pad)
        worst_s = max(worst_s, np.linalg.norm(S - X) / np.linalg.norm(X))
        worst_v = max(worst_v, np.max(np.abs(got[i] - want)) / (abs(X[0, 0]) * dx * dy))
    print("n %d pad %.1f: worst spectrum rel L2 %.2e, sampled max |d|/area %.2e"
          % (n, pad, worst_s, worst_v))
    assert worst_s < 1e-5
    assert worst_v < 1e-5


def test_graph_replay_and_autotune(pkg, oracle, ctx):
    """Graph replays equal the eager result on every replay, with eager work
    and synchronisations between replays; the autotuner's choice gives the
    same Mueller grid as the default launch shape (to FP32 rounding for the
    mixed kernel, whose summation order changes with the shape)."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    for flags in (abi.PREC_FP64, abi.PREC_MIXED):
        eager, eager_ns = ctx.mueller(flags, noshadow=True)
        ctx.set_graph(True)
        try:
            for rep in range(4):
                m, mn = ctx.mueller(flags, noshadow=True)
                np.testing.assert_array_equal(m, eager)
                np.testing.assert_array_equal(mn, eager_ns)
                ctx.orientation_set(0, 1000)  # eager work and a sync between replays
        finally:
            ctx.set_graph(False)
    for flags, tol in ((abi.PREC_FP64, 1e-12), (abi.PREC_MIXED, 2e-5)):
        ref, _ = ctx.mueller(flags)
        table = ctx.autotune(flags)
        best = min(table, key=lambda e: e[3])
        print("autotune flags=%d: %d shapes, best %s" % (flags, len(table), best))
        assert len(table) >= 6
        m, _ = ctx.mueller(flags)
        assert M.grid_l2(m, ref, grid.theta) < tol
    ctx.set_target_blocks(65536)  # back to the default shape for later tests


def test_mfma_occupancy_variants_agree(pkg, ctx):
    """The MFMA kernel with its registers held to 4, 5 and 6 waves per SIMD
    (the autotuner's occupancy dimension, applied with mbs_apply_tune_entry)
    computes the same arithmetic: bit-identical grids."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    out = {}
    for w in (4, 5, 6):
        ctx.apply_tune_entry(abi.PREC_MIXED, 65536, abi.WORK_SEGMENTS, 1, w)
        out[w], _ = ctx.mueller(abi.PREC_MIXED)
    np.testing.assert_array_equal(out[4], out[5])
    np.testing.assert_array_equal(out[6], out[5])
    with pytest.raises(pkg.engine.EngineError):
        ctx.apply_tune_entry(abi.PREC_MIXED, 65536, abi.WORK_QUEUE, 1, 6)
    ctx.set_target_blocks(65536)  # drops the applied shape


def test_mlp_gradient_against_oracle(pkg, ctx):
    """Device forward and backward (BF16 MFMA GEMMs, FP32 accumulation)
    against the FP64 oracle evaluated on the same BF16-rounded operands; and
    device Adam training reduces the loss of a smooth target.  The remaining
    differences come from FP32 accumulation order and from the rare BF16
    rounding that an FP32-level difference in an activation flips."""
    from oracle import mlp_oracle as mo
    n_in, width, depth, n_out = 5, 256, 4, 181
    rng = np.random.default_rng(9)
    with pkg.engine.Mlp(ctx, n_in, width, depth, n_out, seed=3) as net:
        shapes = mo.layer_shapes(n_in, width, depth, n_out)
        p = net.params().astype(np.float64)
        x = rng.normal(size=(200, n_in)).astype(np.float32)
        t = rng.normal(size=(200, n_out)).astype(np.float32)
        y = net.forward(x)
        y_o, _ = mo.forward(p, shapes, x, bf=True)
        ey = np.linalg.norm(y - y_o) / np.linalg.norm(y_o)
        loss, g = net.gradient(x, t)
        loss_o, g_o = mo.loss_and_grad(p, shapes, x, t, bf=True)
        eg = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
        print("forward rel L2 %.2e; loss %.6e vs %.6e; gradient rel L2 %.2e"
              % (ey, loss, loss_o, eg))
        assert ey < 1e-3
        assert abs(loss - loss_o) < 1e-4 * loss_o
        assert eg < 2e-3
    # training: a smooth function of the inputs
    xs = rng.uniform(-1, 1, (4096, n_in)).astype(np.float32)
    ts = np.stack([np.sin(xs @ rng.normal(size=n_in)) for _ in range(n_out)], 1).astype(np.float32)
    with pkg.engine.Mlp(ctx, n_in, 64, 2, n_out, seed=5) as net2:
        l0, _ = net2.gradient(xs[:512], ts[:512])
        l1 = net2.train(xs, ts, batch=256, epochs=30, lr=3e-3, seed=1)
        print("training loss %.4e -> %.4e" % (l0, l1))
        assert l1 < 0.1 * l0


def test_staged_upload_matches_plain_upload(pkg, ctx):
    """mbs_stage_beams / mbs_upload_staged (page-locked staging, asynchronous
    copies) give the same device batch as mbs_upload_beams: results bit for
    bit, for a staged batch diffracted at once and for one whose staging
    area is reused two stagings later while the other area is in flight."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, full, consts = _setup(pkg, fx)
    parts = [full.subset([0, 1, 2]), full.subset([3, 4]), full.subset([5, 6, 7])]
    ctx.set_grid(grid)
    want = []
    for b in parts:
        ctx.upload(b, consts)
        want.append(ctx.mueller(abi.PREC_FP64, noshadow=True))
    for i, b in enumerate(parts):
        ctx.upload_staged(b, consts)
        m, mn = ctx.mueller(abi.PREC_FP64, noshadow=True)
        np.testing.assert_array_equal(m, want[i][0])
        np.testing.assert_array_equal(mn, want[i][1])
    # a staged batch that is empty of beams, then a plain upload again
    empty = abi.BeamBatch([0, 0], [1.0], np.zeros((0, 36)), np.zeros(0, np.int32),
                          np.zeros(0, np.int32), [0], np.zeros(0), np.zeros(0), np.zeros(0),
                          np.zeros(0), np.zeros(0, np.uint8))
    ctx.upload_staged(empty, consts)
    assert not ctx.mueller(0)[0].any()
    ctx.upload(parts[0], consts)
    np.testing.assert_array_equal(ctx.mueller(abi.PREC_FP64, noshadow=True)[0], want[0][0])


def test_work_queue_matches_segments(pkg, oracle, ctx):
    """The device-planned, persistent work-queue launch (MBS_WORK_QUEUE)
    makes the host plan's segments and writes one partial per segment, so it
    equals the host-planned segment launch bit for bit -- FP64, mixed with
    MFMA and with the vector-ALU sum, full and no-shadow -- and the oracle."""
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid, batch, consts = _setup(pkg, fx)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    for flags in (abi.PREC_FP64, abi.PREC_MIXED, abi.PREC_MIXED | abi.NO_MFMA):
        out = {}
        for mode in (abi.WORK_SEGMENTS, abi.WORK_QUEUE):
            ctx.set_work_mode(mode)
            out[mode] = ctx.mueller(flags, noshadow=True)
        ctx.set_work_mode(0)
        for k in range(2):
            np.testing.assert_array_equal(out[abi.WORK_QUEUE][k], out[abi.WORK_SEGMENTS][k])
    ctx.set_work_mode(abi.WORK_QUEUE)
    m, _ = ctx.mueller(abi.PREC_FP64)
    ctx.set_work_mode(0)
    o, _, _ = oracle.diffract(grid, batch, consts)
    check_fp64(m, o, grid.theta, "work queue vs oracle")


def test_queue_graph_replays_new_uploads(pkg, ctx):
    """16384 orientations with beam counts spread log-uniformly over 5..400
    in one persistent launch under graph replay.  The segment plan is made on
    the device inside the graph, so after a new upload with other beam counts
    (same orientation count, within the buffers) the same graph replays --
    one capture for every launch -- and follows the new batch: every replay
    equals the eager host-planned launch of its batch bit for bit, with eager
    work and synchronisations between the replays."""
    abi, syn = pkg.abi, pkg.synthetic
    fx = load_fixture("sobol_average")
    src, consts = abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx)
    grid = syn.uniform_grid(36, 8)
    n_orient = 16384
    rng = np.random.default_rng(3)
    batches = []
    for seed, hi in ((1, 400), (2, 300)):
        counts = np.exp(rng.uniform(np.log(5), np.log(hi), n_orient)).astype(np.int64)
        batches.append(syn.resample(src, counts, seed=seed))
    print("batches: %d orientations, %d and %d beams"
          % (n_orient, batches[0].n_beams, batches[1].n_beams), flush=True)
    ctx.set_grid(grid)
    want = []
    for b in batches:
        ctx.upload(b, consts)
        ctx.set_work_mode(abi.WORK_SEGMENTS)
        want.append(ctx.mueller(abi.PREC_MIXED)[0])
    ctx.set_work_mode(abi.WORK_QUEUE)
    ctx.set_graph(True)
    ctx.upload(batches[0], consts)   # the larger batch first: later uploads fit its buffers
    ctx.mueller(abi.PREC_MIXED)
    c0, l0 = ctx.graph_stats()
    try:
        for rep in range(3):
            for i, b in enumerate(batches):
                ctx.upload(b, consts)
                got = ctx.mueller(abi.PREC_MIXED)[0]
                np.testing.assert_array_equal(got, want[i])
                # eager work of the program between the replays: the
                # azimuth average of the result, with a synchronisation
                rows = ctx.azimuth_average()
                assert np.all(np.isfinite(rows))
        c1, l1 = ctx.graph_stats()
        print("graph: %d capture(s) for %d launches" % (c1 - c0 + 1, l1 - l0 + 1))
        assert c1 == c0 and l1 - l0 == 6   # six replays of the one graph
    finally:
        ctx.set_graph(False)
        ctx.set_work_mode(0)
