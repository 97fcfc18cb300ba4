"""The near-forward F source (mbs_spectrum_source + MBS_SPECTRUM) on the
large-particle configuration: the hexagonal column D=200 um, L=400 um, 12
reflections, the reference tracer's beams (tests/golden/config5_near_forward.npz,
4 orientations, 5870 beams) on theta 0-5 deg at 0.005 deg x 48 phi.

Pinned by the exact transform: the F values each selected beam takes from
its spectrum are compared with the oracle's aperture integral (the edge-sum
formula, itself pinned by closed forms in test_oracle.py) at the same
(A, B), times e^{i psi} / complWave (the kernels' convention).  The bar is
relative to the beam's forward value |F(0, 0)| = k^2 area: 1e-4.  The
Mueller grid with the spectrum source is held to the mixed path's P11 grid
L2 bar (1e-4) against the FP64 oracle; its per-row figures are printed
(the spectrum's error is relative to a beam's peak, so rows where those
beams are faint carry more of it).

The spectra are FFTs of the tent-weighted raster (pixel = polygon integral
of the tent of half-width h, exact) with the tent's sinc^2 divided out, read
by non-uniform FFT interpolation (exponential-of-semicircle kernel), the
method of the reference's aperture probe (tools/fft_aperture_probe.cu:
130-169 raster, :549-551 FFT, :50-102 sampling) with an anti-aliased raster."""
import numpy as np
import pytest

import metrics as M
from conftest import load_fixture

pytestmark = pytest.mark.gpu

N_PHI = 48


@pytest.fixture(scope="module")
def case(pkg):
    fx = load_fixture("config5_near_forward")
    abi = pkg.abi
    theta = np.radians(np.round(np.arange(0, 1001) * 0.005, 10))
    return (abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx),
            abi.Grid.build(theta, N_PHI))


@pytest.fixture(scope="module")
def ctx(pkg, case):
    batch, consts, grid = case
    c = pkg.engine.Context(0)
    c.set_grid(grid)
    c.upload(batch, consts)
    yield c
    c.close()


def _directions(grid):
    """Unit directions of the device's cells (theta-major: t * n_az + p)."""
    t = np.repeat(np.arange(grid.n_zen + 1), grid.n_az)
    p = np.tile(np.arange(grid.n_az), grid.n_zen + 1)
    th, ph = grid.theta[t], p * (2 * np.pi / grid.n_az)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)], -1)


def test_spectrum_values_against_exact_transform(pkg, oracle, ctx, case):
    batch, consts, grid = case
    info = ctx.spectrum_source(a_max=0.1, oversample=8.0)
    print("spectrum source: %d of %d beams (%.1f %% of the aperture area), sizes %s, "
          "%.1f MB of spectra; scan %.2f ms, raster+FFT %.2f ms, sampling %.2f ms"
          % (info["n_selected"], batch.n_beams, 100 * info["selected_area"] / info["total_area"],
             {k: v for k, v in info["n_by_size"].items() if v},
             info["spectrum_bytes"] / 1e6, info["scan_ms"], info["fft_ms"], info["sample_ms"]))
    assert info["n_selected"] > 0
    beams, f = ctx.spectrum_values(info["n_selected"])
    d = _directions(grid)
    k = consts.wave_index
    cw = complex(consts.compl_wave[0], consts.compl_wave[1])
    rng = np.random.default_rng(3)
    worst = 0.0
    table = []
    for slot, b in enumerate(beams):
        s = batch.beam_scalars[b]
        hor, ver, bdir, cen, area = s[0:3], s[3:6], s[6:9], s[9:12], s[12]
        v0, v1 = batch.vert_offsets[b], batch.vert_offsets[b + 1]
        cells = rng.choice(grid.cells, 64, replace=False)
        A = bdir @ hor - d[cells] @ hor
        B = bdir @ ver - d[cells] @ ver
        psi = 0.0 if batch.beam_external[b] else None
        err = 0.0
        for j, c in enumerate(cells):
            ph = 0.0 if psi is not None else -k * (d[c] @ cen)
            want = oracle.aperture(batch.vx[v0:v1], batch.vy[v0:v1], batch.slope_yx[v0:v1],
                                   batch.slope_xy[v0:v1], batch.edge_flags[v0:v1], A[j], B[j],
                                   consts, area) / cw * np.exp(1j * ph)
            err = max(err, abs(f[slot, c] - want))
        rel = err / (k * k * abs(area))
        worst = max(worst, rel)
        ext = max(np.ptp(batch.vx[v0:v1]), np.ptp(batch.vy[v0:v1]))
        amax = max(np.abs(bdir @ hor - d @ hor).max(), np.abs(bdir @ ver - d @ ver).max())
        table.append((rel, b, v1 - v0, ext, area, int(batch.beam_external[b]), amax))
    table.sort(reverse=True)
    print("worst beams (|dF|/(k^2 area), beam, nv, extent um, area um^2, external, "
          "max(|A|,|B|)):")
    for t in table[:8]:
        print("  %.2e %5d %2d %8.2f %10.2f %d %.4f" % t)
    print("F from the spectra vs the exact transform: worst |dF| / (k^2 area) %.2e over %d "
          "beams" % (worst, len(beams)))
    assert worst < 1e-4


def test_spectrum_source_mueller(pkg, oracle, ctx, case):
    abi = pkg.abi
    batch, consts, grid = case
    info = ctx.spectrum_source(a_max=0.1, oversample=8.0)
    ms, _ = ctx.mueller(abi.PREC_MIXED | abi.SPECTRUM)
    md, _ = ctx.mueller(abi.PREC_MIXED)
    o, _, _ = oracle.diffract(grid, batch, consts)
    l2 = M.grid_l2(ms[..., :1], o[..., :1], grid.theta)
    l2d = M.grid_l2(md[..., :1], o[..., :1], grid.theta)
    a_s, a_o = oracle.azimuth_average(grid, ms), oracle.azimuth_average(grid, o)
    rows = np.abs(a_s[:, 2] - a_o[:, 2]) / np.abs(a_o[:, 2])
    q = [0, 20, 100, 200, 400, 600, 800, 1000]
    print("spectrum source (%d beams): P11 grid_l2 %.2e (direct mixed %.2e); azimuth-avg P11 "
          "rel error at theta %s deg: %s" % (info["n_selected"], l2, l2d,
                                             [round(float(np.degrees(grid.theta[i])), 3)
                                              for i in q],
                                             " ".join("%.1e" % rows[i] for i in q)))
    assert l2 < 1e-4
    # the other beams are untouched: with nothing selected the variant is
    # the direct mixed path bit for bit
    ctx.spectrum_source(a_max=1e-9)
    m0, _ = ctx.mueller(abi.PREC_MIXED | abi.SPECTRUM)
    np.testing.assert_array_equal(m0, md)


def test_spectrum_source_is_dropped_on_new_grid(pkg, ctx, case):
    abi = pkg.abi
    batch, consts, grid = case
    ctx.spectrum_source(a_max=0.1)
    ctx.set_grid(grid)
    with pytest.raises(pkg.engine.EngineError):
        ctx.diffract(abi.PREC_MIXED | abi.SPECTRUM)
    with pytest.raises(pkg.engine.EngineError):
        ctx.diffract(abi.PREC_FP64 | abi.SPECTRUM)
