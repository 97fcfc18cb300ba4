"""The large-particle near-forward configuration through the engine: the
hexagonal column D=200 um, L=400 um, m=1.3116+1e-3i, 12 reflections, on the
theta 0-5 deg grid at 0.005 deg (1001 rows) x 48 phi.  The beams are the
reference tracer's own (tests/golden/config5_near_forward.npz: 4 Sobol
orientations, seed 1, 5870 beams; the oracle reproduces that fixture's
reference grids, tests/test_oracle.py).

Bars: FP64 vs the oracle as in test_engine_gpu.py (grid L2 1e-10, row 2e-8,
cell 2e-7); mixed precision: P11 grid L2 1e-4 and every azimuth-averaged P11
row within 1e-4.

The size scan (16 sizes, D log-spaced 10-200 um) re-derives the beam set on
the device for every size (mbs_rescale_beams: vertices and centres times s,
areas times s^2) and checks each size against the oracle on the host-scaled
batch, on every 5th theta row."""
import time

import numpy as np
import pytest

import metrics as M
from conftest import load_fixture

pytestmark = pytest.mark.gpu

N_PHI = 48


def config5_grid(pkg, every=1):
    theta = np.radians(np.round(np.arange(0, 1001, every) * 0.005, 10))
    return pkg.abi.Grid.build(theta, N_PHI)


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.engine.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def case(pkg):
    fx = load_fixture("config5_near_forward")
    return pkg.abi.BeamBatch.from_fixture(fx), pkg.abi.consts_from_fixture(fx)


def _timed(ctx, flags, reps=3):
    ctx.diffract(flags)
    ctx.set_timing(True)
    ms = []
    for _ in range(reps):
        ctx.diffract(flags)
        ms.append(ctx.timing()["kernel_ms"])
    ctx.set_timing(False)
    return float(np.median(ms))


def test_near_forward_grid_against_oracle(pkg, oracle, ctx, case):
    abi = pkg.abi
    batch, consts = case
    grid = config5_grid(pkg)
    assert grid.n_zen == 1000 and grid.theta[-1] == pytest.approx(np.radians(5.0))
    print("config 5: %d orientations, %d beams x %d directions" % (batch.n_orient, batch.n_beams,
                                                                   grid.cells), flush=True)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    m, mn = ctx.mueller(abi.PREC_FP64, noshadow=True)
    t0 = time.perf_counter()
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    t_oracle = time.perf_counter() - t0
    for a, b, label in ((m, o, "full"), (mn, on, "no-shadow")):
        print("config 5 FP64 %s vs oracle: %s" % (label, M.summary(a, b, grid.theta)))
        assert M.grid_l2(a, b, grid.theta) < 1e-10
        assert M.row(a, b) < 2e-8
        assert M.cell(a, b) < 2e-7
    mm, _ = ctx.mueller(abi.PREC_MIXED)
    a_m, a_o = oracle.azimuth_average(grid, mm), oracle.azimuth_average(grid, o)
    l2 = M.grid_l2(mm[..., :1], o[..., :1], grid.theta)
    avg11 = float(np.max(np.abs(a_m[:, 2] - a_o[:, 2]) / np.abs(a_o[:, 2])))
    print("config 5 mixed vs oracle: P11 grid_l2 %.2e, azimuth-avg P11 max rel %.2e, per-cell "
          "P11 max rel %.2e" % (l2, avg11, M.p11_rel(mm, o)))
    assert l2 < 1e-4
    assert avg11 < 1e-4
    units = float(batch.n_beams) * grid.cells
    for name, flags in (("mixed", abi.PREC_MIXED), ("fp64", abi.PREC_FP64)):
        ms = _timed(ctx, flags)
        print("config 5 %s: kernel %.3f ms, %.3g beam.dir evals/s" % (name, ms,
                                                                       units / (ms * 1e-3)))
    print("oracle (CPU, %d threads): %.2f s, %.3g evals/s" % (oracle.threads(), t_oracle,
                                                              units / t_oracle))


def _scaled(abi, batch, s):
    """The host-scaled batch of a size scale s (the rule mbs_rescale_beams
    applies with unit Jones amplitudes: vertices and centres times s, areas
    times s^2, J_phased unchanged)."""
    sc = batch.beam_scalars.copy()
    sc[:, abi.B_CEN:abi.B_CEN + 3] *= s
    sc[:, abi.B_AREA] *= s * s
    return abi.BeamBatch(batch.orient_offsets, batch.weights, sc, batch.beam_nv,
                         batch.beam_external, batch.vert_offsets, batch.vx * s, batch.vy * s,
                         batch.slope_yx, batch.slope_xy, batch.edge_flags)


def test_size_scan_against_oracle(pkg, oracle, ctx, case):
    """16 sizes, D = 10..200 um log-spaced, from one upload of the D=200 um
    beam set: per size the evals/s and the error against the oracle."""
    abi = pkg.abi
    full, consts = case
    batch = full.subset([0])
    grid = config5_grid(pkg, every=5)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    nb = batch.n_beams
    jp = batch.beam_scalars[:, abi.B_JP:abi.B_JP + 8]
    # scaling data that leaves J_phased as uploaded: raw J = J_phased, path 0,
    # sign +1, no absorption paths
    ctx.upload_scaling(jp, np.zeros(nb), np.ones(nb), np.zeros(nb + 1, np.int64), np.zeros(0),
                       0.0, consts.wave_index)
    sizes = np.geomspace(10.0, 200.0, 16)
    units = float(nb) * grid.cells
    worst = (0.0, 0.0)
    for d in sizes:
        s = d / 200.0
        ctx.rescale(s)
        m, _ = ctx.mueller(abi.PREC_FP64)
        o, _, _ = oracle.diffract(grid, _scaled(abi, batch, s), consts)
        e64 = M.grid_l2(m, o, grid.theta)
        mm, _ = ctx.mueller(abi.PREC_MIXED)
        e32 = M.grid_l2(mm[..., :1], o[..., :1], grid.theta)
        ms = _timed(ctx, abi.PREC_MIXED, reps=2)
        print("D %6.2f um: FP64 grid_l2 %.2e, mixed P11 grid_l2 %.2e; mixed %.3g evals/s"
              % (d, e64, e32, units / (ms * 1e-3)))
        assert e64 < 1e-10
        assert e32 < 1e-4
        worst = (max(worst[0], e64), max(worst[1], e32))
    print("size scan worst: FP64 %.2e, mixed P11 %.2e" % worst)
