"""The batched non-convex aggregate through the engine: the 8-column aggregate
(seed 7, D ~ U(5, 30) um, L/D ~ U(1, 4), m=1.3116+1e-4i, 6 reflections).

* The reference tracer's own beams (tests/golden/config3_aggregate.npz: 2
  Sobol orientations, seed 1, 9055 beams with up to 12 vertices) against the
  oracle (FP64 bars as in test_engine_gpu.py; mixed: P11 grid L2 1e-4, every
  azimuth-averaged P11 row where P11 is at least 1e-6 of the forward peak
  within 1e-4, the deeper ones within 5e-4) and against the reference CLI's
  own azimuth-averaged output (gate L2 1e-6).
* A batch of the configuration's shape: 16,384 orientations with uneven beam
  counts (log-uniform over 50..400 here, a synthetic resample of the
  aggregate's traced beams; bench.py --config3 runs the stated 50..40k) on
  configuration 2's 541 x 96 grid, in one persistent work-queue launch under
  graph replay.  Properties: every replay equals the eager beam-segment launch
  bit for bit, and the result equals the sum of its two halves (orientations
  0..8191 and 8192..16383 uploaded as batches of their own) to rounding."""
import numpy as np
import pytest

import metrics as M
from conftest import load_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.engine.Context(0)
    yield c
    c.close()


def test_aggregate_against_oracle_and_reference(pkg, oracle, ctx):
    abi = pkg.abi
    fx = load_fixture("config3_aggregate")
    grid, batch = abi.Grid.from_fixture(fx), abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    print("aggregate: %d orientations, %d beams, nv %d..%d, %d directions"
          % (batch.n_orient, batch.n_beams, batch.beam_nv.min(), batch.beam_nv.max(),
             grid.cells), flush=True)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    m, mn = ctx.mueller(abi.PREC_FP64, noshadow=True)
    o, on, _ = oracle.diffract(grid, batch, consts, noshadow=True)
    for a, b, label in ((m, o, "full"), (mn, on, "no-shadow")):
        print("aggregate FP64 %s vs oracle: %s" % (label, M.summary(a, b, grid.theta)))
        assert M.grid_l2(a, b, grid.theta) < 1e-10
        assert M.row(a, b) < 2e-8
        assert M.cell(a, b) < 2e-7
    for mm, key in ((m, "ref_total_dat"), (mn, "ref_total_dat_noshadow")):
        g = M.gate_l2(oracle.azimuth_average(grid, mm), fx[key])
        print("aggregate FP64 %s gate_l2 vs reference .dat: %.2e" % (key, g))
        assert g < 1e-6
    mm, _ = ctx.mueller(abi.PREC_MIXED)
    a_m, a_o = oracle.azimuth_average(grid, mm), oracle.azimuth_average(grid, o)
    l2 = M.grid_l2(mm[..., :1], o[..., :1], grid.theta)
    rel = np.abs(a_m[:, 2] - a_o[:, 2]) / np.abs(a_o[:, 2])
    depth = np.abs(a_o[:, 2]) / np.abs(a_o[:, 2]).max()
    worst = int(np.argmax(rel))
    print("aggregate mixed vs oracle: P11 grid_l2 %.2e, azimuth-avg P11 max rel %.2e at theta "
          "%.1f deg (P11 there %.1e of the peak); rows at >= 1e-6 of the peak: max rel %.2e"
          % (l2, rel[worst], np.degrees(grid.theta[worst]), depth[worst],
             rel[depth >= 1e-6].max()))
    assert l2 < 1e-4
    # Each beam's FP32 F carries ~2e-7 of its own P11 (measured beam by beam);
    # where the coherent sum of the ~4,500 beams of an orientation cancels to
    # below 1e-6 of the forward peak (the exact backscatter pole of this
    # 2-orientation average: 6.8e-7) that rounding reaches ~1e-4 of the row
    assert rel[depth >= 1e-6].max() < 1e-4
    assert rel.max() < 5e-4


def _batch(pkg, counts, seed):
    abi, syn = pkg.abi, pkg.synthetic
    fx = load_fixture("config3_aggregate")
    src = abi.BeamBatch.from_fixture(fx)
    return syn.resample(src, counts, seed=seed), abi.consts_from_fixture(fx)


def test_aggregate_16384_orientations_one_launch(pkg, ctx):
    abi, syn = pkg.abi, pkg.synthetic
    n_orient = 16384
    rng = np.random.default_rng(7)
    counts = np.exp(rng.uniform(np.log(50), np.log(400), n_orient)).astype(np.int64)
    batch, consts = _batch(pkg, counts, seed=7)
    grid = syn.config2_grid()
    print("batch: %d orientations, %d beams (%d..%d per orientation), %d directions"
          % (n_orient, batch.n_beams, counts.min(), counts.max(), grid.cells), flush=True)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_work_mode(abi.WORK_SEGMENTS)
    want = ctx.mueller(abi.PREC_MIXED)[0]
    ctx.set_work_mode(abi.WORK_QUEUE)
    ctx.set_graph(True)
    try:
        for rep in range(2):
            got = ctx.mueller(abi.PREC_MIXED)[0]
            np.testing.assert_array_equal(got, want)
            assert np.all(np.isfinite(ctx.azimuth_average()))  # eager work between replays
    finally:
        ctx.set_graph(False)
        ctx.set_work_mode(0)
    # the two halves as batches of their own
    half = n_orient // 2
    parts = []
    for lo, hi in ((0, half), (half, n_orient)):
        b0, b1 = batch.orient_offsets[lo], batch.orient_offsets[hi]
        v0, v1 = batch.vert_offsets[b0], batch.vert_offsets[b1]
        sub = abi.BeamBatch(batch.orient_offsets[lo:hi + 1] - b0, batch.weights[lo:hi],
                            batch.beam_scalars[b0:b1], batch.beam_nv[b0:b1],
                            batch.beam_external[b0:b1], batch.vert_offsets[b0:b1 + 1] - v0,
                            batch.vx[v0:v1], batch.vy[v0:v1], batch.slope_yx[v0:v1],
                            batch.slope_xy[v0:v1], batch.edge_flags[v0:v1])
        ctx.upload(sub, consts)
        parts.append(ctx.mueller(abi.PREC_MIXED)[0])
    s = parts[0] + parts[1]
    scale = np.abs(want[..., :1]).max()
    err = float(np.abs(s - want).max() / scale)
    print("sum of the halves vs the whole: max |d| / max |M11| = %.2e" % err)
    assert err < 1e-5
