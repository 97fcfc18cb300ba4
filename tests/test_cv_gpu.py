"""The learned control variate, end to end on the device, and its ablation.

Beams come from the reference's own tracer and PrepareBeams
(oracle/_ref/mbs_po_dump, the reference CLI with the fixture-writing GPU
slot) for Owen-Sobol orientation sets of a D = 10 um, L = 20 um hexagonal
column (lambda = 0.532 um, m = 1.3116 + 1e-4 i, 6 reflections, 181 theta x 32
phi): seeds 1 and 2 (1024 orientations each) train the network, seed 3 (1024)
is the estimation set and seed 4 (4096) the converged reference.  The
reference samples these sets as TracerPOTotal::TraceFromSobolSeed does with
the hexagonal prism's symmetry domain beta <= pi/2, gamma <= pi/3
(src/particle/Hexagonal.cpp:13); the engine's own orientation-set kernel
reproduces the same orientations (its Sobol points are pinned to the
reference's), which gives the network its inputs.

Checked: with c = 0 the estimator is the plain orientation sum, bit for bit;
the trained network's estimate is compared with the plain one at N = 64 ..
1024 against the converged reference (2 pi dcos-weighted L2 of P11 and of
P_ij / P11) -- the table is printed, and the control variate must not do
worse than the plain sum overall.  Parity of the network's gradients with
the oracle is in test_engine_gpu.py.
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP_BIN = os.path.join(ROOT, "oracle", "_ref", "mbs_po_dump")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DUMP_BIN),
                                 reason="reference CLI build (oracle/_ref) is not present")]

BETA_SYM, GAMMA_SYM = np.pi / 2, np.pi / 3


def _trace(wd, n, seed):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_fixtures
    d = os.path.join(wd, "s%d" % seed)
    os.makedirs(d)
    dump = os.path.join(d, "dump.bin")
    threads = max(1, min(16, os.cpu_count() or 1))
    env = dict(os.environ, MBS_FIXTURE_DUMP=dump, MBS_FIXTURE_NO_EXPECTED="1")
    subprocess.run([DUMP_BIN, "--method", "po", "--particle", "1", "20", "10",
                    "--refractive-index", "1.3116", "1e-4", "--wavelength-um", "0.532",
                    "--max-reflections", "6", "--sobol-seed", str(n), str(seed),
                    "--scattering-grid", "0", "180", "32", "180", "--cutoff-profile", "off",
                    "--threads", str(threads), "--close", "--backend", "gpu", "--output", "dmp"],
                   cwd=d, env=env, stdin=subprocess.DEVNULL, capture_output=True, check=True,
                   timeout=600)
    fx = make_fixtures.read_dump(dump)
    os.remove(dump)
    return fx


def _orientation_rows(pkg, ctx, fx):
    """Per-orientation azimuth-averaged Mueller rows [n, 181, 16], weights
    divided out, computed by the engine (FP64)."""
    abi = pkg.abi
    grid, batch = abi.Grid.from_fixture(fx), abi.BeamBatch.from_fixture(fx)
    ctx.set_grid(grid)
    ctx.upload(batch, abi.consts_from_fixture(fx))
    ctx.diffract(abi.PREC_FP64 | abi.PER_ORIENT)
    rows = ctx.azimuth_average_per_orient()
    return pkg.cv.unweighted_rows(rows, batch.weights), rows[0, :, 1]


def _err(est, truth, w):
    e = [np.sqrt(np.sum(w * (est[:, 0] - truth[:, 0]) ** 2) / np.sum(w * truth[:, 0] ** 2))]
    for j in range(1, 16):
        d = est[:, j] / est[:, 0] - truth[:, j] / truth[:, 0]
        e.append(np.sqrt(np.sum(w * d * d) / np.sum(w)))
    return np.array(e)


def test_control_variate_ablation(pkg, oracle):
    cvmod = pkg.cv
    wd = tempfile.mkdtemp(prefix="mbs_cv_")
    t0 = time.perf_counter()
    sets = {s: _trace(wd, n, s) for s, n in ((1, 1024), (2, 1024), (3, 1024), (4, 4096))}
    print("traced 7168 orientations (reference tracer, host) in %.1f s"
          % (time.perf_counter() - t0), flush=True)
    with pkg.engine.Context(0) as ctx:
        t0 = time.perf_counter()
        M, w2pi = {}, None
        for s, fx in sets.items():
            M[s], w2pi = _orientation_rows(pkg, ctx, fx)
        t_engine = time.perf_counter() - t0
        X = {s: cvmod.features(ctx.orientation_set(0, len(M[s]), seed=s, beta_sym=BETA_SYM,
                                                   gamma_sym=GAMMA_SYM), np.log(10.0))
             for s in M}
        cv = cvmod.ControlVariate(ctx, M[1].shape[1], width=256, depth=4, seed=11)
        try:
            tot_plain, tot_cv = _ablation(cv, ctx, M, X, w2pi, t_engine)
        finally:
            cv.close()
    assert tot_cv <= 1.05 * tot_plain


def _ablation(cv, ctx, M, X, w2pi, t_engine):
    """Trains the network, then prints and sums the errors of the plain and
    control-variate estimates for growing N."""
    t0 = time.perf_counter()
    loss = cv.fit(np.concatenate([X[1], X[2]]), np.concatenate([M[1], M[2]]),
                  epochs=200, batch=128, lr=1e-3, seed=5)
    Eg = cv.expectation(0, 65536, 7, BETA_SYM, GAMMA_SYM, np.log(10.0))
    t_cv = time.perf_counter() - t0
    print("engine (7168 per-orientation grids) %.2f s; network training + E[g] %.2f s, "
          "final loss %.3f" % (t_engine, t_cv, loss))
    truth = M[4].mean(0)
    tot_plain = tot_cv = 0.0
    for n in (64, 128, 256, 512, 1024):
        w = np.full(n, 1.0 / n)
        plain = np.tensordot(w, M[3][:n], axes=1)
        zero = cv.estimate(M[3][:n], X[3][:n], w, Eg, c=0.0)
        np.testing.assert_array_equal(zero, plain)
        est = cv.estimate(M[3][:n], X[3][:n], w, Eg)
        ep, ec = _err(plain, truth, w2pi), _err(est, truth, w2pi)
        tot_plain += ep.max()
        tot_cv += ec.max()
        print("N=%4d  plain: max %.4f P11 %.4f | control variate: max %.4f P11 %.4f"
              % (n, ep.max(), ep[0], ec.max(), ec[0]))
    print("sum over N of the max element error: plain %.4f, control variate %.4f"
          % (tot_plain, tot_cv))
    return tot_plain, tot_cv
