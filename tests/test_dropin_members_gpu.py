"""The drop-in's HandlerPO members that the reference CLI does not reach on a
one-GPU host, called from a C++ harness (tests/dropin_harness.cpp, built
into oracle/_ref/dropin_harness) on the reference's prepared beams of the
sobol_average fixture:

* HandleOrientationsToLocalGpuOnDevice with a pack-cache token (the
  reference's packed-beam cache, src/cuda/GpuDiffraction.cu:4274-4282): the
  same token, start, count, scale and wave index reuse the resident batch on
  a new grid -- bit-identical to a fresh pack, with no upload -- and any
  other start packs afresh;
* DiffractThetasGpu, the theta probe (GpuDiffraction.cu:4780-4965), against
  the oracle's incoherent sum on the probe grid (theta at a pole moved 1e-6
  rad inside, vf = (-sin phi, cos phi, 0)) scaled as the probe scales it:
  4 x the sum over azimuths / (nAz + 1).
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import metrics as M
from conftest import load_fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "dropin_harness")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(HARNESS),
                                 reason="drop-in harness (oracle/_ref) is not built")]


def _write_batch(path, batch, wave, n_re, n_im):
    with open(path, "wb") as f:
        f.write(b"MBSBAT01")
        np.array([batch.n_orient, batch.n_beams], np.int32).tofile(f)
        np.array([batch.vert_offsets[-1]], np.int64).tofile(f)
        np.array([wave, n_re, n_im], np.float64).tofile(f)
        for a, t in ((batch.orient_offsets, np.int64), (batch.weights, np.float64),
                     (batch.beam_scalars, np.float64), (batch.beam_nv, np.int32),
                     (batch.beam_external, np.int32), (batch.vert_offsets, np.int64),
                     (batch.vx, np.float64), (batch.vy, np.float64),
                     (batch.slope_yx, np.float64), (batch.slope_xy, np.float64),
                     (batch.edge_flags, np.uint8)):
            np.ascontiguousarray(a, t).tofile(f)


def _case(pkg):
    fx = load_fixture("sobol_average")
    abi = pkg.abi
    batch, consts = abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx)
    wave = 2.0 * np.pi / consts.wave_index
    wd = tempfile.mkdtemp(prefix="mbs_members_")
    path = os.path.join(wd, "batch.bin")
    # the fixture's particle: m = 1.3116 + 1e-4 i
    _write_batch(path, batch, wave, 1.3116, 1e-4)
    return batch, consts, wd, path


def _harness(args, wd):
    env = dict(os.environ, MBS_GPU_TIMING="1", MBS_HIP_PRECISION="fp64")
    p = subprocess.run([HARNESS] + args, cwd=wd, env=env, stdin=subprocess.DEVNULL,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, "harness rc %d:\n%s\n%s" % (p.returncode, p.stdout[-2000:],
                                                        p.stderr[-2000:])
    return p.stderr


def test_pack_cache_token(pkg, oracle):
    abi = pkg.abi
    batch, consts, wd, path = _case(pkg)
    out = os.path.join(wd, "token.bin")
    log = _harness(["token", path, out], wd)
    # how every call ran: the GPU timing line that follows its label
    calls = dict(re.findall(r"call (\S+)\n(GPU timing: [^\n]*)", log))
    for label, line in sorted(calls.items()):
        print("%-15s %s" % (label, line))
    upload = {k: float(m.group(1)) for k, v in calls.items()
              for m in [re.search(r"upload ([0-9.]+) ms", v)] if m}
    assert upload["token-B"] == 0.0                 # the cached batch: no upload
    for label in ("token-A", "token-B-start1"):     # a token names a batch packed whole
        assert upload[label] > 0.0
    for label in ("fresh-B", "fresh-B-start1"):     # no token: pipelined staging
        assert "page-locked staging" in calls[label]
    g = np.fromfile(out, dtype=np.float64).reshape(4, 16, 46, 16)
    fresh, cached, other_tok, other_fresh = g
    np.testing.assert_array_equal(cached, fresh)
    np.testing.assert_array_equal(other_tok, other_fresh)
    assert np.abs(other_fresh - fresh).max() > 1e-6 * np.abs(fresh).max()
    grid = abi.Grid.build(np.linspace(0.0, np.pi / 2, 46), 16)
    o, _, _ = oracle.diffract(grid, batch, consts)
    print("cached batch on grid B vs oracle: " + M.summary(cached, o, grid.theta))
    assert M.grid_l2(cached, o, grid.theta) < 1e-10
    o1, _, _ = oracle.diffract(grid, batch.subset(list(range(1, batch.n_orient))), consts)
    assert M.grid_l2(other_tok, o1, grid.theta) < 1e-10


def test_theta_probe_member(pkg, oracle):
    abi = pkg.abi
    batch, consts, wd, path = _case(pkg)
    out = os.path.join(wd, "probe.bin")
    n_phi = 24
    theta = np.r_[0.0, np.radians([0.25, 1, 3, 10, 22, 46, 90, 135, 170, 179.75]), np.pi]
    _harness(["probe", path, out, str(n_phi)] + ["%.17g" % t for t in theta], wd)
    m11 = np.fromfile(out, dtype=np.float64)
    th = theta.copy()
    th[th <= 1e-14] = 1e-6
    th[th >= np.pi - 1e-14] = np.pi - 1e-6
    phi = np.arange(n_phi) * (2 * np.pi / n_phi)
    vf = np.zeros((n_phi, len(th), 3))
    vf[:, :, 0] = -np.sin(phi)[:, None]
    vf[:, :, 1] = np.cos(phi)[:, None]
    inc = oracle.incoherent(abi.Grid(th, 2 * np.pi / n_phi, vf), batch, consts)
    want = 4.0 * inc[..., 0].sum(axis=0) / (n_phi + 1)
    print("theta probe member vs oracle: max rel %.2e" % np.max(np.abs(m11 - want) / want))
    np.testing.assert_allclose(m11, want, rtol=1e-10, atol=0)
