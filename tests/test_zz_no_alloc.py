"""The entry point that allocates nothing (mbs_diffract_noalloc) against
mbs_diffract on the same uploaded batch, through the C ABI.

The caller owns the output: the test allocates it with hipMalloc before the
calls, as a caller that cannot allocate later would.  Two calls in a row,
the second on a smaller batch (fewer orientations and beams), so the
library's fixed scratch array is reused with stale contents from the first.

Bars: FP64 -- the two entry points differ only in how the batch is cut into
work items, i.e. in the summation order of the same FP64 terms: grid L2 and
per-cell error 1e-12 (relative to the grid's largest M11).  Mixed precision
-- the FP32 partial sums follow the cut: P11 grid L2 1e-5 (the MFMA vs
vector-ALU agreement bar).
"""
import ctypes as C

import numpy as np
import pytest

import metrics as M
from conftest import load_fixture

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p.value

    def free(self, p):
        assert self.L.hipFree(C.c_void_p(p)) == 0

    def fill(self, p, byte, nbytes):
        assert self.L.hipMemset(C.c_void_p(p), C.c_int(byte), C.c_size_t(nbytes)) == 0

    def to_host(self, p, out):
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p),
                                C.c_size_t(out.nbytes), 2) == 0  # hipMemcpyDeviceToHost
        return out


def test_noalloc_matches_diffract_twice(pkg):
    abi = pkg.abi
    fx = load_fixture("sobol_average")
    grid = abi.Grid.from_fixture(fx)
    big = abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    small = big.subset([0, 1, 2])
    assert small.n_beams < big.n_beams and small.n_orient < big.n_orient
    hip = _Hip()
    cells = grid.cells
    nbytes = cells * 16 * 8
    d_full, d_ns = hip.malloc(nbytes), hip.malloc(nbytes)
    try:
        with pkg.engine.Context(0) as ctx:
            ctx.set_grid(grid)
            for step, (batch, flags, noshadow) in enumerate(
                    [(big, abi.PREC_FP64, True), (small, abi.PREC_FP64, False),
                     (big, abi.PREC_MIXED, False), (small, abi.PREC_MIXED, True)]):
                ctx.upload(batch, consts)
                want, want_ns = ctx.mueller(flags, noshadow)
                # stale bytes in the output: the entry point must overwrite them
                hip.fill(d_full, 0x7f, nbytes)
                hip.fill(d_ns, 0x7f, nbytes)
                print("noalloc step %d: %d orientations, %d beams, flags %d, noshadow %s"
                      % (step, batch.n_orient, batch.n_beams, flags, noshadow), flush=True)
                ctx.diffract_noalloc(flags | (abi.NOSHADOW if noshadow else 0), d_full,
                                     d_ns if noshadow else None)
                got = hip.to_host(d_full, np.empty((grid.n_az, grid.n_zen + 1, 16)))
                pairs = [(got, want)]
                if noshadow:
                    pairs.append((hip.to_host(d_ns, np.empty_like(got)), want_ns))
                for g, w in pairs:
                    print("  " + M.summary(g, w, grid.theta))
                    assert np.all(np.isfinite(g))
                    if flags == abi.PREC_FP64:
                        assert M.grid_l2(g, w, grid.theta) < 1e-12
                        assert np.max(np.abs(g - w)) < 1e-12 * np.max(np.abs(w[..., 0]))
                    else:
                        assert M.grid_l2(g[..., :1], w[..., :1], grid.theta) < 1e-5
    finally:
        hip.free(d_full)
        hip.free(d_ns)


def test_noalloc_rejects_what_does_not_fit(pkg):
    """A grid whose one work item needs more than the fixed array: a clear
    capacity error, no launch."""
    abi = pkg.abi
    fx = load_fixture("fixed_absorbing")
    batch, consts = abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx)
    n_az, n_zen = 2048, 1023  # 2^21 cells: 256 MiB of Mueller grid per item
    theta = np.linspace(0.0, np.pi, n_zen + 1)
    grid = abi.Grid.build(theta, n_az)
    hip = _Hip()
    d_full = hip.malloc(16)
    try:
        with pkg.engine.Context(0) as ctx:
            ctx.set_grid(grid)
            ctx.upload(batch, consts)
            with pytest.raises(pkg.engine.EngineError) as e:
                ctx.diffract_noalloc(abi.PREC_FP64, d_full)
            assert e.value.status == abi.MBS_ERR_CAPACITY
            assert "MBS_NOALLOC_SCRATCH_BYTES" in str(e.value)
    finally:
        hip.free(d_full)
