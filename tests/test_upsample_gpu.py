"""Fourier upsampling of a result grid on the device (mbs_upsample_result),
the interpolation step of the reference's angular FFT path, against the
reference's rule restated with numpy's FFT (tests/fourier.py).

Bar: 1e-12 of the grid's largest |M11| per cell (both are FP64 FFTs).
Cases: even and odd azimuth counts, prime line lengths (13 -> 29), factors
2 and 3 in zenith, zenith then azimuth together, no-shadow grids, the
azimuth-averaged rows of the upsampled grid, and the size limits."""
import numpy as np
import pytest

from conftest import load_fixture
from fourier import upsample_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.engine.Context(0)
    yield c
    c.close()


def _result(pkg, ctx, grid, noshadow=False):
    fx = load_fixture("sobol_average")
    batch, consts = pkg.abi.BeamBatch.from_fixture(fx), pkg.abi.consts_from_fixture(fx)
    ctx.set_grid(grid)
    ctx.upload(batch.subset([0, 1, 2]), consts)
    return ctx.mueller(0, noshadow)


@pytest.mark.parametrize("n_az, n_zen, n_az_out, n_zen_out", [
    (24, 60, 48, 60),      # azimuth x2, even count
    (24, 60, 96, 60),      # azimuth x4
    (45, 60, 90, 60),      # odd count
    (13, 30, 29, 30),      # prime lengths
    (16, 60, 16, 120),     # zenith x2
    (16, 60, 16, 180),     # zenith x3
    (18, 40, 54, 80),      # both
])
def test_upsampling_matches_the_reference_rule(pkg, oracle, ctx, n_az, n_zen, n_az_out,
                                               n_zen_out):
    syn = pkg.synthetic
    grid = syn.uniform_grid(n_zen, n_az)
    m, mn = _result(pkg, ctx, grid, noshadow=True)
    th_out = np.linspace(0.0, np.pi, n_zen_out + 1)
    for src, ns in ((m, False), (mn, True)):
        want = upsample_grid(src, n_az_out, n_zen_out)
        got, rows = ctx.upsample(n_az_out, n_zen_out, noshadow=ns, theta_out=th_out, rows=True)
        scale = np.abs(want[..., 0]).max()
        err = np.abs(got - want).max() / scale
        print("%d x %d -> %d x %d (noshadow %s): max |d| / max|M11| %.2e"
              % (n_az, n_zen + 1, n_az_out, n_zen_out + 1, ns, err))
        assert err < 1e-12
        # samples of the input grid pass through (the interpolation is exact there)
        if n_zen_out % n_zen == 0 and n_az_out % n_az == 0:
            sub = got[::n_az_out // n_az, ::n_zen_out // n_zen]
            assert np.abs(sub - src).max() / scale < 1e-12
        g = pkg.abi.Grid.build(th_out, n_az_out)
        want_rows = oracle.azimuth_average(g, want)
        # theta in degrees on the device (the .dat column), radians in the oracle
        np.testing.assert_allclose(rows[:, 0], np.degrees(want_rows[:, 0]), rtol=1e-14,
                                   atol=1e-12)
        np.testing.assert_allclose(rows[:, 1], want_rows[:, 1], rtol=1e-13, atol=0)
        assert np.abs(rows[:, 2:] - want_rows[:, 2:]).max() / np.abs(want_rows[:, 2]).max() < 1e-12


def test_upsampling_limits(pkg, ctx):
    grid = pkg.synthetic.uniform_grid(20, 16)
    _result(pkg, ctx, grid)
    E = pkg.engine.EngineError
    for args in ((8, 20), (16, 10)):                    # sizes may only grow
        with pytest.raises(E):
            ctx.upsample(*args)
    with pytest.raises(E) as e:                         # longer lines than the LDS holds
        ctx.upsample(4100, 20)
    assert e.value.status == pkg.abi.MBS_ERR_CAPACITY
    with pytest.raises(E):                              # no no-shadow result
        ctx.upsample(32, 20, noshadow=True)
    same, _ = ctx.upsample(16, 20)                      # nothing to do: the grid itself
    m = np.zeros_like(same)
    ctx.accumulate(m.reshape(-1))
    np.testing.assert_array_equal(same, m)
