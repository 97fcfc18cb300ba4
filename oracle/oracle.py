"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never by the engine.  Functions mirror oracle/po_oracle.c, which
cites the reference lines each one restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compiles liboracle.so with gcc (make -C oracle)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.mbs_oracle_diffract_batch.argtypes = [vp, vp, vp, dp, dp, dp]
        L.mbs_oracle_diffract_batch.restype = C.c_int
        L.mbs_oracle_orientation_jones.argtypes = [vp, vp, vp, C.c_int, dp, dp]
        L.mbs_oracle_orientation_jones.restype = C.c_int
        L.mbs_oracle_incoherent_batch.argtypes = [vp, vp, vp, dp]
        L.mbs_oracle_incoherent_batch.restype = C.c_int
        L.mbs_oracle_add_mueller.argtypes = [dp, dp, C.c_double]
        L.mbs_oracle_add_mueller.restype = None
        L.mbs_oracle_azimuth_average.argtypes = [vp, dp, dp]
        L.mbs_oracle_azimuth_average.restype = None
        L.mbs_oracle_sobol.argtypes = [C.c_int, C.c_uint32, dp, dp]
        L.mbs_oracle_sobol.restype = None
        L.mbs_oracle_orientation_set.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_double,
                                                 C.c_double, dp]
        L.mbs_oracle_orientation_set.restype = None
        L.mbs_oracle_aperture_raster.argtypes = [dp, dp, C.c_int, C.c_int, C.c_double, dp, dp]
        L.mbs_oracle_aperture_raster.restype = None
        L.mbs_oracle_2pi_dcos.argtypes = [dp, C.c_int, C.c_int]
        L.mbs_oracle_2pi_dcos.restype = C.c_double
        L.mbs_oracle_small_phase_factor.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double,
                                                    C.c_double, dp, dp]
        L.mbs_oracle_small_phase_factor.restype = C.c_int
        L.mbs_oracle_aperture.argtypes = [dp, dp, dp, dp, C.POINTER(C.c_uint8), C.c_int,
                                          C.c_double, C.c_double, vp, C.c_double, dp, dp]
        L.mbs_oracle_aperture.restype = None
        L.mbs_oracle_set_ref_sincos.argtypes = [C.c_int]
        L.mbs_oracle_set_ref_sincos.restype = None
        L.mbs_oracle_threads.restype = C.c_int
        L.mbs_oracle_check_batch.argtypes = [vp]
        L.mbs_oracle_check_batch.restype = C.c_int
        _LIB = L
    return _LIB


class ref_sincos:
    """Context manager: evaluate phases with the reference CPU's polynomial
    sincos (HandlerPO_fast.h:116-146) instead of libm."""

    def __enter__(self):
        lib().mbs_oracle_set_ref_sincos(1)
        return self

    def __exit__(self, *exc):
        lib().mbs_oracle_set_ref_sincos(0)
        return False


def threads():
    return lib().mbs_oracle_threads()


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def diffract(grid, batch, consts, noshadow=False, per_orient=False):
    """Weighted Mueller sum over the batch's orientations (and optionally the
    no-shadow sum and the per-orientation grids).  Returns (m, m_ns, per)."""
    cells = grid.cells
    m = np.zeros((grid.n_az, grid.n_zen + 1, 16))
    m_ns = np.zeros_like(m) if noshadow else None
    per = np.zeros((batch.n_orient, grid.n_az, grid.n_zen + 1, 16)) if per_orient else None
    rc = lib().mbs_oracle_diffract_batch(
        C.byref(grid.struct()), C.byref(batch.struct()), C.byref(consts), _d(m),
        _d(m_ns) if noshadow else None, _d(per) if per_orient else None)
    if rc != 0:
        raise RuntimeError("oracle status %d" % rc)
    assert m.size == cells * 16
    return m, m_ns, per


def incoherent(grid, batch, consts):
    """sum_o w_o sum_b Mueller(d_b) per cell: single beams' Mueller matrices
    added without interference (po_oracle.c, mbs_oracle_incoherent_batch)."""
    m = np.zeros((grid.n_az, grid.n_zen + 1, 16))
    rc = lib().mbs_oracle_incoherent_batch(C.byref(grid.struct()), C.byref(batch.struct()),
                                           C.byref(consts), _d(m))
    if rc != 0:
        raise RuntimeError("oracle status %d" % rc)
    return m


def orientation_jones(grid, batch, consts, orient, noshadow=False):
    j = np.zeros((grid.n_az, grid.n_zen + 1, 8))
    jn = np.zeros_like(j) if noshadow else None
    rc = lib().mbs_oracle_orientation_jones(C.byref(grid.struct()), C.byref(batch.struct()),
                                            C.byref(consts), orient, _d(j),
                                            _d(jn) if noshadow else None)
    if rc != 0:
        raise RuntimeError("oracle status %d" % rc)
    return j, jn


def mueller_from_jones(j8, weight=1.0):
    """Mueller matrices (..., 16) of Jones matrices given as (..., 8) reals."""
    j8 = np.ascontiguousarray(j8, dtype=np.float64)
    flat = j8.reshape(-1, 8)
    out = np.zeros((flat.shape[0], 16))
    L = lib()
    for i in range(flat.shape[0]):
        L.mbs_oracle_add_mueller(_d(out[i]), _d(flat[i]), weight)
    return out.reshape(j8.shape[:-1] + (16,))


def azimuth_average(grid, m):
    """Rows [theta, 2*pi*dcos, M11..M44] of the azimuth-averaged grid."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros((grid.n_zen + 1, 18))
    lib().mbs_oracle_azimuth_average(C.byref(grid.struct()), _d(m), _d(out))
    return out


def aperture(vx, vy, slope_yx, slope_xy, flags, A, B, consts, area):
    vx, vy = [np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy)]
    syx, sxy = [np.ascontiguousarray(a, dtype=np.float64) for a in (slope_yx, slope_xy)]
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    fr, fi = C.c_double(), C.c_double()
    lib().mbs_oracle_aperture(_d(vx), _d(vy), _d(syx), _d(sxy),
                              fl.ctypes.data_as(C.POINTER(C.c_uint8)), len(vx),
                              A, B, C.byref(consts), area, C.byref(fr), C.byref(fi))
    return complex(fr.value, fi.value)


def sobol(n, seed):
    """First n Owen-scrambled 2-D Sobol points (x, y) of Sobol2D(seed)."""
    x, y = np.zeros(n), np.zeros(n)
    lib().mbs_oracle_sobol(int(n), int(seed), _d(x), _d(y))
    return x, y


def orientation_set(kind, n, seed=1, beta_sym=np.pi, gamma_sym=2 * np.pi):
    """[n, 16] orientation rows (beta, gamma, weight, q (zero for Euler sets),
    R row-major) as mbs_orientation_set defines them."""
    out = np.zeros((int(n), 16))
    lib().mbs_oracle_orientation_set(int(kind), int(n), int(seed), float(beta_sym),
                                     float(gamma_sym), _d(out))
    return out


def aperture_raster(vx, vy, n, pad):
    """(raster [n, n] of 0/1 as float64, (xc, yc, dx, dy)) of a convex polygon,
    as the device aperture FFT rasterises it."""
    vx, vy = [np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy)]
    grid = np.zeros((n, n))
    geom = np.zeros(4)
    lib().mbs_oracle_aperture_raster(_d(vx), _d(vy), len(vx), int(n), float(pad), _d(grid),
                                     _d(geom))
    return grid, geom


def _lanczos(x, radius):
    x = np.asarray(x, dtype=np.float64)
    out = np.sinc(x) * np.sinc(x / radius)
    return np.where(np.abs(x) >= radius, 0.0, out)


def aperture_fft(vx, vy, n, pad, ab, k, radius=4):
    """I(A, B) = integral of exp(ik(Ax + By)) over the polygon, from the FFT
    of its raster (numpy FFT, FP64) sampled by phase-corrected Lanczos
    interpolation -- the restatement of the device path
    (tools/fft_aperture_probe.cu:50-102 for the sampling).  Returns
    (values [n_points] complex, spectrum [n, n] complex)."""
    grid, (xc, yc, dx, dy) = aperture_raster(vx, vy, n, pad)
    X = np.fft.fft2(grid)  # rows = y, columns = x; exp(-2 pi i j m / n)
    ab = np.asarray(ab, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(len(ab), dtype=complex)
    corr = np.pi * (n - 1) / n
    for p, (A, B) in enumerate(ab):
        fx, fy = -k * A / (2 * np.pi), -k * B / (2 * np.pi)
        ux, uy = fx * n * dx, fy * n * dy
        cx, cy = int(np.floor(ux)), int(np.floor(uy))
        jx = np.arange(cx - radius + 1, cx + radius + 1)
        jy = np.arange(cy - radius + 1, cy + radius + 1)
        wx, wy = _lanczos(ux - jx, radius), _lanczos(uy - jy, radius)
        w = np.outer(wy, wx)
        z = X[np.ix_(jy % n, jx % n)] * np.exp(1j * corr * (jy[:, None] + jx[None, :]))
        val = np.sum(w * z) / np.sum(w)
        out[p] = dx * dy * val * np.exp(-2j * np.pi * (fx * xc + fy * yc))
    return out, X
