"""ctypes mirror of include/mbs_hip.h and helpers that build its structs
from numpy arrays.

The arrays of a beam batch follow the fields of the reference's PreparedBeam
(src/handler/HandlerPO.h:14-46); ``BeamBatch`` keeps the numpy arrays alive for
as long as the ctypes struct that points into them.
"""
import ctypes as C

import numpy as np

MBS_OK = 0
MBS_ERR_INVALID = 1
MBS_ERR_BAD_BEAM = 2
MBS_ERR_DEVICE = 3
MBS_ERR_NO_GRID = 4
MBS_ERR_NO_BEAMS = 5
MBS_ERR_CAPACITY = 6
STATUS_NAMES = {0: "MBS_OK", 1: "MBS_ERR_INVALID", 2: "MBS_ERR_BAD_BEAM",
                3: "MBS_ERR_DEVICE", 4: "MBS_ERR_NO_GRID", 5: "MBS_ERR_NO_BEAMS",
                6: "MBS_ERR_CAPACITY"}

B_HOR, B_VER, B_DIR, B_CEN, B_AREA = 0, 3, 6, 9, 12
B_NT, B_NP, B_NXDT, B_NXDP, B_JP, B_NORMAL = 13, 16, 19, 22, 25, 33
BEAM_SCALARS = 36
WORK_AUTO, WORK_ORIENTATIONS, WORK_SEGMENTS, WORK_QUEUE = 0, 1, 2, 3
EDGE_VALID_X, EDGE_VALID_Y = 1, 2
MIN_VERTICES, MAX_VERTICES = 3, 31

PREC_FP64 = 0
PREC_MIXED = 1
NOSHADOW = 4
PER_ORIENT = 8
NO_MFMA = 16
INCOHERENT = 32
SPECTRUM = 64

ORIENT_SOBOL, ORIENT_HAMMERSLEY, ORIENT_SHOEMAKE = 0, 1, 2
ORIENT_STRIDE = 16
ORIENT_BETA, ORIENT_GAMMA, ORIENT_W, ORIENT_Q, ORIENT_R = 0, 1, 2, 3, 7

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class mbs_beam_batch(C.Structure):
    _fields_ = [("n_orient", C.c_int32), ("n_beams", C.c_int32),
                ("orient_offsets", _i64p), ("weights", _dp),
                ("beam_scalars", _dp), ("beam_nv", _i32p),
                ("beam_external", _i32p), ("vert_offsets", _i64p),
                ("vx", _dp), ("vy", _dp), ("slope_yx", _dp), ("slope_xy", _dp),
                ("edge_flags", _u8p)]


class mbs_beam_scaling(C.Structure):
    _fields_ = [("raw_jones", _dp), ("path", _dp), ("sign", _dp),
                ("abs_offsets", C.POINTER(C.c_int64)), ("abs_paths", _dp),
                ("c_abs", C.c_double), ("wave_index", C.c_double)]


class mbs_beam_staging(C.Structure):
    _fields_ = [("n_orient", C.c_int32), ("n_beams", C.c_int32), ("n_vert", C.c_int64),
                ("orient_offsets", C.POINTER(C.c_int64)), ("weights", C.POINTER(C.c_double)),
                ("beam_scalars", C.POINTER(C.c_double)), ("beam_nv", C.POINTER(C.c_int32)),
                ("beam_external", C.POINTER(C.c_int32)),
                ("vert_offsets", C.POINTER(C.c_int64)), ("vx", C.POINTER(C.c_double)),
                ("vy", C.POINTER(C.c_double)), ("slope_yx", C.POINTER(C.c_double)),
                ("slope_xy", C.POINTER(C.c_double)), ("edge_flags", C.POINTER(C.c_uint8))]


class mbs_grid(C.Structure):
    _fields_ = [("n_az", C.c_int32), ("n_zen", C.c_int32), ("theta", _dp),
                ("az_step", C.c_double), ("vf", _dp)]


class mbs_consts(C.Structure):
    _fields_ = [("wave_index", C.c_double), ("eps1", C.c_double),
                ("compl_wave", C.c_double * 2), ("inv_compl_wave", C.c_double * 2),
                ("legacy_sign", C.c_int32), ("reserved", C.c_int32)]


class mbs_device_info(C.Structure):
    _fields_ = [("device_id", C.c_int32), ("visible_devices", C.c_int32),
                ("total_mem", C.c_int64), ("free_mem", C.c_int64),
                ("compute_units", C.c_int32), ("clock_khz", C.c_int32),
                ("runtime_version", C.c_int32), ("driver_version", C.c_int32),
                ("name", C.c_char * 128), ("arch", C.c_char * 64)]


class mbs_timing(C.Structure):
    _fields_ = [("upload_ms", C.c_float), ("kernel_ms", C.c_float),
                ("reduce_ms", C.c_float), ("n_items", C.c_int32), ("n_blocks", C.c_int32)]


class mbs_spectrum_params(C.Structure):
    _fields_ = [("a_max", C.c_double), ("oversample", C.c_double), ("min_n", C.c_int32),
                ("max_n", C.c_int32)]


class mbs_spectrum_info(C.Structure):
    _fields_ = [("n_selected", C.c_int32), ("n_by_size", C.c_int32 * 6), ("reserved", C.c_int32),
                ("spectrum_bytes", C.c_int64), ("selected_area", C.c_double),
                ("total_area", C.c_double), ("scan_ms", C.c_float), ("fft_ms", C.c_float),
                ("sample_ms", C.c_float), ("reserved2", C.c_float)]


class mbs_tune_entry(C.Structure):
    _fields_ = [("target_blocks", C.c_int32), ("work_mode", C.c_int32), ("mfma", C.c_int32),
                ("ms", C.c_float), ("waves", C.c_int32)]


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class BeamBatch:
    """Prepared beams of one or more orientations as contiguous arrays.

    Fields mirror ``mbs_beam_batch``.  ``from_fixture`` reads the .npz files
    under tests/golden; ``subset`` selects orientations."""

    def __init__(self, orient_offsets, weights, beam_scalars, beam_nv,
                 beam_external, vert_offsets, vx, vy, slope_yx, slope_xy,
                 edge_flags):
        self.orient_offsets = _c(orient_offsets, np.int64)
        self.weights = _c(weights, np.float64)
        self.beam_scalars = _c(beam_scalars, np.float64).reshape(-1, BEAM_SCALARS)
        self.beam_nv = _c(beam_nv, np.int32)
        self.beam_external = _c(beam_external, np.int32)
        self.vert_offsets = _c(vert_offsets, np.int64)
        self.vx = _c(vx, np.float64)
        self.vy = _c(vy, np.float64)
        self.slope_yx = _c(slope_yx, np.float64)
        self.slope_xy = _c(slope_xy, np.float64)
        self.edge_flags = _c(edge_flags, np.uint8)
        n_b = len(self.beam_nv)
        if (self.beam_scalars.shape[0] != n_b or len(self.beam_external) != n_b
                or len(self.vert_offsets) != n_b + 1
                or self.orient_offsets[-1] != n_b
                or len(self.orient_offsets) != len(self.weights) + 1):
            raise ValueError("inconsistent beam batch arrays")
        self._s = None

    @property
    def n_orient(self):
        return len(self.weights)

    @property
    def n_beams(self):
        return len(self.beam_nv)

    @classmethod
    def from_fixture(cls, fx):
        return cls(fx["orient_offsets"], fx["weights"], fx["beam_scalars"],
                   fx["beam_nv"], fx["beam_external"], fx["vert_offsets"],
                   fx["vx"], fx["vy"], fx["slope_yx"], fx["slope_xy"],
                   fx["edge_flags"])

    def subset(self, orients):
        """New batch holding the given orientation indices, in that order."""
        oo, vo = self.orient_offsets, self.vert_offsets
        beams = np.concatenate([np.arange(oo[o], oo[o + 1]) for o in orients]) \
            if len(orients) else np.zeros(0, np.int64)
        counts = [oo[o + 1] - oo[o] for o in orients]
        nv = self.beam_nv[beams]
        verts = np.concatenate([np.arange(vo[b], vo[b + 1]) for b in beams]) \
            if len(beams) else np.zeros(0, np.int64)
        return BeamBatch(np.concatenate([[0], np.cumsum(counts)]),
                         self.weights[list(orients)], self.beam_scalars[beams], nv,
                         self.beam_external[beams], np.concatenate([[0], np.cumsum(nv)]),
                         self.vx[verts], self.vy[verts], self.slope_yx[verts],
                         self.slope_xy[verts], self.edge_flags[verts])

    def struct(self):
        if self._s is None:
            self._s = mbs_beam_batch(
                self.n_orient, self.n_beams, _ptr(self.orient_offsets, C.c_int64),
                _ptr(self.weights, C.c_double), _ptr(self.beam_scalars, C.c_double),
                _ptr(self.beam_nv, C.c_int32), _ptr(self.beam_external, C.c_int32),
                _ptr(self.vert_offsets, C.c_int64), _ptr(self.vx, C.c_double),
                _ptr(self.vy, C.c_double), _ptr(self.slope_yx, C.c_double),
                _ptr(self.slope_xy, C.c_double), _ptr(self.edge_flags, C.c_uint8))
        return self._s


class Grid:
    """Scattering grid (ScatteringRange, src/handler/Handler.h:15-189)."""

    def __init__(self, theta, az_step, vf):
        self.theta = _c(theta, np.float64)
        self.az_step = float(az_step)
        self.vf = _c(vf, np.float64)
        self.n_zen = len(self.theta) - 1
        self.n_az = self.vf.shape[0]
        if self.vf.shape != (self.n_az, self.n_zen + 1, 3):
            raise ValueError("vf must be [n_az, n_zen+1, 3]")
        self._s = None

    @property
    def cells(self):
        return self.n_az * (self.n_zen + 1)

    @classmethod
    def from_fixture(cls, fx):
        return cls(fx["theta"], fx["az_step"], fx["vf"])

    @classmethod
    def build(cls, theta, n_az, pol_basis=(0.0, 1.0, 0.0)):
        """Grid with the polarisation basis of ScatteringRange::
        ComputeSphereDirections (Handler.h:127-174): vf = (-sin phi, cos phi, 0)
        except at the poles (within 256*DBL_EPSILON), where vf = +pol (forward)
        or -pol (backward); pol is the incident polarisation basis (0,1,0),
        src/Tracer.cpp:126-127."""
        theta = np.asarray(theta, dtype=np.float64)
        step = 2.0 * np.pi / n_az
        phi = np.arange(n_az) * step
        vf = np.zeros((n_az, len(theta), 3))
        vf[:, :, 0] = -np.sin(phi)[:, None]
        vf[:, :, 1] = np.cos(phi)[:, None]
        tol = 256.0 * np.finfo(np.float64).eps
        pol = np.asarray(pol_basis, dtype=np.float64)
        bwd = np.abs(theta - np.pi) <= tol
        fwd = np.abs(theta) <= tol
        vf[:, bwd, :] = -pol
        vf[:, fwd & ~bwd, :] = pol
        return cls(theta, step, vf)

    def struct(self):
        if self._s is None:
            self._s = mbs_grid(self.n_az, self.n_zen, _ptr(self.theta, C.c_double),
                               self.az_step, _ptr(self.vf, C.c_double))
        return self._s


def consts_from_fixture(fx):
    c = mbs_consts()
    c.wave_index = float(fx["wave_index"])
    c.eps1 = float(fx["eps1"])
    c.compl_wave[0], c.compl_wave[1] = [float(v) for v in fx["compl_wave"]]
    c.inv_compl_wave[0], c.inv_compl_wave[1] = [float(v) for v in fx["inv_compl_wave"]]
    c.legacy_sign = int(fx["legacy_sign"])
    return c


def make_consts(wavelength, legacy_sign=False):
    """Constants of Handler::Handler (src/handler/Handler.cpp:102-126):
    k = 2 pi / lambda, complWave = -i lambda / (4 pi^2), invComplWave = i/lambda,
    eps1 = 1e9 * DBL_EPSILON."""
    c = mbs_consts()
    c.wave_index = 2.0 * np.pi / wavelength
    c.eps1 = 1e9 * np.finfo(np.float64).eps
    c.compl_wave[0], c.compl_wave[1] = 0.0, -wavelength / (2.0 * np.pi) ** 2
    c.inv_compl_wave[0], c.inv_compl_wave[1] = 0.0, 1.0 / wavelength
    c.legacy_sign = 1 if legacy_sign else 0
    return c
