"""ctypes binding of libmbs_hip.so (the gfx950 engine, include/mbs_hip.h).

There is no CPU fallback: if the library is missing, or no HIP device is
visible, every entry point raises.  The library is loaded from this package
directory (built in place by build.py), never from site-packages.
"""
import ctypes as C
import os
import sys

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbs_hip.so")
_LIB = None


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, status), message))
        self.status = status


def lib():
    """Loads libmbs_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(abi.MBS_ERR_DEVICE,
                              "libmbs_hip.so is not built (run __graft_entry__.build() "
                              "or python mbs-fast_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.mbs_abi_version.restype = C.c_int
        L.mbs_build_info.restype = C.c_char_p
        L.mbs_visible_device_count.restype = C.c_int
        L.mbs_check_runtime.argtypes = [C.POINTER(abi.mbs_device_info)]
        L.mbs_check_runtime.restype = C.c_int
        L.mbs_query_memory.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.mbs_query_memory.restype = C.c_int
        L.mbs_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mbs_ctx_create.restype = C.c_int
        L.mbs_ctx_destroy.argtypes = [vp]
        L.mbs_ctx_destroy.restype = C.c_int
        L.mbs_last_error.argtypes = [vp]
        L.mbs_last_error.restype = C.c_char_p
        L.mbs_set_grid.argtypes = [vp, C.POINTER(abi.mbs_grid)]
        L.mbs_set_grid.restype = C.c_int
        L.mbs_upload_beams.argtypes = [vp, C.POINTER(abi.mbs_beam_batch),
                                       C.POINTER(abi.mbs_consts)]
        L.mbs_upload_beams.restype = C.c_int
        L.mbs_stage_beams.argtypes = [vp, C.c_int32, C.c_int32, C.c_int64,
                                      C.POINTER(abi.mbs_beam_staging)]
        L.mbs_stage_beams.restype = C.c_int
        L.mbs_upload_staged.argtypes = [vp, C.POINTER(abi.mbs_consts)]
        L.mbs_upload_staged.restype = C.c_int
        L.mbs_upload_beam_scaling.argtypes = [vp, C.POINTER(abi.mbs_beam_scaling)]
        L.mbs_upload_beam_scaling.restype = C.c_int
        L.mbs_rescale_beams.argtypes = [vp, C.c_double]
        L.mbs_rescale_beams.restype = C.c_int
        L.mbs_diffract.argtypes = [vp, C.c_uint32]
        L.mbs_diffract.restype = C.c_int
        L.mbs_diffract_noalloc.argtypes = [vp, C.c_uint32, vp, vp]
        L.mbs_diffract_noalloc.restype = C.c_int
        L.mbs_accumulate_mueller.argtypes = [vp, dp, dp]
        L.mbs_accumulate_mueller.restype = C.c_int
        L.mbs_fetch_per_orient.argtypes = [vp, dp]
        L.mbs_fetch_per_orient.restype = C.c_int
        L.mbs_diffract_batch.argtypes = [vp, C.POINTER(abi.mbs_grid),
                                         C.POINTER(abi.mbs_beam_batch),
                                         C.POINTER(abi.mbs_consts), C.c_uint32, dp, dp]
        L.mbs_diffract_batch.restype = C.c_int
        L.mbs_orientation_set.argtypes = [vp, C.c_int, C.c_int32, C.c_uint32, C.c_double,
                                          C.c_double, dp]
        L.mbs_orientation_set.restype = C.c_int
        L.mbs_aperture_fft.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32, C.c_int32,
                                       C.c_double, C.c_int32, dp, C.c_int32, dp]
        L.mbs_aperture_fft.restype = C.c_int
        L.mbs_spectrum_source.argtypes = [vp, C.POINTER(abi.mbs_spectrum_params),
                                          C.POINTER(abi.mbs_spectrum_info)]
        L.mbs_spectrum_source.restype = C.c_int
        L.mbs_spectrum_values.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.mbs_spectrum_values.restype = C.c_int
        L.mbs_aperture_fft_spectrum.argtypes = [vp, C.c_int32, C.POINTER(C.c_float)]
        L.mbs_aperture_fft_spectrum.restype = C.c_int
        L.mbs_aperture_fft_timing.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                              C.POINTER(C.c_float)]
        L.mbs_aperture_fft_timing.restype = C.c_int
        L.mbs_autotune.argtypes = [vp, C.c_uint32, C.POINTER(abi.mbs_tune_entry), C.c_int32,
                                   C.POINTER(C.c_int32)]
        L.mbs_autotune.restype = C.c_int
        L.mbs_graph_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.mbs_graph_stats.restype = C.c_int
        L.mbs_apply_tune_entry.argtypes = [vp, C.c_uint32, C.POINTER(abi.mbs_tune_entry)]
        L.mbs_apply_tune_entry.restype = C.c_int
        L.mbs_set_graph.argtypes = [vp, C.c_int]
        L.mbs_set_graph.restype = C.c_int
        fp = C.POINTER(C.c_float)
        L.mbs_mlp_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_uint64, C.POINTER(vp)]
        L.mbs_mlp_create.restype = C.c_int
        L.mbs_mlp_destroy.argtypes = [vp]
        L.mbs_mlp_destroy.restype = C.c_int
        L.mbs_mlp_param_count.argtypes = [vp]
        L.mbs_mlp_param_count.restype = C.c_int64
        L.mbs_mlp_set_params.argtypes = [vp, fp]
        L.mbs_mlp_set_params.restype = C.c_int
        L.mbs_mlp_get_params.argtypes = [vp, fp]
        L.mbs_mlp_get_params.restype = C.c_int
        L.mbs_mlp_forward.argtypes = [vp, fp, C.c_int32, fp]
        L.mbs_mlp_forward.restype = C.c_int
        L.mbs_mlp_gradient.argtypes = [vp, fp, fp, C.c_int32, dp, fp]
        L.mbs_mlp_gradient.restype = C.c_int
        L.mbs_mlp_train.argtypes = [vp, fp, fp, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                    C.c_uint64, dp]
        L.mbs_mlp_train.restype = C.c_int
        L.mbs_azimuth_average_per_orient.argtypes = [vp, dp]
        L.mbs_azimuth_average_per_orient.restype = C.c_int
        L.mbs_upsample_result.argtypes = [vp, C.c_int, C.c_int32, C.c_int32, dp, C.c_int, dp, dp]
        L.mbs_upsample_result.restype = C.c_int
        L.mbs_azimuth_average.argtypes = [vp, C.c_int, dp]
        L.mbs_azimuth_average.restype = C.c_int
        L.mbs_device_result.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
        L.mbs_device_result.restype = C.c_int
        L.mbs_ctx_stream.argtypes = [vp]
        L.mbs_ctx_stream.restype = vp
        L.mbs_batch_stats.argtypes = [vp, C.POINTER(C.c_int64), dp, C.POINTER(C.c_int64), dp]
        L.mbs_batch_stats.restype = C.c_int
        L.mbs_set_target_blocks.argtypes = [vp, C.c_int]
        L.mbs_set_target_blocks.restype = C.c_int
        L.mbs_set_work_mode.argtypes = [vp, C.c_int]
        L.mbs_set_work_mode.restype = C.c_int
        L.mbs_set_timing.argtypes = [vp, C.c_int]
        L.mbs_set_timing.restype = C.c_int
        L.mbs_get_timing.argtypes = [vp, C.POINTER(abi.mbs_timing)]
        L.mbs_get_timing.restype = C.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def device_info():
    info = abi.mbs_device_info()
    rc = lib().mbs_check_runtime(C.byref(info))
    if rc:
        raise EngineError(rc, lib().mbs_last_error(None).decode())
    return {"device_id": info.device_id, "visible": info.visible_devices,
            "name": info.name.decode(), "arch": info.arch.decode(),
            "cus": info.compute_units, "clock_khz": info.clock_khz,
            "total_mem": info.total_mem, "free_mem": info.free_mem}


class Context:
    """One engine context (one HIP device, one stream)."""

    def __init__(self, device=0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.mbs_ctx_create(device, C.byref(h))
        if rc:
            raise EngineError(rc, self._L.mbs_last_error(None).decode())
        self._h = h
        self.grid = None
        self.batch = None

    def _check(self, rc):
        if rc:
            raise EngineError(rc, self._L.mbs_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.mbs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        # No HIP calls while the interpreter shuts down: the runtime may
        # already be torn down; the process exit releases the device.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    @property
    def stream(self):
        return self._L.mbs_ctx_stream(self._h)

    def set_grid(self, grid):
        self._check(self._L.mbs_set_grid(self._h, C.byref(grid.struct())))
        self.grid = grid

    def upload(self, batch, consts):
        self._check(self._L.mbs_upload_beams(self._h, C.byref(batch.struct()), C.byref(consts)))
        self.batch = batch

    def upload_staged(self, batch, consts):
        """The batch through the context's page-locked staging
        (mbs_stage_beams, filled here, then mbs_upload_staged): the copies
        are asynchronous and the call returns once they are queued."""
        st = abi.mbs_beam_staging()
        nv = int(batch.vert_offsets[-1])
        self._check(self._L.mbs_stage_beams(self._h, batch.n_orient, batch.n_beams, nv,
                                            C.byref(st)))
        for name, a, n in (("orient_offsets", batch.orient_offsets, batch.n_orient + 1),
                           ("weights", batch.weights, batch.n_orient),
                           ("beam_scalars", batch.beam_scalars.reshape(-1),
                            batch.n_beams * abi.BEAM_SCALARS),
                           ("beam_nv", batch.beam_nv, batch.n_beams),
                           ("beam_external", batch.beam_external, batch.n_beams),
                           ("vert_offsets", batch.vert_offsets, batch.n_beams + 1),
                           ("vx", batch.vx, nv), ("vy", batch.vy, nv),
                           ("slope_yx", batch.slope_yx, nv), ("slope_xy", batch.slope_xy, nv),
                           ("edge_flags", batch.edge_flags, nv)):
            if n:
                dst = np.ctypeslib.as_array(getattr(st, name), shape=(n,))
                dst[:] = a
        self._check(self._L.mbs_upload_staged(self._h, C.byref(consts)))
        self.batch = batch

    def upload_scaling(self, raw_jones, path, sign, abs_offsets, abs_paths, c_abs, wave_index):
        """Per-beam data of the uploaded batch for mbs_rescale_beams."""
        self._sc_arrays = [np.ascontiguousarray(raw_jones, np.float64).reshape(-1),
                           np.ascontiguousarray(path, np.float64),
                           np.ascontiguousarray(sign, np.float64),
                           np.ascontiguousarray(abs_offsets, np.int64),
                           np.ascontiguousarray(abs_paths, np.float64)]
        r, p, sg, ao, ap = self._sc_arrays
        sc = abi.mbs_beam_scaling(_d(r), _d(p), _d(sg), ao.ctypes.data_as(C.POINTER(C.c_int64)),
                                  _d(ap) if ap.size else None, float(c_abs), float(wave_index))
        self._check(self._L.mbs_upload_beam_scaling(self._h, C.byref(sc)))

    def rescale(self, scale):
        self._check(self._L.mbs_rescale_beams(self._h, float(scale)))

    def diffract(self, flags=0):
        self._check(self._L.mbs_diffract(self._h, flags))

    def diffract_noalloc(self, flags, d_m_full, d_m_noshadow=None):
        """mbs_diffract_noalloc: writes the Mueller sum of the uploaded batch
        into caller-owned device memory (device addresses as ints)."""
        self._check(self._L.mbs_diffract_noalloc(self._h, flags, C.c_void_p(d_m_full),
                                                 C.c_void_p(d_m_noshadow)
                                                 if d_m_noshadow else None))

    def accumulate(self, m_full, m_noshadow=None):
        for a in (m_full, m_noshadow):
            if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous
                                  or a.size != self.grid.cells * 16):
                raise ValueError("Mueller arrays must be C-contiguous float64 [cells*16]")
        self._check(self._L.mbs_accumulate_mueller(
            self._h, _d(m_full), _d(m_noshadow) if m_noshadow is not None else None))

    def mueller(self, flags=0, noshadow=False):
        """Runs the uploaded batch; returns (M, M_noshadow) as [n_az, n_zen+1, 16]."""
        if noshadow:
            flags |= abi.NOSHADOW
        self.diffract(flags)
        g = self.grid
        m = np.zeros((g.n_az, g.n_zen + 1, 16))
        mn = np.zeros_like(m) if noshadow else None
        self.accumulate(m, mn)
        return m, mn

    def upsample(self, n_az_out, n_zen_out, noshadow=False, theta_out=None, rows=False):
        """mbs_upsample_result: the last result Fourier-upsampled on the
        device to n_az_out x (n_zen_out + 1) cells.  Returns (grid
        [n_az_out, n_zen_out+1, 16], rows [n_zen_out+1, 18] or None)."""
        grid = np.zeros((int(n_az_out), int(n_zen_out) + 1, 16))
        r = np.zeros((int(n_zen_out) + 1, 18)) if rows else None
        th = np.ascontiguousarray(theta_out, np.float64) if theta_out is not None else None
        self._check(self._L.mbs_upsample_result(
            self._h, 1 if noshadow else 0, int(n_az_out), int(n_zen_out),
            _d(th) if th is not None else None, 0, _d(grid), _d(r) if r is not None else None))
        return grid, r

    def azimuth_average(self, noshadow=False):
        """Rows [n_zen+1, 18] (theta deg, 2 pi dcos, M11..M44) of the last
        result, averaged over azimuth on the device."""
        out = np.zeros((self.grid.n_zen + 1, 18))
        self._check(self._L.mbs_azimuth_average(self._h, 1 if noshadow else 0, _d(out)))
        return out

    def orientation_set(self, kind, n, seed=1, beta_sym=np.pi, gamma_sym=2 * np.pi):
        """[n, 16] rows (beta, gamma, weight, qx, qy, qz, qw, R row-major),
        generated on the device; kind is abi.ORIENT_SOBOL, ORIENT_HAMMERSLEY
        or ORIENT_SHOEMAKE."""
        out = np.zeros((int(n), abi.ORIENT_STRIDE))
        self._check(self._L.mbs_orientation_set(self._h, int(kind), int(n), int(seed),
                                                float(beta_sym), float(gamma_sym), _d(out)))
        return out

    def aperture_fft(self, beams, n, pad=4.0, ab=None, radius=4):
        """I(A, B) of each listed beam at the (A, B) points, [n_beams, n_points]
        complex, from the device aperture FFT (see mbs_aperture_fft)."""
        beams = np.ascontiguousarray(beams, dtype=np.int64)
        pts = np.zeros((0, 2)) if ab is None else np.ascontiguousarray(ab, dtype=np.float64)
        out = np.zeros((len(beams), len(pts), 2))
        self._check(self._L.mbs_aperture_fft(
            self._h, beams.ctypes.data_as(C.POINTER(C.c_int64)), len(beams), int(n),
            float(pad), len(pts), _d(pts), int(radius), _d(out)))
        return out[..., 0] + 1j * out[..., 1]

    def spectrum_source(self, a_max=0.1, oversample=8.0, min_n=64, max_n=2048):
        """mbs_spectrum_source on the uploaded batch and the current grid: the
        near-forward beams' F from their spectra for mbs_diffract(... |
        abi.SPECTRUM).  Returns the selection summary as a dict."""
        p = abi.mbs_spectrum_params(float(a_max), float(oversample), int(min_n), int(max_n))
        info = abi.mbs_spectrum_info()
        self._check(self._L.mbs_spectrum_source(self._h, C.byref(p), C.byref(info)))
        out = {f: getattr(info, f) for f, _ in abi.mbs_spectrum_info._fields_
               if not f.startswith("reserved")}
        out["n_by_size"] = {64 << i: int(v) for i, v in enumerate(info.n_by_size)}
        return out

    def spectrum_clear(self):
        self._check(self._L.mbs_spectrum_source(self._h, None, None))

    def spectrum_values(self, n_selected):
        """(beams [n_selected], F [n_selected, cells] complex64) of the
        current selection, F in the kernels' convention s k^2 I e^{i psi}."""
        beams = np.zeros(int(n_selected), np.int32)
        f = np.zeros((int(n_selected), self.grid.cells, 2), np.float32)
        self._check(self._L.mbs_spectrum_values(
            self._h, beams.ctypes.data_as(C.POINTER(C.c_int32)),
            f.ctypes.data_as(C.POINTER(C.c_float))))
        return beams, f[..., 0] + 1j * f[..., 1]

    def aperture_fft_spectrum(self, index, n):
        out = np.zeros((n, n, 2), dtype=np.float32)
        self._check(self._L.mbs_aperture_fft_spectrum(
            self._h, int(index), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out[..., 0] + 1j * out[..., 1].astype(np.complex64)

    def aperture_fft_timing(self):
        r, c, s = C.c_float(), C.c_float(), C.c_float()
        self._check(self._L.mbs_aperture_fft_timing(self._h, C.byref(r), C.byref(c), C.byref(s)))
        return {"rows_ms": r.value, "cols_ms": c.value, "sample_ms": s.value}

    def azimuth_average_per_orient(self):
        """Rows [n_orient, n_zen+1, 18] of the last MBS_PER_ORIENT result."""
        out = np.zeros((self.batch.n_orient, self.grid.n_zen + 1, 18))
        self._check(self._L.mbs_azimuth_average_per_orient(self._h, _d(out)))
        return out

    def per_orient(self):
        g = self.grid
        out = np.zeros((self.batch.n_orient, g.n_az, g.n_zen + 1, 16))
        self._check(self._L.mbs_fetch_per_orient(self._h, _d(out)))
        return out

    def device_result(self):
        a, b = C.c_void_p(), C.c_void_p()
        self._check(self._L.mbs_device_result(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stats(self):
        nb, units, mean_e = C.c_int64(), C.c_double(), C.c_double()
        hist = np.zeros(32, np.int64)
        self._check(self._L.mbs_batch_stats(self._h, C.byref(nb), C.byref(units),
                                            hist.ctypes.data_as(C.POINTER(C.c_int64)),
                                            C.byref(mean_e)))
        return {"n_beams": nb.value, "units": units.value, "nv_hist": hist,
                "mean_edges": mean_e.value}

    def set_target_blocks(self, n):
        self._check(self._L.mbs_set_target_blocks(self._h, int(n)))

    def set_work_mode(self, mode):
        """0 automatic, 1 whole orientations per item, 2 beam segments, 3 beam
        segments planned on the device and run from a device work queue."""
        self._check(self._L.mbs_set_work_mode(self._h, int(mode)))

    def autotune(self, flags=0):
        """Times candidate launch shapes, keeps the fastest; returns the list
        of (target_blocks, work_mode, mfma, ms, waves)."""
        buf = (abi.mbs_tune_entry * 128)()
        n = C.c_int32()
        self._check(self._L.mbs_autotune(self._h, flags, buf, 128, C.byref(n)))
        return [(e.target_blocks, e.work_mode, e.mfma, e.ms, e.waves) for e in buf[:n.value]]

    def apply_tune_entry(self, flags, target_blocks, work_mode, mfma, waves=0):
        """mbs_apply_tune_entry: use this launch shape for later calls."""
        e = abi.mbs_tune_entry(int(target_blocks), int(work_mode), int(mfma), 0.0, int(waves))
        self._check(self._L.mbs_apply_tune_entry(self._h, flags, C.byref(e)))

    def set_graph(self, enable=True):
        self._check(self._L.mbs_set_graph(self._h, 1 if enable else 0))

    def graph_stats(self):
        """(captures, launches) of graph replay on this context."""
        c, l = C.c_int64(), C.c_int64()
        self._check(self._L.mbs_graph_stats(self._h, C.byref(c), C.byref(l)))
        return c.value, l.value

    def set_timing(self, enable=True):
        self._check(self._L.mbs_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        t = abi.mbs_timing()
        self._check(self._L.mbs_get_timing(self._h, C.byref(t)))
        return {"upload_ms": t.upload_ms, "kernel_ms": t.kernel_ms, "reduce_ms": t.reduce_ms,
                "n_items": t.n_items, "n_blocks": t.n_blocks}


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """Device MLP of the learned control variate (mbs_mlp_*): n_in -> depth x
    width GELU -> n_out, trained with the library's BF16-MFMA kernels."""

    def __init__(self, ctx, n_in, width, depth, n_out, seed=0):
        self.ctx = ctx
        self.shape = (n_in, width, depth, n_out)
        h = C.c_void_p()
        ctx._check(ctx._L.mbs_mlp_create(ctx._h, n_in, width, depth, n_out, seed, C.byref(h)))
        self._h = h
        self.n_params = ctx._L.mbs_mlp_param_count(h)

    def close(self):
        if self._h:
            self.ctx._L.mbs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self.ctx._check(self.ctx._L.mbs_mlp_get_params(self._h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.ctx._check(self.ctx._L.mbs_mlp_set_params(self._h, _f(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.shape[3]), dtype=np.float32)
        self.ctx._check(self.ctx._L.mbs_mlp_forward(self._h, _f(x), x.shape[0], _f(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        self.ctx._check(self.ctx._L.mbs_mlp_gradient(self._h, _f(x), _f(t), x.shape[0],
                                                     C.byref(loss), _f(g)))
        return loss.value, g

    def train(self, x, t, batch=256, epochs=10, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double()
        self.ctx._check(self.ctx._L.mbs_mlp_train(self._h, _f(x), _f(t), x.shape[0],
                                                  int(batch), int(epochs), float(lr), int(seed),
                                                  C.byref(loss)))
        return loss.value
