"""Host-side mirror of the reference's HandlerPO GPU interface.

The reference exposes its GPU diffraction backend as member functions of
HandlerPO (src/handler/HandlerPO.h:98-146), implemented in
src/cuda/GpuDiffraction.cu and stubbed to ``false`` in CPU builds
(src/handler/HandlerPO.cpp:2798-2883).  ``HandlerPOGpu`` offers the same
members, argument meanings and error behaviour on top of the HIP engine:

* results are ACCUMULATED (``+=``) into caller-owned grids, cell (p, t) for
  p < nAz, t <= nZen (GpuDiffraction.cu:4751-4766);
* the no-shadow grid is written only when ``ComputeNoShadow()`` is true
  (``!m_fullOnly``, HandlerPO.cpp:982-985);
* a batch the backend cannot take (incoherent mode, a beam with fewer than 3
  or more than 31 aperture vertices) returns ``False`` and the caller raises,
  as TracerPOTotal.cpp:5934-5939 does;
* a device error prints a message and exits with status 2 unless
  ``MBS_GPU_ALLOW_FALLBACK=1``, in which case the call returns ``False``
  (GpuDiffraction.cu:694-710).

``prepared`` is a :class:`abi.BeamBatch` (the flat form of
``std::vector<PreparedOrientation>``); ``start``/``count`` select
orientations as in the reference.
"""
import os
import sys

import numpy as np

from . import abi
from .engine import Context, EngineError


class HandlerPOGpu:
    def __init__(self, grid, consts, device=0, precision="fp64"):
        self.grid = grid
        self.consts = consts
        self.isCoh = True
        self.m_fullOnly = True
        self.precision = precision
        self._ctx = None
        self._device = device

    # -- HandlerPO settings mirrored from the reference -------------------
    def SetFullOnly(self, value):
        self.m_fullOnly = bool(value)

    def IsFullOnly(self):
        return self.m_fullOnly

    def ComputeNoShadow(self):
        return not self.m_fullOnly

    # -- internals -----------------------------------------------------------
    def _context(self):
        if self._ctx is None:
            self._ctx = Context(self._device)
            self._ctx.set_grid(self.grid)
        return self._ctx

    def _flags(self):
        f = abi.PREC_MIXED if self.precision == "mixed" else abi.PREC_FP64
        if self.ComputeNoShadow():
            f |= abi.NOSHADOW
        return f

    def _device_failure(self, err):
        sys.stderr.write("GPU diffraction error: %s\n" % err)
        if os.environ.get("MBS_GPU_ALLOW_FALLBACK") == "1":
            return False
        sys.exit(2)

    @staticmethod
    def _cells_view(m, n_az, n_zen1):
        """Accepts Arr2D-shaped [nAz+1, nZen+1, 16] or [nAz, nZen+1, 16]
        float64 arrays; returns the [nAz, nZen+1, 16] part."""
        if m.dtype != np.float64 or m.ndim != 3 or m.shape[1:] != (n_zen1, 16):
            raise ValueError("Mueller grid must be float64 [nAz(+1), nZen+1, 16]")
        if m.shape[0] not in (n_az, n_az + 1):
            raise ValueError("Mueller grid has %d phi rows, expected %d or %d"
                             % (m.shape[0], n_az, n_az + 1))
        return m[:n_az]

    # -- HandlerPO GPU members -------------------------------------------------
    def HandleOrientationsToLocalGpu(self, prepared, localM, localM_noshadow, start=0,
                                     count=None, scale=1.0, waveIndex=0.0):
        """HandlerPO::HandleOrientationsToLocalGpu (HandlerPO.h:109-120)."""
        if not self.isCoh:
            return False
        if count is None:
            count = prepared.n_orient - start
        if start < 0 or count < 0 or start > prepared.n_orient:
            return False
        if scale != 1.0 or waveIndex != 0.0:
            return False  # size-rescaled batches are the multi-size path
        n = min(count, prepared.n_orient - start)
        g = self.grid
        if n == 0 or g.cells == 0:
            return True
        batch = prepared if (start == 0 and n == prepared.n_orient) else \
            prepared.subset(list(range(start, start + n)))
        mv = self._cells_view(localM, g.n_az, g.n_zen + 1)
        mnv = self._cells_view(localM_noshadow, g.n_az, g.n_zen + 1) \
            if self.ComputeNoShadow() else None
        try:
            ctx = self._context()
            ctx.upload(batch, self.consts)
            ctx.diffract(self._flags())
            tmp = np.zeros((g.n_az, g.n_zen + 1, 16))
            tmpn = np.zeros_like(tmp) if mnv is not None else None
            ctx.accumulate(tmp, tmpn)
        except EngineError as e:
            if e.status in (abi.MBS_ERR_BAD_BEAM, abi.MBS_ERR_INVALID):
                return False
            return self._device_failure(e)
        mv += tmp
        if mnv is not None:
            mnv += tmpn
        return True

    def HandleOrientationsToLocalGpuOnDevice(self, prepared, start, count, device, localM,
                                             localM_noshadow, scale=1.0, waveIndex=0.0,
                                             cacheToken=0):
        """HandlerPO::HandleOrientationsToLocalGpuOnDevice (HandlerPO.h:121-130).
        One GPU per process here: any device index other than the handler's is
        refused."""
        if device != self._device:
            return False
        return self.HandleOrientationsToLocalGpu(prepared, localM, localM_noshadow, start,
                                                 count, scale, waveIndex)

    def HandleBeamsToLocalGpu(self, prepared_orientation, localM, localM_noshadow):
        """HandlerPO::HandleBeamsToLocalGpu (HandlerPO.h:103-105): one orientation."""
        return self.HandleOrientationsToLocalGpu(prepared_orientation, localM,
                                                 localM_noshadow, 0, 1)

    def SelectGpuOrientationBatchSize(self, prepared, start, maxCount):
        """HandlerPO::SelectGpuOrientationBatchSize (HandlerPO.h:145-147): the
        largest count from ``start`` whose device working set fits in 80 % of
        free device memory (partials + packed beams), at least 1."""
        from .engine import lib
        import ctypes as C
        fr, tot = C.c_int64(), C.c_int64()
        if lib().mbs_query_memory(C.byref(fr), C.byref(tot)) != 0:
            return max(1, maxCount)
        budget = 0.8 * fr.value
        cells = self.grid.cells
        n = max(1, maxCount)
        oo = prepared.orient_offsets
        while n > 1:
            end = min(prepared.n_orient, start + n)
            beams = int(oo[end] - oo[start])
            need = beams * (36 + 40 + 14 + 6 * 31) * 8 + 64 * cells * 16 * 8 * 2
            if need <= budget:
                break
            n //= 2
        return n

    def HandleOrientationsToLocalGpuFftPhi(self, prepared, localM, localM_noshadow, start=0,
                                           count=None, scale=1.0, waveIndex=0.0,
                                           fftPhiFactor=0):
        """HandlerPO::HandleOrientationsToLocalGpuFftPhi (GpuDiffraction.cu:3504-3839):
        the batch on nLow = nFull / factor azimuths (at least 16), then the
        Mueller grids Fourier-upsampled on the device (mbs_upsample_result)
        and added into the handler's grids.  factor: ``fftPhiFactor``, else MBS_FFT_PHI_FACTOR,
        else by nFull (choose_fft_phi_factor, :712-727); a factor of 1,
        nFull < 32 or nLow >= nFull runs the direct path.  With
        MBS_FFT_THETA_FACTOR the coarse grid also has nZen/factor rows,
        upsampled in zenith first.  The validation and adaptive options return
        False."""
        if not self.isCoh:
            return False
        g = self.grid
        n_full = g.n_az
        factor = fftPhiFactor
        if factor <= 0:
            env = os.environ.get("MBS_FFT_PHI_FACTOR", "")
            factor = int(env) if env.isdigit() and 1 <= int(env) <= 64 else 0
        if factor <= 0:
            factor = 8 if n_full >= 2400 else 6 if n_full >= 600 else 3 if n_full >= 240 else 2
        n_low = n_full // factor if factor > 1 else n_full
        if n_low < 16:
            n_low = min(n_full, 16)
        if factor <= 1 or n_full < 32 or n_low >= n_full:
            return self.HandleOrientationsToLocalGpu(prepared, localM, localM_noshadow, start,
                                                     count, scale, waveIndex)
        if (os.environ.get("MBS_FFT_CHECK") == "1"
                or os.environ.get("MBS_FFT_ADAPTIVE_PHI", "0") not in ("", "0")):
            return False
        # theta factor (gpu_fft_theta_factor, :532-542; :3543-3550): a uniform
        # zenith grid of at least 32 rows that the factor divides
        env = os.environ.get("MBS_FFT_THETA_FACTOR", "")
        tf = int(env) if env.isdigit() and 1 <= int(env) <= 16 else 1
        steps = np.diff(g.theta)
        if tf > 1 and (g.n_zen < 32 or g.n_zen % tf or not np.allclose(steps, steps[0])):
            tf = 1
        n_zen_low = max(2, g.n_zen // tf) if tf > 1 else g.n_zen
        theta_low = np.linspace(g.theta[0], g.theta[-1], n_zen_low + 1) if tf > 1 else g.theta
        low = abi.Grid.build(theta_low, n_low)
        ns = self.ComputeNoShadow()
        count = prepared.n_orient - start if count is None else count
        if start < 0 or count < 0 or start > prepared.n_orient:
            return False
        orients = list(range(start, min(prepared.n_orient, start + count)))
        if not orients:
            return True
        if scale != 1.0 or waveIndex != 0.0:
            return False  # size-rescaled batches are the multi-size path
        batch = prepared.subset(orients)
        ctx = self._context()
        try:
            ctx.set_grid(low)
            ctx.upload(batch, self.consts)
            ctx.diffract(self._flags())
            # the device upsampler (mbs_upsample_result): zenith, then azimuth
            up, _ = ctx.upsample(n_full, g.n_zen)
            up_ns = ctx.upsample(n_full, g.n_zen, noshadow=True)[0] if ns else None
        except EngineError as e:
            return self._device_failure(e)
        finally:
            ctx.set_grid(g)
        self._cells_view(localM, g.n_az, g.n_zen + 1)[:] += up
        if ns:
            self._cells_view(localM_noshadow, g.n_az, g.n_zen + 1)[:] += up_ns
        return True

    def DiffractThetasGpu(self, prepared, theta_rads, nPoints=None):
        """HandlerPO::DiffractThetasGpu (HandlerPO.h:159-162), the reference's
        theta probe (GpuDiffraction.cu:4780-4965).  Returns (ok, m11_out).

        For every theta: the M11 of each single beam (no interference between
        beams), weighted by its orientation's weight, summed over the beams
        and the handler's nAz azimuths p * azimuth step, divided by nAz + 1
        (theta_m11_kernel, :1909-1957).  Theta within 1e-14 of a pole is moved
        1e-6 rad inside (:4845-4850); vf = (-sin phi, cos phi, 0) except at
        the poles (:4862-4876).  The reference's GPU rotation leaves out the
        1/2 of the CPU one and its Mueller kernels undo that with weight/4,
        but the probe sums 0.5 |d|^2 unscaled: its values are 4 times the M11
        of the CPU convention, and are reproduced so.  m11_out is untouched
        when there is nothing to do (no points, orientations or azimuths) and
        zero-filled otherwise; a batch the backend cannot take returns False."""
        theta = np.asarray(theta_rads, dtype=np.float64)
        n = len(theta) if nPoints is None else int(nPoints)
        if n <= 0:
            return True, None
        g = self.grid
        if prepared.n_orient == 0 or g.n_az <= 0:
            return True, None
        if np.any((prepared.beam_nv <= 0) | (prepared.beam_nv > 32)):
            return False, None  # as :4800-4802; the engine refuses nv < 3 or > 31 below
        m11 = np.zeros(n)
        if prepared.n_beams == 0:
            return True, m11
        th = theta[:n].copy()
        th[th <= 1e-14] = 1e-6
        th[th >= np.pi - 1e-14] = np.pi - 1e-6
        phi = np.arange(g.n_az) * g.az_step
        dz = -np.cos(th)
        vf = np.zeros((g.n_az, n, 3))
        vf[:, :, 0] = -np.sin(phi)[:, None]
        vf[:, :, 1] = np.cos(phi)[:, None]
        vf[:, dz >= 1.0 - 1e-15] = (0.0, -1.0, 0.0)
        vf[:, dz <= -1.0 + 1e-15] = (0.0, 1.0, 0.0)
        probe = abi.Grid(th, g.az_step, vf)
        try:
            ctx = self._context()
            ctx.set_grid(probe)
            try:
                ctx.upload(prepared, self.consts)
                ctx.diffract(abi.INCOHERENT)
                tmp = np.zeros((g.n_az, n, 16))
                ctx.accumulate(tmp, None)
            finally:
                ctx.set_grid(g)
        except EngineError as e:
            if e.status in (abi.MBS_ERR_BAD_BEAM, abi.MBS_ERR_INVALID):
                return False, None
            return self._device_failure(e), None
        m11[:] = 4.0 * tmp[:, :, 0].sum(axis=0) / (g.n_az + 1)
        return True, m11

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None
