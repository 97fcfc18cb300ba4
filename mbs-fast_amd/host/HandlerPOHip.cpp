// MI355X backend of MBS-fast's PO handler: the GPU members of HandlerPO and
// the GpuSupport queries, implemented over the engine's C ABI
// (include/mbs_hip.h).
//
// This translation unit takes the place of the reference's
// src/cuda/GpuDiffraction.cu and src/cuda/GpuSupport.cpp in a build of the
// reference whose HandlerPO.cpp and main.cpp are compiled with -DUSE_CUDA
// (that define only removes the CPU stubs of these members,
// src/handler/HandlerPO.cpp:2798-2883, and enables the GPU CLI options).  It
// is compiled against the reference's own headers and links
// libmbs_hip.so; it contains no HIP code itself.
//
// Interfaces mirrored (signatures, argument meaning, error behaviour):
//   src/handler/HandlerPO.h:98-146    HandlerPO GPU members
//   src/cuda/GpuSupport.h:30-34       CheckGpuRuntime, QueryActiveGpuMemory,
//                                     VisibleGpuDeviceCount, FormatGpuInfo
// Semantics kept from the reference driver
// (src/cuda/GpuDiffraction.cu:4159-4778):
//   * results are ADDED into the caller's localM / localM_noshadow cells
//     (p < nAzimuth, t <= nZenith); localM_noshadow is touched only when
//     ComputeNoShadow();
//   * incoherent handlers (!isCoh) and beams with invalid edge data or more
//     than 31 vertices are refused (return false, the caller throws);
//   * a device error prints a message and exits with status 2 unless
//     MBS_GPU_ALLOW_FALLBACK=1, in which case the call returns false;
//   * one engine context (device buffers, stream, grid tables) per host
//     thread and device, so concurrent callers on different threads and
//     devices are independent;
//   * MBS_GPU_TIMING=1 prints a per-call timing line to stderr.
// Precision: FP64 by default; built with -DMBS_GPU_FP32 (the reference's own
// FP32-storage switch, src/cuda/GpuDiffraction.cu:15-41) it runs the mixed
// kernel (FP32 storage, FP64 phases and moments).  MBS_HIP_PRECISION=fp64 or
// =fp32 overrides the build default at run time.
// Not provided (return false, as in a reference build without these paths):
// the phi-FFT, multi-size and theta-probe variants.
#include "HandlerPO.h"
#include "GpuSupport.h"

#include "mbs_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace
{
bool EnvIs(const char *name, const char *value)
{
    const char *v = std::getenv(name);
    return v && std::strcmp(v, value) == 0;
}

uint32_t PrecisionFlag()
{
    if (EnvIs("MBS_HIP_PRECISION", "fp64"))
        return MBS_PREC_FP64;
    if (EnvIs("MBS_HIP_PRECISION", "fp32"))
        return MBS_PREC_MIXED;
#if defined(MBS_GPU_FP32)
    return MBS_PREC_MIXED;
#else
    return MBS_PREC_FP64;
#endif
}

// A device failure follows the reference's convention
// (src/cuda/GpuDiffraction.cu:694-710): abort the run with status 2, unless
// the user allowed a fallback, in which case the call reports failure.
bool DeviceFailure(const char *what, const char *message)
{
    std::fprintf(stderr, "ERROR: HIP diffraction backend: %s failed: %s\n", what,
                 message ? message : "(no message)");
    if (EnvIs("MBS_GPU_ALLOW_FALLBACK", "1"))
        return false;
    std::fprintf(stderr, "Fix: check the ROCm installation and the visible device, "
                         "or set MBS_GPU_ALLOW_FALLBACK=1.\n");
    std::exit(2);
}

// Per-thread engine contexts, one per device, destroyed at thread exit.
struct ThreadContexts
{
    struct Entry
    {
        mbs_ctx *ctx = nullptr;
        std::vector<double> theta, vf;
        double azStep = 0.0;
        int nAz = -1, nZen = -1;
    };
    std::map<int, Entry> byDevice;
    ~ThreadContexts()
    {
        for (auto &kv : byDevice)
            if (kv.second.ctx)
                mbs_ctx_destroy(kv.second.ctx);
    }
};
thread_local ThreadContexts t_contexts;

// Flat arrays of one batch, in the C ABI's layout (mbs_beam_batch).
struct PackedBatch
{
    std::vector<int64_t> orientOffsets, vertOffsets;
    std::vector<double> weights, scalars, vx, vy, syx, sxy;
    std::vector<int32_t> nv, external;
    std::vector<uint8_t> flags;

    mbs_beam_batch View() const
    {
        mbs_beam_batch b;
        b.n_orient = (int32_t)weights.size();
        b.n_beams = (int32_t)nv.size();
        b.orient_offsets = orientOffsets.data();
        b.weights = weights.data();
        b.beam_scalars = scalars.data();
        b.beam_nv = nv.data();
        b.beam_external = external.data();
        b.vert_offsets = vertOffsets.data();
        b.vx = vx.data();
        b.vy = vy.data();
        b.slope_yx = syx.data();
        b.slope_xy = sxy.data();
        b.edge_flags = flags.data();
        return b;
    }
};

// The job of pack_prepared_gpu_beam (src/cuda/GpuDiffraction.cu:273-358):
// PreparedBeam (src/handler/HandlerPO.h:14-46) -> structure of arrays.
// Returns false for a beam the device path cannot take.
bool Pack(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
          PackedBatch &pk)
{
    pk.orientOffsets.assign(1, 0);
    pk.vertOffsets.assign(1, 0);
    for (int o = start; o < start + nOrient; ++o)
    {
        const PreparedOrientation &po = prepared[o];
        pk.weights.push_back(po.sinZenith);
        for (const PreparedBeam &pb : po.beams)
        {
            const BeamEdgeData &ed = pb.edgeData;
            if (!ed.valid || ed.nVertices < MBS_MIN_VERTICES
                || ed.nVertices > MBS_MAX_VERTICES)
                return false;
            const double s[MBS_BEAM_SCALARS] = {
                pb.horAx, pb.horAy, pb.horAz, pb.verAx, pb.verAy, pb.verAz,
                pb.bdx, pb.bdy, pb.bdz, pb.cenx, pb.ceny, pb.cenz,
                pb.beam_area,
                pb.pNTx, pb.pNTy, pb.pNTz, pb.pNPx, pb.pNPy, pb.pNPz,
                pb.pnxDTx, pb.pnxDTy, pb.pnxDTz, pb.pnxDPx, pb.pnxDPy, pb.pnxDPz,
                pb.jp00r, pb.jp00i, pb.jp01r, pb.jp01i,
                pb.jp10r, pb.jp10i, pb.jp11r, pb.jp11i,
                pb.normx, pb.normy, pb.normz};
            pk.scalars.insert(pk.scalars.end(), s, s + MBS_BEAM_SCALARS);
            pk.nv.push_back(ed.nVertices);
            pk.external.push_back(pb.isExternal ? 1 : 0);
            for (int v = 0; v < ed.nVertices; ++v)
            {
                pk.vx.push_back(ed.x[v]);
                pk.vy.push_back(ed.y[v]);
                pk.syx.push_back(ed.edge_valid_x[v] ? ed.slope_yx[v] : 0.0);
                pk.sxy.push_back(ed.edge_valid_y[v] ? ed.slope_xy[v] : 0.0);
                pk.flags.push_back((uint8_t)((ed.edge_valid_x[v] ? MBS_EDGE_VALID_X : 0)
                                             | (ed.edge_valid_y[v] ? MBS_EDGE_VALID_Y : 0)));
            }
            pk.vertOffsets.push_back((int64_t)pk.vx.size());
        }
        pk.orientOffsets.push_back((int64_t)pk.nv.size());
    }
    return true;
}
}  // namespace

// ---- GpuSupport (src/cuda/GpuSupport.h:30-34) ------------------------------
bool CheckGpuRuntime(GpuDeviceInfo &info, std::string &error)
{
    mbs_device_info d;
    std::memset(&d, 0, sizeof d);
    if (mbs_check_runtime(&d) != MBS_OK)
    {
        error = mbs_last_error(nullptr);
        return false;
    }
    info.deviceId = d.device_id;
    info.name = d.name;
    info.totalGlobalMem = d.total_mem;
    info.freeGlobalMem = d.free_mem;
    info.computeMajor = 9;   // gfx9 family (gfx950)
    info.computeMinor = 5;
    info.runtimeVersion = d.runtime_version;
    info.driverVersion = d.driver_version;
    info.buildArch = 950;
    info.visibleDeviceCount = d.visible_devices;
    info.visibleDevices.clear();
    GpuVisibleDeviceInfo v;
    v.deviceId = d.device_id;
    v.name = d.name;
    v.totalGlobalMem = d.total_mem;
    v.computeMajor = 9;
    v.computeMinor = 5;
    info.visibleDevices.push_back(v);
    error.clear();
    return true;
}

bool QueryActiveGpuMemory(long long &freeBytes, long long &totalBytes, std::string &error)
{
    int64_t f = 0, t = 0;
    if (mbs_query_memory(&f, &t) != MBS_OK)
    {
        error = mbs_last_error(nullptr);
        return false;
    }
    freeBytes = f;
    totalBytes = t;
    error.clear();
    return true;
}

int VisibleGpuDeviceCount() { return mbs_visible_device_count(); }

std::string FormatGpuInfo(const GpuDeviceInfo &info)
{
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "HIP device %d: %s (gfx950 engine), %.1f GiB total, %.1f GiB free, "
                  "%d visible device(s), HIP runtime %d",
                  info.deviceId, info.name.c_str(), info.totalGlobalMem / 1073741824.0,
                  info.freeGlobalMem / 1073741824.0, info.visibleDeviceCount,
                  info.runtimeVersion);
    return buf;
}

// ---- HandlerPO GPU members (src/handler/HandlerPO.h:98-146) -----------------
bool HandlerPO::HandleOrientationsToLocalGpuOnDevice(
    const std::vector<PreparedOrientation> &prepared, int start, int count, int device,
    Arr2D &localM, Arr2D &localM_noshadow, double scale, double waveIndex,
    unsigned long long /*cacheToken*/)
{
    if (!isCoh)
        return false;
    // Rescaled beam sets belong to the multi-size path, not provided here.
    if (scale != 1.0 || (waveIndex != 0.0 && waveIndex != m_waveIndex))
        return false;
    if (start < 0 || count < 0 || start > (int)prepared.size())
        return false;
    const int nOrient = std::min(count, (int)prepared.size() - start);
    const int nAz = m_sphere.nAzimuth;
    const int nZen = m_sphere.nZenith;
    if (nOrient == 0 || nAz * (nZen + 1) == 0)
        return true;

    PackedBatch pk;
    if (!Pack(prepared, start, nOrient, pk))
        return false;
    if (pk.nv.empty())
        return true;

    ThreadContexts::Entry &e = t_contexts.byDevice[device];
    if (!e.ctx && mbs_ctx_create(device, &e.ctx) != MBS_OK)
    {
        e.ctx = nullptr;
        return DeviceFailure("mbs_ctx_create", mbs_last_error(nullptr));
    }
    // Grid: zenith values, azimuth step and the transverse basis vf of
    // ScatteringRange (src/handler/Handler.h:15-189); uploaded when changed.
    std::vector<double> theta(nZen + 1), vf((size_t)nAz * (nZen + 1) * 3);
    for (int t = 0; t <= nZen; ++t)
        theta[t] = m_sphere.GetZenith(t);
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            double *w = &vf[((size_t)p * (nZen + 1) + t) * 3];
            w[0] = m_sphere.vf[p][t].x;
            w[1] = m_sphere.vf[p][t].y;
            w[2] = m_sphere.vf[p][t].z;
        }
    if (e.nAz != nAz || e.nZen != nZen || e.azStep != m_sphere.azinuthStep
        || e.theta != theta || e.vf != vf)
    {
        mbs_grid g;
        g.n_az = nAz;
        g.n_zen = nZen;
        g.theta = theta.data();
        g.az_step = m_sphere.azinuthStep;
        g.vf = vf.data();
        const int rc = mbs_set_grid(e.ctx, &g);
        if (rc == MBS_ERR_DEVICE)
            return DeviceFailure("mbs_set_grid", mbs_last_error(e.ctx));
        if (rc != MBS_OK)
            return false;
        e.nAz = nAz;
        e.nZen = nZen;
        e.azStep = m_sphere.azinuthStep;
        e.theta.swap(theta);
        e.vf.swap(vf);
    }

    // Handler constants (src/handler/Handler.cpp:102-126).
    mbs_consts c;
    std::memset(&c, 0, sizeof c);
    c.wave_index = m_waveIndex;
    c.eps1 = m_eps1;
    c.compl_wave[0] = real(m_complWave);
    c.compl_wave[1] = imag(m_complWave);
    c.inv_compl_wave[0] = real(m_invComplWave);
    c.inv_compl_wave[1] = imag(m_invComplWave);
    c.legacy_sign = m_legacySign ? 1 : 0;

    const bool noShadow = ComputeNoShadow();
    const bool timing = EnvIs("MBS_GPU_TIMING", "1");
    mbs_set_timing(e.ctx, timing ? 1 : 0);
    const mbs_beam_batch view = pk.View();
    int rc = mbs_upload_beams(e.ctx, &view, &c);
    if (rc == MBS_OK)
        rc = mbs_diffract(e.ctx, PrecisionFlag() | (noShadow ? MBS_NOSHADOW : 0u));
    std::vector<double> mFull((size_t)nAz * (nZen + 1) * 16, 0.0);
    std::vector<double> mNs(noShadow ? mFull.size() : 0, 0.0);
    if (rc == MBS_OK)
        rc = mbs_accumulate_mueller(e.ctx, mFull.data(), noShadow ? mNs.data() : nullptr);
    if (rc == MBS_ERR_DEVICE)
        return DeviceFailure("diffraction", mbs_last_error(e.ctx));
    if (rc != MBS_OK)
    {
        std::fprintf(stderr, "HIP diffraction backend: %s\n", mbs_last_error(e.ctx));
        return false;
    }
    if (timing)
    {
        mbs_timing tm;
        if (mbs_get_timing(e.ctx, &tm) == MBS_OK)
            std::fprintf(stderr,
                         "GPU timing: %d orientations, %zu beams: upload %.3f ms, kernel %.3f "
                         "ms, reduce %.3f ms (%d work items, %d workgroups)\n",
                         nOrient, pk.nv.size(), tm.upload_ms, tm.kernel_ms, tm.reduce_ms,
                         tm.n_items, tm.n_blocks);
    }
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const size_t cell = ((size_t)p * (nZen + 1) + t) * 16;
            double *dst = localM.RawCell(p, t);
            for (int k = 0; k < 16; ++k)
                dst[k] += mFull[cell + k];
            if (noShadow)
            {
                double *dn = localM_noshadow.RawCell(p, t);
                for (int k = 0; k < 16; ++k)
                    dn[k] += mNs[cell + k];
            }
        }
    return true;
}

bool HandlerPO::HandleOrientationsToLocalGpu(const std::vector<PreparedOrientation> &prepared,
                                             int start, int count, Arr2D &localM,
                                             Arr2D &localM_noshadow, double scale,
                                             double waveIndex)
{
    return HandleOrientationsToLocalGpuOnDevice(prepared, start, count, 0, localM,
                                                localM_noshadow, scale, waveIndex, 0);
}

bool HandlerPO::HandleOrientationsToLocalGpu(const std::vector<PreparedOrientation> &prepared,
                                             Arr2D &localM, Arr2D &localM_noshadow)
{
    return HandleOrientationsToLocalGpu(prepared, 0, (int)prepared.size(), localM,
                                        localM_noshadow, 1.0, 0.0);
}

bool HandlerPO::HandleBeamsToLocalGpu(const PreparedOrientation &prepared, Arr2D &localM,
                                      Arr2D &localM_noshadow)
{
    const std::vector<PreparedOrientation> one(1, prepared);
    return HandleOrientationsToLocalGpu(one, 0, 1, localM, localM_noshadow, 1.0, 0.0);
}

bool HandlerPO::HandleOrientationsToLocalGpuFftPhi(const std::vector<PreparedOrientation> &,
                                                   int, int, Arr2D &, Arr2D &, double,
                                                   double)
{
    return false;
}

bool HandlerPO::HandleOrientationsToLocalGpuMultiK(const std::vector<PreparedOrientation> &,
                                                   int, int, const std::vector<double> &,
                                                   double, std::vector<Arr2D> &)
{
    return false;
}

bool HandlerPO::HandleOrientationsToLocalGpuFftPhiMultiK(
    const std::vector<PreparedOrientation> &, int, int, const std::vector<double> &, double,
    std::vector<Arr2D> &)
{
    return false;
}

bool HandlerPO::DiffractThetasGpu(const std::vector<PreparedOrientation> &, const double *,
                                  int, std::vector<double> &)
{
    return false;
}

// Batches are bounded by beam count so that one call's vertex and beam
// records stay well inside device memory; the engine itself splits work
// evenly, so the batch size only bounds host packing memory.
int HandlerPO::SelectGpuOrientationBatchSize(const std::vector<PreparedOrientation> &prepared,
                                             int start, int maxCount) const
{
    if (maxCount <= 0)
        return 1;
    const size_t kMaxBeams = (size_t)1 << 22;
    size_t beams = 0;
    int n = 0;
    for (int o = start; o < (int)prepared.size() && n < maxCount; ++o, ++n)
    {
        beams += prepared[o].beams.size();
        if (n > 0 && beams > kMaxBeams)
            break;
    }
    return std::max(n, 1);
}
