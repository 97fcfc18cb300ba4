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
//   src/handler/HandlerPO.h:98-162    HandlerPO GPU members
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
// Multi-size (rescaled beam sets): the scale / waveIndex arguments with the
// reference's scale-on-pack rule, and the fused multi-scale member, which
// packs once and applies the same rule on the device.
// Theta probe (DiffractThetasGpu): on the engine's incoherent FP64 sum.
// Phi-FFT interpolation (HandleOrientationsToLocalGpuFftPhi and its
// multi-size form): the direct path on a coarse azimuth grid and the
// reference's phi and theta upsampling, validation and adaptive rows.
#include "HandlerPO.h"
#include "GpuSupport.h"

#include "mbs_hip.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
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

// Flat arrays of one batch, in the C ABI's layout (mbs_beam_batch).
struct PackedBatch
{
    std::vector<int64_t> orientOffsets, vertOffsets;
    std::vector<double> weights, scalars, vx, vy, syx, sxy;
    std::vector<int32_t> nv, external;
    std::vector<uint8_t> flags;
    // per-beam data of a device rescale (mbs_beam_scaling), when packed
    std::vector<double> rawJones, path, sign, absPaths;
    std::vector<int64_t> absOffsets;

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

// Per-thread engine contexts, one per device, destroyed at thread exit.
struct ThreadContexts
{
    struct Entry
    {
        mbs_ctx *ctx = nullptr;
        std::vector<double> theta, vf;
        double azStep = 0.0;
        int nAz = -1, nZen = -1;
        // host staging of the packed batch, kept so that later calls reuse
        // its pages instead of allocating and faulting in ~2 GB per batch
        PackedBatch pk;
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

// A rescaled beam set (the multi-size path): the particle, and with it every
// beam, is scaled by `scale` about the origin.  As the reference's pack does
// (src/cuda/GpuDiffraction.cu:273-358, applied when |scale - 1| > 1e-15,
// :4324), aperture vertices and the beam centre are multiplied by the scale,
// the area by its square, and J_phased is rebuilt from the beam's raw Jones
// matrix: J_phased = sign * a(scale) * J * exp(i waveIndex * path * scale),
// path = the optical path (external beams) or the projected length
// (internal beams), sign = -1 for external beams and for internal beams with
// an odd number of interactions, a(scale) the mean over the beam's
// absorption paths p of exp(cAbs * p * scale) (1 for p <= DBL_EPSILON).
// Slopes, axes and polarisation data do not change with size.
struct PackScale
{
    bool on = false;
    double scale = 1.0, waveIndex = 0.0, cAbs = 0.0;

    PackScale() = default;
    PackScale(double s, double k, double c)
        : on(std::fabs(s - 1.0) > 1e-15), scale(s), waveIndex(k), cAbs(c) {}
};

double AbsorptionFactor(const PreparedBeam &pb, double scale, double cAbs)
{
    if (pb.absorptionPaths.empty())
        return 1.0;
    double sum = 0.0;
    for (double path : pb.absorptionPaths)
        sum += path > DBL_EPSILON ? std::exp(cAbs * path * scale) : 1.0;
    return sum / (double)pb.absorptionPaths.size();
}

void ScaledJones(const PreparedBeam &pb, const PackScale &sc, double jp[8])
{
    const double path = (pb.isExternal ? pb.origBeam.opticalPath : pb.info.projLenght)
                        * sc.scale;
    double sign = pb.isExternal ? -1.0 : 1.0;
    if (!pb.isExternal && (pb.origBeam.nActs & 1))
        sign = -sign;
    const double amp = sign * AbsorptionFactor(pb, sc.scale, sc.cAbs);
    const double sn = std::sin(sc.waveIndex * path), cs = std::cos(sc.waveIndex * path);
    const ::complex raw[4] = {pb.origBeam.J.m11, pb.origBeam.J.m12, pb.origBeam.J.m21,
                              pb.origBeam.J.m22};
    for (int k = 0; k < 4; ++k)
    {
        const double jr = real(raw[k]) * amp, ji = imag(raw[k]) * amp;
        jp[2 * k] = jr * cs - ji * sn;
        jp[2 * k + 1] = jr * sn + ji * cs;
    }
}

// The job of pack_prepared_gpu_beam (src/cuda/GpuDiffraction.cu:273-358):
// PreparedBeam (src/handler/HandlerPO.h:14-46) -> structure of arrays.
// Two passes over the orientations, both OpenMP-parallel as the reference's
// packing is (:4407-4428): beam and vertex counts (which also validate every
// beam), then, after a prefix sum, every orientation fills its own slice.
// Returns false for a beam the device path cannot take.  With `scaling`
// the per-beam data of a device rescale (mbs_beam_scaling) are packed too:
// the raw Jones matrix, the path and sign of ScaledJones, the absorption
// paths.
bool Pack(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
          const PackScale &sc, PackedBatch &pk, bool scaling = false)
{
    std::vector<int64_t> nb(nOrient), nvert(nOrient), nabs(nOrient);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
    for (int i = 0; i < nOrient; ++i)
    {
        int64_t v = 0, a = 0;
        for (const PreparedBeam &pb : prepared[start + i].beams)
        {
            const BeamEdgeData &ed = pb.edgeData;
            if (!ed.valid || ed.nVertices < MBS_MIN_VERTICES
                || ed.nVertices > MBS_MAX_VERTICES)
                bad = 1;
            v += ed.nVertices;
            a += (int64_t)pb.absorptionPaths.size();
        }
        nb[i] = (int64_t)prepared[start + i].beams.size();
        nvert[i] = v;
        nabs[i] = a;
    }
    if (bad)
        return false;
    pk.orientOffsets.assign(nOrient + 1, 0);
    std::vector<int64_t> vbase(nOrient + 1, 0), abase(nOrient + 1, 0);
    for (int i = 0; i < nOrient; ++i)
    {
        pk.orientOffsets[i + 1] = pk.orientOffsets[i] + nb[i];
        vbase[i + 1] = vbase[i] + nvert[i];
        abase[i + 1] = abase[i] + nabs[i];
    }
    if (scaling)
    {
        const size_t n = (size_t)pk.orientOffsets[nOrient];
        pk.rawJones.resize(n * 8);
        pk.path.resize(n);
        pk.sign.resize(n);
        pk.absOffsets.resize(n + 1);
        pk.absOffsets[n] = abase[nOrient];
        pk.absPaths.resize((size_t)abase[nOrient]);
    }
    const size_t nBeams = (size_t)pk.orientOffsets[nOrient];
    const size_t nVerts = (size_t)vbase[nOrient];
    pk.weights.resize(nOrient);
    pk.scalars.resize(nBeams * MBS_BEAM_SCALARS);
    pk.nv.resize(nBeams);
    pk.external.resize(nBeams);
    pk.vertOffsets.resize(nBeams + 1);
    pk.vx.resize(nVerts);
    pk.vy.resize(nVerts);
    pk.syx.resize(nVerts);
    pk.sxy.resize(nVerts);
    pk.flags.resize(nVerts);
    pk.vertOffsets[nBeams] = (int64_t)nVerts;
    const double g = sc.on ? sc.scale : 1.0, g2 = g * g;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nOrient; ++i)
    {
        const PreparedOrientation &po = prepared[start + i];
        pk.weights[i] = po.sinZenith;
        size_t bi = (size_t)pk.orientOffsets[i];
        size_t vi = (size_t)vbase[i];
        size_t ai = (size_t)abase[i];
        for (const PreparedBeam &pb : po.beams)
        {
            const BeamEdgeData &ed = pb.edgeData;
            if (scaling)
            {
                const ::complex raw[4] = {pb.origBeam.J.m11, pb.origBeam.J.m12,
                                          pb.origBeam.J.m21, pb.origBeam.J.m22};
                for (int k = 0; k < 4; ++k)
                {
                    pk.rawJones[bi * 8 + 2 * k] = real(raw[k]);
                    pk.rawJones[bi * 8 + 2 * k + 1] = imag(raw[k]);
                }
                pk.path[bi] = pb.isExternal ? pb.origBeam.opticalPath : pb.info.projLenght;
                double sg = pb.isExternal ? -1.0 : 1.0;
                if (!pb.isExternal && (pb.origBeam.nActs & 1))
                    sg = -sg;
                pk.sign[bi] = sg;
                pk.absOffsets[bi] = (int64_t)ai;
                for (double p : pb.absorptionPaths)
                    pk.absPaths[ai++] = p;
            }
            double jp[8] = {pb.jp00r, pb.jp00i, pb.jp01r, pb.jp01i,
                            pb.jp10r, pb.jp10i, pb.jp11r, pb.jp11i};
            if (sc.on)
                ScaledJones(pb, sc, jp);
            const double s[MBS_BEAM_SCALARS] = {
                pb.horAx, pb.horAy, pb.horAz, pb.verAx, pb.verAy, pb.verAz,
                pb.bdx, pb.bdy, pb.bdz, pb.cenx * g, pb.ceny * g, pb.cenz * g,
                pb.beam_area * g2,
                pb.pNTx, pb.pNTy, pb.pNTz, pb.pNPx, pb.pNPy, pb.pNPz,
                pb.pnxDTx, pb.pnxDTy, pb.pnxDTz, pb.pnxDPx, pb.pnxDPy, pb.pnxDPz,
                jp[0], jp[1], jp[2], jp[3], jp[4], jp[5], jp[6], jp[7],
                pb.normx, pb.normy, pb.normz};
            std::memcpy(&pk.scalars[bi * MBS_BEAM_SCALARS], s, sizeof s);
            pk.nv[bi] = ed.nVertices;
            pk.external[bi] = pb.isExternal ? 1 : 0;
            pk.vertOffsets[bi] = (int64_t)vi;
            for (int v = 0; v < ed.nVertices; ++v, ++vi)
            {
                pk.vx[vi] = ed.x[v] * g;
                pk.vy[vi] = ed.y[v] * g;
                pk.syx[vi] = ed.edge_valid_x[v] ? ed.slope_yx[v] : 0.0;
                pk.sxy[vi] = ed.edge_valid_y[v] ? ed.slope_xy[v] : 0.0;
                pk.flags[vi] = (uint8_t)((ed.edge_valid_x[v] ? MBS_EDGE_VALID_X : 0)
                                         | (ed.edge_valid_y[v] ? MBS_EDGE_VALID_Y : 0));
            }
            ++bi;
        }
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
namespace
{
// Grid of ScatteringRange (src/handler/Handler.h:15-189): zenith values,
// azimuth step and the transverse basis vf.
struct GridSpec
{
    int nAz = 0, nZen = 0;
    double azStep = 0.0;
    std::vector<double> theta, vf;
};

GridSpec MakeGrid(const ScatteringRange &sphere)
{
    GridSpec g;
    g.nAz = sphere.nAzimuth;
    g.nZen = sphere.nZenith;
    g.azStep = sphere.azinuthStep;
    g.theta.resize(g.nZen + 1);
    g.vf.resize((size_t)g.nAz * (g.nZen + 1) * 3);
    for (int t = 0; t <= g.nZen; ++t)
        g.theta[t] = sphere.GetZenith(t);
    for (int p = 0; p < g.nAz; ++p)
        for (int t = 0; t <= g.nZen; ++t)
        {
            double *w = &g.vf[((size_t)p * (g.nZen + 1) + t) * 3];
            w[0] = sphere.vf[p][t].x;
            w[1] = sphere.vf[p][t].y;
            w[2] = sphere.vf[p][t].z;
        }
    return g;
}

// This thread's context on `device`, created on first use, with the grid
// uploaded when it changed.  Returns null after a failure (the caller
// returns `ok`).
ThreadContexts::Entry *Context(int device, GridSpec &g, bool &ok)
{
    ThreadContexts::Entry &e = t_contexts.byDevice[device];
    if (!e.ctx && mbs_ctx_create(device, &e.ctx) != MBS_OK)
    {
        e.ctx = nullptr;
        ok = DeviceFailure("mbs_ctx_create", mbs_last_error(nullptr));
        return nullptr;
    }
    if (e.nAz != g.nAz || e.nZen != g.nZen || e.azStep != g.azStep || e.theta != g.theta
        || e.vf != g.vf)
    {
        mbs_grid mg;
        mg.n_az = g.nAz;
        mg.n_zen = g.nZen;
        mg.theta = g.theta.data();
        mg.az_step = g.azStep;
        mg.vf = g.vf.data();
        const int rc = mbs_set_grid(e.ctx, &mg);
        if (rc != MBS_OK)
        {
            ok = rc == MBS_ERR_DEVICE ? DeviceFailure("mbs_set_grid", mbs_last_error(e.ctx))
                                      : false;
            return nullptr;
        }
        e.nAz = g.nAz;
        e.nZen = g.nZen;
        e.azStep = g.azStep;
        e.theta.swap(g.theta);
        e.vf.swap(g.vf);
    }
    return &e;
}

// Diffracts one packed batch and adds the result into localM (and into
// *localMns when it is given).
bool RunBatch(ThreadContexts::Entry &e, const PackedBatch &pk, const mbs_consts &c,
              Arr2D &localM, Arr2D *localMns, double packMs, bool upload = true)
{
    const int nAz = e.nAz, nZen = e.nZen;
    const bool timing = EnvIs("MBS_GPU_TIMING", "1");
    const auto t0 = std::chrono::steady_clock::now();
    mbs_set_timing(e.ctx, timing ? 1 : 0);
    const mbs_beam_batch view = pk.View();
    int rc = upload ? mbs_upload_beams(e.ctx, &view, &c) : MBS_OK;
    if (rc == MBS_OK)
        rc = mbs_diffract(e.ctx, PrecisionFlag() | (localMns ? MBS_NOSHADOW : 0u));
    std::vector<double> mFull((size_t)nAz * (nZen + 1) * 16, 0.0);
    std::vector<double> mNs(localMns ? mFull.size() : 0, 0.0);
    if (rc == MBS_OK)
        rc = mbs_accumulate_mueller(e.ctx, mFull.data(), localMns ? mNs.data() : nullptr);
    if (rc == MBS_ERR_DEVICE)
        return DeviceFailure("diffraction", mbs_last_error(e.ctx));
    if (rc != MBS_OK)
    {
        std::fprintf(stderr, "HIP diffraction backend: %s\n", mbs_last_error(e.ctx));
        return false;
    }
    const double callMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (timing)
    {
        mbs_timing tm;
        if (mbs_get_timing(e.ctx, &tm) == MBS_OK)
            std::fprintf(stderr,
                         "GPU timing: %d orientations, %zu beams: pack %.3f ms, upload %.3f ms, "
                         "kernel %.3f ms, reduce %.3f ms, engine calls %.3f ms (%d work items, "
                         "%d workgroups)\n",
                         (int)pk.weights.size(), pk.nv.size(), packMs, tm.upload_ms,
                         tm.kernel_ms, tm.reduce_ms, callMs, tm.n_items, tm.n_blocks);
    }
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const size_t cell = ((size_t)p * (nZen + 1) + t) * 16;
            double *dst = localM.RawCell(p, t);
            for (int k = 0; k < 16; ++k)
                dst[k] += mFull[cell + k];
            if (localMns)
            {
                double *dn = localMns->RawCell(p, t);
                for (int k = 0; k < 16; ++k)
                    dn[k] += mNs[cell + k];
            }
        }
    return true;
}
}  // namespace

// Handler constants (src/handler/Handler.cpp:102-126).
static mbs_consts HandlerConsts(double waveIndex, double eps1, const ::complex &complWave,
                                const ::complex &invComplWave, bool legacySign)
{
    mbs_consts c;
    std::memset(&c, 0, sizeof c);
    c.wave_index = waveIndex;
    c.eps1 = eps1;
    c.compl_wave[0] = real(complWave);
    c.compl_wave[1] = imag(complWave);
    c.inv_compl_wave[0] = real(invComplWave);
    c.inv_compl_wave[1] = imag(invComplWave);
    c.legacy_sign = legacySign ? 1 : 0;
    return c;
}

// scale != 1 is the multi-size path's rescaled beam set (see PackScale); the
// aperture integrals keep the handler's own wave index, `waveIndex` only
// re-phases J_phased, as in the reference.
bool HandlerPO::HandleOrientationsToLocalGpuOnDevice(
    const std::vector<PreparedOrientation> &prepared, int start, int count, int device,
    Arr2D &localM, Arr2D &localM_noshadow, double scale, double waveIndex,
    unsigned long long /*cacheToken*/)
{
    if (!isCoh)
        return false;
    if (start < 0 || count < 0 || start > (int)prepared.size())
        return false;
    const int nOrient = std::min(count, (int)prepared.size() - start);
    if (nOrient == 0 || m_sphere.nAzimuth * (m_sphere.nZenith + 1) == 0)
        return true;

    GridSpec g = MakeGrid(m_sphere);
    bool ok = false;
    ThreadContexts::Entry *e = Context(device, g, ok);
    if (!e)
        return ok;
    PackedBatch &pk = e->pk;
    const auto tp = std::chrono::steady_clock::now();
    if (!Pack(prepared, start, nOrient, PackScale(scale, waveIndex, AbsorptionCoefficient()),
              pk))
        return false;
    const double packMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
    if (pk.nv.empty())
        return true;
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    return RunBatch(*e, pk, c, localM, ComputeNoShadow() ? &localM_noshadow : nullptr, packMs);
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

namespace
{
int EnvInt(const char *name, int lo, int hi, int fallback)
{
    const char *v = std::getenv(name);
    if (!v || !*v)
        return fallback;
    char *end = nullptr;
    const long x = std::strtol(v, &end, 10);
    return (end && *end == '\0' && x >= lo && x <= hi) ? (int)x : fallback;
}

// choose_fft_phi_factor (src/cuda/GpuDiffraction.cu:712-727, with
// gpu_fft_phi_factor :513-531): the handler's factor, else MBS_FFT_PHI_FACTOR
// (1..64, "auto" = unset), else by the full azimuth count.
int ChoosePhiFactor(int nFull, int configured)
{
    if (configured > 0)
        return configured;
    const int requested = EnvInt("MBS_FFT_PHI_FACTOR", 1, 64, 0);
    if (requested > 0)
        return requested;
    return nFull >= 2400 ? 8 : nFull >= 600 ? 6 : nFull >= 240 ? 3 : 2;
}

// Adds to dst(q, t), q < nFull, t <= nZen, the band-limited interpolation in
// phi of low(p, t), p < nLow: the upsampling of the reference's phi-FFT path
// (fft_upsample_phi_arr2d :3195-3330, fft_phi_pad_kernel :794-814): forward
// DFT of the nLow samples; frequency k at k for k <= nLow/2 (the Nyquist term
// of an even nLow on the positive side), else at k - nLow; scaled by 1/nLow;
// inverse DFT over nFull; real part.  As a real matrix,
// y[q] = sum_p K(q/nFull - p/nLow) x[p], K(u) = (1/nLow) sum_k cos(2 pi f_k u),
// with the sum in closed form (Dirichlet kernel): sin((M + 1/2) a)/sin(a/2) for
// nLow = 2M + 1, sin((M - 1/2) a)/sin(a/2) + cos(M a) for nLow = 2M,
// a = 2 pi u, and nLow at u = 0.
std::vector<double> UpsamplingMatrix(int nLow, int nFull)
{
    std::vector<double> K((size_t)nFull * nLow);
    const int M = nLow / 2;
    for (int q = 0; q < nFull; ++q)
        for (int p = 0; p < nLow; ++p)
        {
            double k = 1.0;
            if ((int64_t)q * nLow != (int64_t)p * nFull)
            {
                const double a = 2.0 * M_PI * ((double)q / nFull - (double)p / nLow);
                const double h = std::sin(0.5 * a);
                k = (nLow & 1) ? std::sin((M + 0.5) * a) / h / nLow
                               : (std::sin((M - 0.5) * a) / h + std::cos(M * a)) / nLow;
            }
            K[(size_t)q * nLow + p] = k;
        }
    return K;
}

void UpsamplePhi(const Arr2D &low, int nLow, int nFull, int nZen, Arr2D &dst)
{
    const std::vector<double> K = UpsamplingMatrix(nLow, nFull);
#pragma omp parallel for schedule(static)
    for (int t = 0; t <= nZen; ++t)
    {
        std::vector<double> x((size_t)nLow * 16);
        for (int p = 0; p < nLow; ++p)
            std::memcpy(&x[(size_t)p * 16], low.RawCell(p, t), 16 * sizeof(double));
        for (int q = 0; q < nFull; ++q)
        {
            double acc[16] = {0};
            const double *kq = &K[(size_t)q * nLow];
            for (int p = 0; p < nLow; ++p)
                for (int e = 0; e < 16; ++e)
                    acc[e] += kq[p] * x[(size_t)p * 16 + e];
            double *cell = dst.RawCell(q, t);
            for (int e = 0; e < 16; ++e)
                cell[e] += acc[e];
        }
    }
}
// The theta upsampling of the same path (fft_upsample_theta_arr2d,
// :3343-3420) for a uniform zenith grid: the nThetaLow + 1 rows of each
// azimuth column extended evenly to a period of 2 nThetaLow (row t and row
// 2 nThetaLow - t), upsampled by the rule above to 2 nThetaFull, rows
// 0..nThetaFull added into dst (p < nAz).
void UpsampleTheta(const Arr2D &low, int nAz, int nThetaLow, int nThetaFull, Arr2D &dst)
{
    const int nLow = 2 * nThetaLow, nFull = 2 * nThetaFull;
    const std::vector<double> K = UpsamplingMatrix(nLow, nFull);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < nAz; ++p)
    {
        std::vector<double> x((size_t)nLow * 16);
        for (int t = 0; t < nLow; ++t)
            std::memcpy(&x[(size_t)t * 16], low.RawCell(p, t <= nThetaLow ? t : nLow - t),
                        16 * sizeof(double));
        for (int q = 0; q <= nThetaFull; ++q)
        {
            double acc[16] = {0};
            const double *kq = &K[(size_t)q * nLow];
            for (int t = 0; t < nLow; ++t)
                for (int e = 0; e < 16; ++e)
                    acc[e] += kq[t] * x[(size_t)t * 16 + e];
            double *cell = dst.RawCell(p, q);
            for (int e = 0; e < 16; ++e)
                cell[e] += acc[e];
        }
    }
}

// The coarse-grid sphere of the phi-FFT path (:3563-3570): nLow azimuths and,
// with a theta factor, nZen / factor zenith rows.
ScatteringRange CoarseSphere(const ScatteringRange &full, int nLow, int nZenLow,
                             const Light &incident)
{
    ScatteringRange low = full;
    low.nAzimuth = nLow;
    low.azinuthStep = 2.0 * M_PI / nLow;
    low.nZenith = nZenLow;
    low.zenithStep = (low.zenithEnd - low.zenithStart) / nZenLow;
    low.ComputeSphereDirections(incident);
    return low;
}

// The theta factor (gpu_fft_theta_factor, :532-542, and :3543-3550): 1 for
// non-uniform or short grids or when it does not divide nZen.
int ThetaFactor(const ScatteringRange &sphere)
{
    const int f = EnvInt("MBS_FFT_THETA_FACTOR", 1, 16, 1);
    if (f <= 1 || sphere.isNonUniform || sphere.nZenith < 32 || sphere.nZenith % f != 0)
        return 1;
    return f;
}

// report_fft_check (src/cuda/GpuDiffraction.cu:3142-3195): the upsampled
// grid against a reference grid -- M11 errors relative to the local M11 (and
// to the cells with M11 >= 1e-3 of its maximum), and all 16 elements relative
// to max(|M11|, that floor) -- printed; the estimate is the larger of the
// significant-M11 and the element figures.
double ReportFftCheck(const Arr2D &fftM, const Arr2D &refM, int nAz, int nZen, const char *name)
{
    double refMax = 0.0;
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
            refMax = std::max(refMax, std::fabs(refM.RawCell(p, t)[0]));
    const double significant = 1e-3 * refMax;
    double maxAbs = 0.0, maxRel = 0.0, sigRel = 0.0, normRel = 0.0, rms = 0.0;
    int samples = 0, sigSamples = 0;
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const double *r = refM.RawCell(p, t), *f = fftM.RawCell(p, t);
            const double absErr = std::fabs(f[0] - r[0]);
            const double relErr = absErr / std::max(std::fabs(r[0]), 1e-12);
            maxAbs = std::max(maxAbs, absErr);
            maxRel = std::max(maxRel, relErr);
            if (std::fabs(r[0]) >= significant)
            {
                sigRel = std::max(sigRel, relErr);
                ++sigSamples;
            }
            const double scale = std::max(std::fabs(r[0]), significant);
            for (int e = 0; e < 16; ++e)
                normRel = std::max(normRel, std::fabs(f[e] - r[e]) / scale);
            rms += relErr * relErr;
            ++samples;
        }
    rms = std::sqrt(rms / std::max(1, samples));
    std::fprintf(stderr,
                 "GPU FFT check %s: M11 max_abs=%.6g max_rel=%.6g significant_max_rel=%.6g "
                 "rms_rel=%.6g, Mueller/M11 max_rel=%.6g, samples=%d significant_samples=%d\n",
                 name, maxAbs, maxRel, sigRel, rms, normRel, samples, sigSamples);
    return std::max(sigRel, normRel);
}

double EnvDouble(const char *name, double fallback)
{
    const char *v = std::getenv(name);
    if (!v || !*v)
        return fallback;
    char *end = nullptr;
    const double x = std::strtod(v, &end);
    return (end && *end == '\0' && x > 0.0) ? x : fallback;
}

// azimuth_average_output_cell (:3052-3084): one theta row's azimuth mean,
// the linear-polarisation columns rotated by L(-phi) off the poles.
void RowAverage(const Arr2D &src, int nAz, int nZen, int t, double out[16])
{
    for (int k = 0; k < 16; ++k)
        out[k] = 0.0;
    const bool pole = t == 0 || t == nZen;
    for (int p = 0; p < nAz; ++p)
    {
        const double *m = src.RawCell(p, t);
        if (pole)
        {
            for (int k = 0; k < 16; ++k)
                out[k] += m[k];
            continue;
        }
        const double a = -p * (2.0 * M_PI / (double)nAz);
        const double cs = std::cos(2.0 * a), sn = std::sin(2.0 * a);
        for (int r = 0; r < 4; ++r)
        {
            out[4 * r] += m[4 * r];
            out[4 * r + 1] += m[4 * r + 1] * cs - m[4 * r + 2] * sn;
            out[4 * r + 2] += m[4 * r + 1] * sn + m[4 * r + 2] * cs;
            out[4 * r + 3] += m[4 * r + 3];
        }
    }
    for (int k = 0; k < 16; ++k)
        out[k] *= 1.0 / (double)nAz;
}

// detect_fft_phi_spike_rows (:3086-3135): interior theta rows whose averaged
// off-diagonal ratios M_ij/M11 (ij = 23, 24, 32, 34, 42, 43, 13, 14) leave
// the mean of the two neighbouring rows by at least `threshold`; the
// `maxRows` highest scores, in row order.
std::vector<int> SpikeRows(const Arr2D &m, int nAz, int nZen, double threshold, int maxRows)
{
    std::vector<int> rows;
    if (maxRows <= 0 || nAz <= 0 || nZen < 2)
        return rows;
    const int elems[8] = {6, 7, 9, 11, 13, 14, 2, 3};
    std::vector<std::pair<double, int>> scored;
    double prev[16], cur[16], next[16];
    for (int t = 1; t < nZen; ++t)
    {
        RowAverage(m, nAz, nZen, t - 1, prev);
        RowAverage(m, nAz, nZen, t, cur);
        RowAverage(m, nAz, nZen, t + 1, next);
        const double c11 = std::max(std::fabs(cur[0]), 1e-30);
        const double p11 = std::max(std::fabs(prev[0]), 1e-30);
        const double n11 = std::max(std::fabs(next[0]), 1e-30);
        double score = 0.0;
        for (int e : elems)
            score = std::max(score, std::fabs(cur[e] / c11 - 0.5 * (prev[e] / p11 + next[e] / n11)));
        if (score >= threshold)
            scored.push_back(std::make_pair(score, t));
    }
    std::sort(scored.begin(), scored.end(),
              [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                  return a.first > b.first;
              });
    for (const auto &s : scored)
    {
        rows.push_back(s.second);
        if ((int)rows.size() >= maxRows)
            break;
    }
    std::sort(rows.begin(), rows.end());
    return rows;
}

void AddCells(const Arr2D &src, int nAz, int nZen, Arr2D &dst)
{
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const double *a = src.RawCell(p, t);
            double *b = dst.RawCell(p, t);
            for (int e = 0; e < 16; ++e)
                b[e] += a[e];
        }
}
}  // namespace

// The reference's angular Fourier interpolation (src/cuda/GpuDiffraction.cu:
// 3504-3839): the batch is diffracted on a grid of nLow = nFull / factor
// azimuths (at least 16) with the handler's zenith rows, and the Mueller grids
// are upsampled in phi to the handler's nFull azimuths and added into localM
// (and localM_noshadow).  A factor of 1, nFull < 32 or nLow >= nFull runs the
// direct path.  With MBS_FFT_CHECK=1 or a tolerance the reference's
// validation and refinement run too, with MBS_FFT_ADAPTIVE_PHI its direct
// recomputation of spiking theta rows, and with MBS_FFT_THETA_FACTOR (a
// uniform zenith grid of at least 32 rows that the factor divides) the
// coarse grid has nZen / factor rows, upsampled in theta first.
bool HandlerPO::HandleOrientationsToLocalGpuFftPhi(const std::vector<PreparedOrientation> &prepared,
                                                   int start, int count, Arr2D &localM,
                                                   Arr2D &localM_noshadow, double scale,
                                                   double waveIndex)
{
    if (!isCoh)
        return false;
    const int nFull = m_sphere.nAzimuth;
    const int nZen = m_sphere.nZenith;
    const int factor = ChoosePhiFactor(nFull, m_fftPhiFactor);
    int nLow = factor > 1 ? nFull / factor : nFull;
    if (nLow < 16)
        nLow = std::min(nFull, 16);
    if (factor <= 1 || nFull < 32 || nLow >= nFull)
    {
        m_fftStatistics->RecordCall(nFull, nFull, 1);
        return HandleOrientationsToLocalGpu(prepared, start, count, localM, localM_noshadow,
                                            scale, waveIndex);
    }
    const char *check = std::getenv("MBS_FFT_CHECK");
    const char *adaptive = std::getenv("MBS_FFT_ADAPTIVE_PHI");
    const int thetaFactor = ThetaFactor(m_sphere);
    const int nZenLow = thetaFactor > 1 ? std::max(2, nZen / thetaFactor) : nZen;
    const bool doCheck = (check && std::strcmp(check, "1") == 0) || m_fftTolerance > 0.0;
    const bool adaptiveRows = adaptive && *adaptive && std::strcmp(adaptive, "0") != 0;
    m_fftStatistics->RecordCall(nFull, nLow, factor);
    static bool printed = false;
    if (!printed)
    {
        std::fprintf(stderr,
                     "GPU FFT angular interpolation: direct Nphi=%d, output Nphi=%d (factor=%d), "
                     "direct Ntheta=%d, output Ntheta=%d (factor=%d). "
                     "This is angular Fourier interpolation, not aperture pFFT/FMM.\n",
                     nLow, nFull, factor, nZenLow + 1, nZen + 1, thetaFactor);
        printed = true;
    }
    const bool ns = ComputeNoShadow();
    const ScatteringRange lowSphere = CoarseSphere(m_sphere, nLow, nZenLow, *m_incidentLight);
    Arr2D coarseM(nLow + 1, nZenLow + 1, 4, 4);
    coarseM.ClearArr();
    Arr2D coarseMns(ns ? nLow + 1 : 0, ns ? nZenLow + 1 : 0, ns ? 4 : 0, ns ? 4 : 0);
    if (ns)
        coarseMns.ClearArr();
    const ScatteringRange saved = m_sphere;
    m_sphere = lowSphere;
    const bool ok = HandleOrientationsToLocalGpu(prepared, start, count, coarseM, coarseMns,
                                                 scale, waveIndex);
    m_sphere = saved;
    if (!ok)
        return false;
    // theta first (to the handler's rows), then phi
    Arr2D lowM(nLow + 1, nZen + 1, 4, 4), lowMns(nLow + 1, nZen + 1, 4, 4);
    if (thetaFactor > 1)
    {
        lowM.ClearArr();
        UpsampleTheta(coarseM, nLow, nZenLow, nZen, lowM);
        if (ns)
        {
            lowMns.ClearArr();
            UpsampleTheta(coarseMns, nLow, nZenLow, nZen, lowMns);
        }
    }
    else
    {
        lowM = coarseM;
        if (ns)
            lowMns = coarseMns;
    }
    if (!doCheck && !adaptiveRows)
    {
        UpsamplePhi(lowM, nLow, nFull, nZen, localM);
        if (ns)
            UpsamplePhi(lowMns, nLow, nFull, nZen, localM_noshadow);
        return true;
    }
    // validation (:3745-3782): a direct run on min(nFull, 2 nLow) azimuths,
    // upsampled the same way, estimates the error; above the tolerance it
    // replaces the interpolated grids
    Arr2D fftM(nFull + 1, nZen + 1, 4, 4), fftMns(nFull + 1, nZen + 1, 4, 4);
    fftM.ClearArr();
    fftMns.ClearArr();
    UpsamplePhi(lowM, nLow, nFull, nZen, fftM);
    if (ns)
        UpsamplePhi(lowMns, nLow, nFull, nZen, fftMns);
    if (doCheck)
    {
        const int nCheck = std::min(nFull, nLow * 2);
        ScatteringRange checkSphere = saved;
        checkSphere.nAzimuth = nCheck;
        checkSphere.azinuthStep = 2.0 * M_PI / nCheck;
        checkSphere.ComputeSphereDirections(*m_incidentLight);
        Arr2D checkLowM(nCheck + 1, nZen + 1, 4, 4), checkLowMns(nCheck + 1, nZen + 1, 4, 4);
        checkLowM.ClearArr();
        checkLowMns.ClearArr();
        m_sphere = checkSphere;
        const bool checkOk = HandleOrientationsToLocalGpu(prepared, start, count, checkLowM,
                                                          checkLowMns, scale, waveIndex);
        m_sphere = saved;
        if (!checkOk)
            return false;
        Arr2D checkM(nFull + 1, nZen + 1, 4, 4), checkMns(nFull + 1, nZen + 1, 4, 4);
        checkM.ClearArr();
        checkMns.ClearArr();
        UpsamplePhi(checkLowM, nCheck, nFull, nZen, checkM);
        if (ns)
            UpsamplePhi(checkLowMns, nCheck, nFull, nZen, checkMns);
        double err = ReportFftCheck(fftM, checkM, nFull, nZen, "full");
        if (ns)
            err = std::max(err, ReportFftCheck(fftMns, checkMns, nFull, nZen, "no-shadow"));
        const bool refine = m_fftTolerance > 0.0 && err > m_fftTolerance;
        m_fftStatistics->RecordValidation(err, refine);
        if (refine)
        {
            std::fprintf(stderr,
                         "GPU FFT refinement: estimated error %.6g above tolerance %.6g, using the "
                         "direct Nphi=%d grid\n",
                         err, m_fftTolerance, nCheck);
            fftM = checkM;
            if (ns)
                fftMns = checkMns;
        }
    }
    // adaptive rows (:3783-3832, MBS_FFT_ADAPTIVE_PHI): theta rows whose
    // averaged polarisation ratios spike against their neighbours are
    // computed directly on all nFull azimuths and replace the interpolated
    // rows
    if (adaptiveRows)
    {
        const std::vector<int> rows =
            SpikeRows(fftM, nFull, nZen, EnvDouble("MBS_FFT_ADAPTIVE_PHI_THRESHOLD", 0.04),
                      EnvInt("MBS_FFT_ADAPTIVE_PHI_MAX_ROWS", 0, 128, 8));
        if (!rows.empty())
        {
            ScatteringRange rowSphere = saved;
            rowSphere.isNonUniform = true;
            rowSphere.thetaValues.clear();
            for (int t : rows)
                rowSphere.thetaValues.push_back(saved.GetZenith(t));
            rowSphere.nAzimuth = nFull;
            rowSphere.azinuthStep = 2.0 * M_PI / nFull;
            rowSphere.nZenith = (int)rows.size() - 1;
            rowSphere.zenithStart = rowSphere.thetaValues.front();
            rowSphere.zenithEnd = rowSphere.thetaValues.back();
            rowSphere.zenithStep = rowSphere.nZenith > 0
                                       ? (rowSphere.zenithEnd - rowSphere.zenithStart) /
                                             rowSphere.nZenith
                                       : 0.0;
            rowSphere.ComputeSphereDirections(*m_incidentLight);
            Arr2D directM(nFull + 1, (int)rows.size(), 4, 4);
            Arr2D directMns(nFull + 1, (int)rows.size(), 4, 4);
            directM.ClearArr();
            directMns.ClearArr();
            m_sphere = rowSphere;
            const bool rowsOk = HandleOrientationsToLocalGpu(prepared, start, count, directM,
                                                             directMns, scale, waveIndex);
            m_sphere = saved;
            if (!rowsOk)
                return false;
            for (size_t i = 0; i < rows.size(); ++i)
                for (int p = 0; p < nFull; ++p)
                {
                    std::memcpy(fftM.RawCell(p, rows[i]), directM.RawCell(p, (int)i),
                                16 * sizeof(double));
                    if (ns)
                        std::memcpy(fftMns.RawCell(p, rows[i]), directMns.RawCell(p, (int)i),
                                    16 * sizeof(double));
                }
            std::fprintf(stderr, "GPU FFT adaptive phi: %zu of %d theta rows computed directly\n",
                         rows.size(), nZen + 1);
        }
    }
    AddCells(fftM, nFull, nZen, localM);
    if (ns)
        AddCells(fftMns, nFull, nZen, localM_noshadow);
    return true;
}

// The multi-size path's fused call (src/scattering/TracerPOTotal.cpp:4380-4432):
// one traced beam set, diffracted for every scale s into localMs[s] (full
// output only; the caller uses it without no-shadow output).  The beam set
// is packed and uploaded once; each scale is re-derived on the device by the
// rule of HandleOrientationsToLocalGpu(..., scale, waveIndex)'s pack, as the
// reference's fused kernel reuses one upload (GpuDiffraction.cu:2332-2410).
bool HandlerPO::HandleOrientationsToLocalGpuMultiK(const std::vector<PreparedOrientation> &prepared,
                                                   int start, int count,
                                                   const std::vector<double> &scales,
                                                   double waveIndex, std::vector<Arr2D> &localMs)
{
    if (!isCoh)
        return false;
    if (start < 0 || count < 0 || start > (int)prepared.size() || scales.empty()
        || localMs.size() < scales.size())
        return false;
    const int nOrient = std::min(count, (int)prepared.size() - start);
    if (nOrient == 0 || m_sphere.nAzimuth * (m_sphere.nZenith + 1) == 0)
        return true;
    GridSpec g = MakeGrid(m_sphere);
    bool ok = false;
    ThreadContexts::Entry *e = Context(0, g, ok);
    if (!e)
        return ok;
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    PackedBatch &pk = e->pk;
    // one pack and upload at scale 1 with the scaling data; every scale is
    // then re-derived on the device (mbs_rescale_beams) by the rule of the
    // per-scale pack (PackScale, ScaledJones)
    const auto tp = std::chrono::steady_clock::now();
    if (!Pack(prepared, start, nOrient, PackScale(), pk, true))
        return false;
    const double packMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
    if (pk.nv.empty())
        return true;
    const mbs_beam_batch view = pk.View();
    mbs_beam_scaling bsc;
    bsc.raw_jones = pk.rawJones.data();
    bsc.path = pk.path.data();
    bsc.sign = pk.sign.data();
    bsc.abs_offsets = pk.absOffsets.data();
    bsc.abs_paths = pk.absPaths.data();
    bsc.c_abs = AbsorptionCoefficient();
    bsc.wave_index = waveIndex;
    int rc = mbs_upload_beams(e->ctx, &view, &c);
    if (rc == MBS_OK)
        rc = mbs_upload_beam_scaling(e->ctx, &bsc);
    for (size_t s = 0; s < scales.size() && rc == MBS_OK; ++s)
    {
        rc = mbs_rescale_beams(e->ctx, scales[s]);
        if (rc == MBS_OK && !RunBatch(*e, pk, c, localMs[s], nullptr, s == 0 ? packMs : 0.0,
                                      false))
            return false;
    }
    if (rc == MBS_ERR_DEVICE)
        return DeviceFailure("rescaled batch", mbs_last_error(e->ctx));
    if (rc != MBS_OK)
    {
        std::fprintf(stderr, "HIP diffraction backend: %s\n", mbs_last_error(e->ctx));
        return false;
    }
    return true;
}

// The multi-size form of the phi-FFT path (src/cuda/GpuDiffraction.cu, the
// member after :3839): only with MBS_GPU_MULTI_K_FULL=1 and full output, as
// in the reference; the fused multi-size call on the coarse azimuth grid,
// then every scale's grid upsampled (theta, then phi) into localMs[s].  As
// in the reference, its validation and adaptive options return false.
bool HandlerPO::HandleOrientationsToLocalGpuFftPhiMultiK(
    const std::vector<PreparedOrientation> &prepared, int start, int count,
    const std::vector<double> &scales, double waveIndex, std::vector<Arr2D> &localMs)
{
    const char *fullK = std::getenv("MBS_GPU_MULTI_K_FULL");
    if (!(fullK && std::strcmp(fullK, "1") == 0))
        return false;
    if (!isCoh || ComputeNoShadow())
        return false;
    const char *check = std::getenv("MBS_FFT_CHECK");
    const char *adaptive = std::getenv("MBS_FFT_ADAPTIVE_PHI");
    if ((check && std::strcmp(check, "1") == 0) || m_fftTolerance > 0.0
        || (adaptive && *adaptive && std::strcmp(adaptive, "0") != 0))
        return false;
    const int nFull = m_sphere.nAzimuth;
    const int nZen = m_sphere.nZenith;
    const int factor = ChoosePhiFactor(nFull, m_fftPhiFactor);
    int nLow = factor > 1 ? nFull / factor : nFull;
    if (nLow < 16)
        nLow = std::min(nFull, 16);
    if (factor <= 1 || nFull < 32 || nLow >= nFull)
    {
        m_fftStatistics->RecordCall(nFull, nFull, 1);
        return HandleOrientationsToLocalGpuMultiK(prepared, start, count, scales, waveIndex,
                                                  localMs);
    }
    if (localMs.size() < scales.size())
        return false;
    m_fftStatistics->RecordCall(nFull, nLow, factor);
    const int thetaFactor = ThetaFactor(m_sphere);
    const int nZenLow = thetaFactor > 1 ? std::max(2, nZen / thetaFactor) : nZen;
    const ScatteringRange lowSphere = CoarseSphere(m_sphere, nLow, nZenLow, *m_incidentLight);
    std::vector<Arr2D> lowMs;
    lowMs.reserve(scales.size());
    for (size_t s = 0; s < scales.size(); ++s)
    {
        lowMs.push_back(Arr2D(nLow + 1, nZenLow + 1, 4, 4));
        lowMs.back().ClearArr();
    }
    const ScatteringRange saved = m_sphere;
    m_sphere = lowSphere;
    const bool ok = HandleOrientationsToLocalGpuMultiK(prepared, start, count, scales,
                                                       waveIndex, lowMs);
    m_sphere = saved;
    if (!ok)
        return false;
    for (size_t s = 0; s < scales.size(); ++s)
    {
        if (thetaFactor > 1)
        {
            Arr2D thetaM(nLow + 1, nZen + 1, 4, 4);
            thetaM.ClearArr();
            UpsampleTheta(lowMs[s], nLow, nZenLow, nZen, thetaM);
            UpsamplePhi(thetaM, nLow, nFull, nZen, localMs[s]);
        }
        else
            UpsamplePhi(lowMs[s], nLow, nFull, nZen, localMs[s]);
    }
    return true;
}

// The reference's theta probe (src/cuda/GpuDiffraction.cu:4780-4965, kernel
// theta_m11_kernel :1909-1957): for every theta of theta_rads, the M11 of
// each single beam (beams do not interfere), weighted by its orientation's
// weight, summed over the beams and the handler's nAz azimuths p * azimuth
// step, divided by nAz + 1.  Theta within 1e-14 of a pole is moved 1e-6 rad
// inside (:4845-4850); vf = (-sin phi, cos phi, 0) except at the poles
// (:4862-4876).  The reference's GPU rotation leaves out the 1/2 of the CPU
// one (its Mueller kernels weight by 1/4 instead), but the probe sums
// 0.5 |d|^2 unscaled: its values are 4 times the M11 of the CPU convention,
// reproduced so.  Evaluated by the engine's FP64 incoherent sum
// (MBS_INCOHERENT) on a probe grid; the context's next call uploads its own
// grid again.  Beams with fewer than 3 vertices, which the reference probe
// would take, are refused like every other entry point here.
bool HandlerPO::DiffractThetasGpu(const std::vector<PreparedOrientation> &prepared,
                                  const double *theta_rads, int nPoints,
                                  std::vector<double> &m11_out)
{
    if (nPoints <= 0)
        return true;
    const int nOrient = (int)prepared.size();
    const int nAz = m_sphere.nAzimuth;
    if (nOrient == 0 || nAz <= 0)
        return true;
    size_t nBeams = 0;
    for (const PreparedOrientation &po : prepared)
        for (const PreparedBeam &pb : po.beams)
        {
            if (!pb.edgeData.valid || pb.edgeData.nVertices <= 0 || pb.edgeData.nVertices > 32)
                return false;
            ++nBeams;
        }
    m11_out.assign(nPoints, 0.0);
    if (nBeams == 0)
        return true;

    GridSpec g;
    g.nAz = nAz;
    g.nZen = nPoints - 1;
    g.azStep = m_sphere.azinuthStep;
    g.theta.resize(nPoints);
    g.vf.resize((size_t)nAz * nPoints * 3);
    for (int t = 0; t < nPoints; ++t)
    {
        double theta = theta_rads[t];
        if (theta <= 1e-14)
            theta = 1e-6;
        else if (theta >= M_PI - 1e-14)
            theta = M_PI - 1e-6;
        g.theta[t] = theta;
    }
    for (int p = 0; p < nAz; ++p)
    {
        const double phi = p * g.azStep;
        for (int t = 0; t < nPoints; ++t)
        {
            const double dz = -std::cos(g.theta[t]);
            double *w = &g.vf[((size_t)p * nPoints + t) * 3];
            w[0] = -std::sin(phi);
            w[1] = std::cos(phi);
            w[2] = 0.0;
            if (dz >= 1.0 - 1e-15 || dz <= -1.0 + 1e-15)
            {
                w[0] = 0.0;
                w[1] = dz > 0.0 ? -1.0 : 1.0;
            }
        }
    }
    bool ok = false;
    ThreadContexts::Entry *e = Context(0, g, ok);
    if (!e)
        return ok;
    PackedBatch &pk = e->pk;
    if (!Pack(prepared, 0, nOrient, PackScale(), pk))
        return false;
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    const mbs_beam_batch view = pk.View();
    int rc = mbs_upload_beams(e->ctx, &view, &c);
    if (rc == MBS_OK)
        rc = mbs_diffract(e->ctx, MBS_INCOHERENT);
    std::vector<double> m((size_t)nAz * nPoints * 16, 0.0);
    if (rc == MBS_OK)
        rc = mbs_accumulate_mueller(e->ctx, m.data(), nullptr);
    if (rc == MBS_ERR_DEVICE)
        return DeviceFailure("theta probe", mbs_last_error(e->ctx));
    if (rc != MBS_OK)
    {
        std::fprintf(stderr, "HIP diffraction backend: %s\n", mbs_last_error(e->ctx));
        return false;
    }
    for (int t = 0; t < nPoints; ++t)
    {
        double sum = 0.0;
        for (int p = 0; p < nAz; ++p)
            sum += m[((size_t)p * nPoints + t) * 16];
        m11_out[t] = 4.0 * sum / (nAz + 1);
    }
    return true;
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
