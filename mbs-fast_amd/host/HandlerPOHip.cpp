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
        // the resident batch's pack-cache key (token 0: not reusable), the
        // reference's packed-beam cache (GpuDiffraction.cu:4274-4282,
        // :4588-4601): the same token, orientation range, scale and wave
        // index mean the same packed beams
        struct CacheKey
        {
            unsigned long long token = 0;
            int start = 0, count = 0;
            double scale = 0.0, waveIndex = 0.0;
            bool operator==(const CacheKey &o) const
            {
                return token == o.token && start == o.start && count == o.count
                       && scale == o.scale && waveIndex == o.waveIndex;
            }
        } resident;
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
// PreparedBeam (src/handler/HandlerPO.h:14-46) -> structure of arrays, in
// two OpenMP-parallel passes as the reference's packing is (:4407-4428):
// Count (beam, vertex and absorption-path counts per orientation, which
// also validates every beam), then, after prefix sums, Fill (every
// orientation writes its own slice of the output arrays).
struct PackCounts
{
    std::vector<int64_t> beamBase, vertBase, absBase;  // [nOrient + 1] prefix sums
    int64_t Beams() const { return beamBase.back(); }
    int64_t Verts() const { return vertBase.back(); }
};

// Where Fill writes: the arrays of an mbs_beam_batch (and, for a device
// rescale, of mbs_beam_scaling), for orientations [0, nOrient).
struct PackTarget
{
    int64_t *orientOffsets, *vertOffsets;
    double *weights, *scalars, *vx, *vy, *syx, *sxy;
    int32_t *nv, *external;
    uint8_t *flags;
    // optional (null when not wanted): the per-beam data of a device rescale
    double *rawJones, *path, *sign, *absPaths;
    int64_t *absOffsets;
};

// Returns false for a beam the device path cannot take.
bool Count(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
           PackCounts &pc)
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
    pc.beamBase.assign(nOrient + 1, 0);
    pc.vertBase.assign(nOrient + 1, 0);
    pc.absBase.assign(nOrient + 1, 0);
    for (int i = 0; i < nOrient; ++i)
    {
        pc.beamBase[i + 1] = pc.beamBase[i] + nb[i];
        pc.vertBase[i + 1] = pc.vertBase[i] + nvert[i];
        pc.absBase[i + 1] = pc.absBase[i] + nabs[i];
    }
    return true;
}

// Fills t for prepared[start, start + nOrient), whose counts are pc.
void Fill(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
          const PackCounts &pc, const PackScale &sc, const PackTarget &t)
{
    const bool scaling = t.rawJones != nullptr;
    for (int i = 0; i <= nOrient; ++i)
        t.orientOffsets[i] = pc.beamBase[i];
    t.vertOffsets[pc.Beams()] = pc.Verts();
    if (scaling)
        t.absOffsets[pc.Beams()] = pc.absBase[nOrient];
    const double g = sc.on ? sc.scale : 1.0, g2 = g * g;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nOrient; ++i)
    {
        const PreparedOrientation &po = prepared[start + i];
        t.weights[i] = po.sinZenith;
        size_t bi = (size_t)pc.beamBase[i];
        size_t vi = (size_t)pc.vertBase[i];
        size_t ai = (size_t)pc.absBase[i];
        for (const PreparedBeam &pb : po.beams)
        {
            const BeamEdgeData &ed = pb.edgeData;
            if (scaling)
            {
                const ::complex raw[4] = {pb.origBeam.J.m11, pb.origBeam.J.m12,
                                          pb.origBeam.J.m21, pb.origBeam.J.m22};
                for (int k = 0; k < 4; ++k)
                {
                    t.rawJones[bi * 8 + 2 * k] = real(raw[k]);
                    t.rawJones[bi * 8 + 2 * k + 1] = imag(raw[k]);
                }
                t.path[bi] = pb.isExternal ? pb.origBeam.opticalPath : pb.info.projLenght;
                double sg = pb.isExternal ? -1.0 : 1.0;
                if (!pb.isExternal && (pb.origBeam.nActs & 1))
                    sg = -sg;
                t.sign[bi] = sg;
                t.absOffsets[bi] = (int64_t)ai;
                for (double p : pb.absorptionPaths)
                    t.absPaths[ai++] = p;
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
            std::memcpy(&t.scalars[bi * MBS_BEAM_SCALARS], s, sizeof s);
            t.nv[bi] = ed.nVertices;
            t.external[bi] = pb.isExternal ? 1 : 0;
            t.vertOffsets[bi] = (int64_t)vi;
            for (int v = 0; v < ed.nVertices; ++v, ++vi)
            {
                t.vx[vi] = ed.x[v] * g;
                t.vy[vi] = ed.y[v] * g;
                t.syx[vi] = ed.edge_valid_x[v] ? ed.slope_yx[v] : 0.0;
                t.sxy[vi] = ed.edge_valid_y[v] ? ed.slope_xy[v] : 0.0;
                t.flags[vi] = (uint8_t)((ed.edge_valid_x[v] ? MBS_EDGE_VALID_X : 0)
                                        | (ed.edge_valid_y[v] ? MBS_EDGE_VALID_Y : 0));
            }
            ++bi;
        }
    }
}

// Count and Fill into pk's host vectors.  With `scaling` the per-beam data
// of a device rescale (mbs_beam_scaling) are packed too.
bool Pack(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
          const PackScale &sc, PackedBatch &pk, bool scaling = false)
{
    PackCounts pc;
    if (!Count(prepared, start, nOrient, pc))
        return false;
    const size_t nBeams = (size_t)pc.Beams(), nVerts = (size_t)pc.Verts();
    pk.orientOffsets.resize(nOrient + 1);
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
    PackTarget t = {pk.orientOffsets.data(), pk.vertOffsets.data(), pk.weights.data(),
                    pk.scalars.data(), pk.vx.data(), pk.vy.data(), pk.syx.data(),
                    pk.sxy.data(), pk.nv.data(), pk.external.data(), pk.flags.data(),
                    nullptr, nullptr, nullptr, nullptr, nullptr};
    if (scaling)
    {
        pk.rawJones.resize(nBeams * 8);
        pk.path.resize(nBeams);
        pk.sign.resize(nBeams);
        pk.absOffsets.resize(nBeams + 1);
        pk.absPaths.resize((size_t)pc.absBase[nOrient]);
        t.rawJones = pk.rawJones.data();
        t.path = pk.path.data();
        t.sign = pk.sign.data();
        t.absOffsets = pk.absOffsets.data();
        t.absPaths = pk.absPaths.data();
    }
    Fill(prepared, start, nOrient, pc, sc, t);
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

// An engine status under the reference's conventions: a device error exits
// (or, with MBS_GPU_ALLOW_FALLBACK=1, returns false); any other failure is
// reported and returns false.
bool Status(int rc, const char *what, mbs_ctx *ctx)
{
    if (rc == MBS_OK)
        return true;
    if (rc == MBS_ERR_DEVICE)
        return DeviceFailure(what, mbs_last_error(ctx));
    std::fprintf(stderr, "HIP diffraction backend: %s\n", mbs_last_error(ctx));
    return false;
}

// Uploads (unless `upload` is false: the batch is resident) and diffracts
// one packed batch on the context's grid; the result stays on the device.
// MBS_GPU_TIMING=1 prints the stage times.
bool DiffractBatch(ThreadContexts::Entry &e, const PackedBatch &pk, const mbs_consts &c,
                   bool ns, double packMs, bool upload = true)
{
    const bool timing = EnvIs("MBS_GPU_TIMING", "1");
    const auto t0 = std::chrono::steady_clock::now();
    mbs_set_timing(e.ctx, timing ? 1 : 0);
    const mbs_beam_batch view = pk.View();
    int rc = upload ? mbs_upload_beams(e.ctx, &view, &c) : MBS_OK;
    if (rc == MBS_OK)
        rc = mbs_diffract(e.ctx, PrecisionFlag() | (ns ? MBS_NOSHADOW : 0u));
    if (!Status(rc, "diffraction", e.ctx))
        return false;
    if (timing)
    {
        mbs_timing tm;
        if (mbs_get_timing(e.ctx, &tm) == MBS_OK)
        {
            const double callMs = std::chrono::duration<double, std::milli>(
                                      std::chrono::steady_clock::now() - t0)
                                      .count();
            std::fprintf(stderr,
                         "GPU timing: %d orientations, %zu beams: pack %.3f ms, upload %.3f ms, "
                         "kernel %.3f ms, reduce %.3f ms, engine calls %.3f ms (%d work items, "
                         "%d workgroups)\n",
                         (int)pk.weights.size(), pk.nv.size(), packMs, tm.upload_ms,
                         tm.kernel_ms, tm.reduce_ms, callMs, tm.n_items, tm.n_blocks);
        }
    }
    return true;
}

// Adds a flat grid [nAz][nZen+1][16] into the Arr2D cells p < nAz, t <= nZen.
void AddGrid(const std::vector<double> &g, int nAz, int nZen, Arr2D &dst)
{
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const double *src = &g[((size_t)p * (nZen + 1) + t) * 16];
            double *cell = dst.RawCell(p, t);
            for (int k = 0; k < 16; ++k)
                cell[k] += src[k];
        }
}

// Adds the device result of the context's last diffraction into localM (and
// *localMns when it is given).
bool AddResult(ThreadContexts::Entry &e, Arr2D &localM, Arr2D *localMns)
{
    const size_t n = (size_t)e.nAz * (e.nZen + 1) * 16;
    std::vector<double> mFull(n, 0.0), mNs(localMns ? n : 0, 0.0);
    if (!Status(mbs_accumulate_mueller(e.ctx, mFull.data(), localMns ? mNs.data() : nullptr),
                "result copy", e.ctx))
        return false;
    AddGrid(mFull, e.nAz, e.nZen, localM);
    if (localMns)
        AddGrid(mNs, e.nAz, e.nZen, *localMns);
    return true;
}

// Diffracts one packed batch and adds the result into localM (and into
// *localMns when it is given).
bool RunBatch(ThreadContexts::Entry &e, const PackedBatch &pk, const mbs_consts &c,
              Arr2D &localM, Arr2D *localMns, double packMs, bool upload = true)
{
    return DiffractBatch(e, pk, c, localMns != nullptr, packMs, upload)
           && AddResult(e, localM, localMns);
}
// Beams per part of a pipelined batch: a batch of more is cut into parts
// of whole orientations, so the host fills part i + 1 into page-locked
// staging while the device copies and diffracts part i.
const int64_t kPartBeams = 65536;

// Diffracts prepared[start, start + nOrient) (counts pc) in parts through
// the engine's page-locked staging (mbs_stage_beams / mbs_upload_staged:
// asynchronous DMA) and adds every part's result into localM (*localMns).
bool RunStaged(ThreadContexts::Entry &e, const std::vector<PreparedOrientation> &prepared,
               int start, int nOrient, const PackCounts &pc, const PackScale &sc,
               const mbs_consts &c, Arr2D &localM, Arr2D *localMns)
{
    const bool timing = EnvIs("MBS_GPU_TIMING", "1");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<int> cuts(1, 0);
    for (int o = 1; o <= nOrient; ++o)
        if (o == nOrient || pc.beamBase[o] - pc.beamBase[cuts.back()] >= kPartBeams)
            cuts.push_back(o);
    double fillMs = 0.0;
    for (size_t part = 0; part + 1 < cuts.size(); ++part)
    {
        const int o0 = cuts[part], n = cuts[part + 1] - o0;
        PackCounts sub;
        sub.beamBase.resize(n + 1);
        sub.vertBase.resize(n + 1);
        sub.absBase.resize(n + 1);
        for (int i = 0; i <= n; ++i)
        {
            sub.beamBase[i] = pc.beamBase[o0 + i] - pc.beamBase[o0];
            sub.vertBase[i] = pc.vertBase[o0 + i] - pc.vertBase[o0];
            sub.absBase[i] = pc.absBase[o0 + i] - pc.absBase[o0];
        }
        mbs_beam_staging st;
        if (!Status(mbs_stage_beams(e.ctx, n, (int32_t)sub.Beams(), sub.Verts(), &st), "staging",
                    e.ctx))
            return false;
        const auto tf = std::chrono::steady_clock::now();
        const PackTarget t = {st.orient_offsets, st.vert_offsets, st.weights, st.beam_scalars,
                              st.vx, st.vy, st.slope_yx, st.slope_xy, st.beam_nv,
                              st.beam_external, st.edge_flags,
                              nullptr, nullptr, nullptr, nullptr, nullptr};
        Fill(prepared, start + o0, n, sub, sc, t);
        fillMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tf)
                      .count();
        // the previous part's result, while this part waits to be copied
        if (part > 0 && !AddResult(e, localM, localMns))
            return false;
        int rc = mbs_upload_staged(e.ctx, &c);
        if (rc == MBS_OK)
            rc = mbs_diffract(e.ctx, PrecisionFlag() | (localMns ? MBS_NOSHADOW : 0u));
        if (!Status(rc, "diffraction", e.ctx))
            return false;
    }
    if (!AddResult(e, localM, localMns))
        return false;
    if (timing)
        std::fprintf(stderr,
                     "GPU timing: %d orientations, %lld beams in %zu part(s): fill %.3f ms, call "
                     "%.3f ms (page-locked staging, asynchronous upload)\n",
                     nOrient, (long long)pc.Beams(), cuts.size() - 1, fillMs,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                         .count());
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
    unsigned long long cacheToken)
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
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    Arr2D *ns = ComputeNoShadow() ? &localM_noshadow : nullptr;
    ThreadContexts::Entry::CacheKey key;
    key.token = cacheToken;
    key.start = start;
    key.count = nOrient;
    key.scale = scale;
    key.waveIndex = waveIndex;
    // the batch this token names is resident: diffract it on this grid
    if (cacheToken != 0 && e->resident == key)
        return pk.nv.empty() || RunBatch(*e, pk, c, localM, ns, 0.0, false);
    e->resident = ThreadContexts::Entry::CacheKey();
    const PackScale sc(scale, waveIndex, AbsorptionCoefficient());
    if (cacheToken == 0)
    {
        // no reuse asked for: pipelined through page-locked staging
        PackCounts pc;
        if (!Count(prepared, start, nOrient, pc))
            return false;
        return pc.Beams() == 0 || RunStaged(*e, prepared, start, nOrient, pc, sc, c, localM, ns);
    }
    // a batch to keep: packed whole on the host (the cache holds pk)
    const auto tp = std::chrono::steady_clock::now();
    if (!Pack(prepared, start, nOrient, sc, pk))
        return false;
    const double packMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
    if (!pk.nv.empty() && !RunBatch(*e, pk, c, localM, ns, packMs))
        return false;
    e->resident = key;
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

double EnvDouble(const char *name, double fallback)
{
    const char *v = std::getenv(name);
    if (!v || !*v)
        return fallback;
    char *end = nullptr;
    const double x = std::strtod(v, &end);
    return (end && *end == '\0' && x > 0.0) ? x : fallback;
}

// Longest line the device upsampler takes (mbs_upsample_result).
const int kMaxUpsampleLine = 4096;

// The azimuth factor of the angular FFT path: the handler's own, else
// MBS_FFT_PHI_FACTOR (1..64; unset or "auto" chooses), else by the full
// azimuth count with the reference's defaults (gpu_fft_phi_factor and
// choose_fft_phi_factor, src/cuda/GpuDiffraction.cu:513-531, :712-727).
int PhiFactor(int nFull, int configured)
{
    static const struct
    {
        int minAzimuths, factor;
    } kDefaults[] = {{2400, 8}, {600, 6}, {240, 3}};
    if (configured > 0)
        return configured;
    const int fromEnv = EnvInt("MBS_FFT_PHI_FACTOR", 1, 64, 0);
    if (fromEnv > 0)
        return fromEnv;
    for (const auto &d : kDefaults)
        if (nFull >= d.minAzimuths)
            return d.factor;
    return 2;
}

// The zenith factor (MBS_FFT_THETA_FACTOR, gpu_fft_theta_factor :532-542 and
// :3543-3550): 1 unless the zenith grid is uniform, has at least 32 rows and
// the factor divides them -- and, here, unless the device upsampler takes
// the extended zenith lines.
int ThetaFactor(const ScatteringRange &sphere)
{
    const int f = EnvInt("MBS_FFT_THETA_FACTOR", 1, 16, 1);
    if (f <= 1 || sphere.isNonUniform || sphere.nZenith < 32 || sphere.nZenith % f != 0
        || 2 * sphere.nZenith > kMaxUpsampleLine)
        return 1;
    return f;
}

// The coarse-grid sphere of the path (:3563-3570): nLow azimuths and nZenLow
// zenith rows over the same range.
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

// Packs prepared[start, start + nOrient) and diffracts it on `sphere` with
// this thread's context on device 0; the result stays on the device.  e is
// left null when there is nothing to diffract (no beams).
bool DiffractResident(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
                      const ScatteringRange &sphere, const PackScale &sc, const mbs_consts &c,
                      bool ns, ThreadContexts::Entry *&e)
{
    e = nullptr;
    GridSpec g = MakeGrid(sphere);
    bool ok = false;
    ThreadContexts::Entry *ctx = Context(0, g, ok);
    if (!ctx)
        return ok;
    ctx->resident = ThreadContexts::Entry::CacheKey();
    const auto tp = std::chrono::steady_clock::now();
    if (!Pack(prepared, start, nOrient, sc, ctx->pk))
        return false;
    const double packMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
    if (ctx->pk.nv.empty())
        return true;
    if (!DiffractBatch(*ctx, ctx->pk, c, ns, packMs))
        return false;
    e = ctx;
    return true;
}

// The batch resident on e diffracted again, on another grid.
bool Rediffract(ThreadContexts::Entry &e, const ScatteringRange &sphere, const mbs_consts &c,
                bool ns)
{
    GridSpec g = MakeGrid(sphere);
    bool ok = false;
    if (!Context(0, g, ok))
        return ok;
    return DiffractBatch(e, e.pk, c, ns, 0.0, false);
}

// The last result on e, Fourier-upsampled on the device to the grid of
// `full` (its azimuths and zenith rows): grid [nAz][nZen+1][16] and, when
// rows is given, the grid's azimuth-averaged rows [nZen+1][18].
bool Upsampled(ThreadContexts::Entry &e, bool ns, const ScatteringRange &full,
               std::vector<double> &grid, std::vector<double> *rows)
{
    const int nAz = full.nAzimuth, nZen = full.nZenith;
    std::vector<double> theta(nZen + 1);
    for (int t = 0; t <= nZen; ++t)
        theta[t] = full.GetZenith(t);
    grid.assign((size_t)nAz * (nZen + 1) * 16, 0.0);
    if (rows)
        rows->assign((size_t)(nZen + 1) * 18, 0.0);
    return Status(mbs_upsample_result(e.ctx, ns ? 1 : 0, nAz, nZen, theta.data(), 0, grid.data(),
                                      rows ? rows->data() : nullptr),
                  "upsampling", e.ctx);
}

// How far an interpolated grid is from a reference grid, in the figures of
// the reference's FFT check (report_fft_check, :3142-3195), whose stderr
// line is kept as it is: M11 errors relative to the cell's own M11, over
// all cells and over the "significant" ones (M11 at least 1e-3 of the
// grid's largest), their RMS, and every element's error relative to
// max(|M11|, that floor).  The estimate is the larger of the significant-M11
// and the element figures.
struct InterpolationError
{
    double m11MaxAbs = 0.0, m11MaxRel = 0.0, m11SignificantRel = 0.0, m11Rms = 0.0;
    double elementRel = 0.0;
    int cells = 0, significantCells = 0;

    double Estimate() const { return std::max(m11SignificantRel, elementRel); }
    void Print(const char *name) const
    {
        std::fprintf(stderr,
                     "GPU FFT check %s: M11 max_abs=%.6g max_rel=%.6g significant_max_rel=%.6g "
                     "rms_rel=%.6g, Mueller/M11 max_rel=%.6g, samples=%d significant_samples=%d\n",
                     name, m11MaxAbs, m11MaxRel, m11SignificantRel, m11Rms, elementRel, cells,
                     significantCells);
    }
};

InterpolationError Compare(const std::vector<double> &got, const std::vector<double> &ref)
{
    InterpolationError r;
    const size_t n = ref.size() / 16;
    double largest = 0.0;
    for (size_t i = 0; i < n; ++i)
        largest = std::max(largest, std::fabs(ref[16 * i]));
    const double floor = 1e-3 * largest;
    double sumSq = 0.0;
    for (size_t i = 0; i < n; ++i)
    {
        const double *a = &got[16 * i], *b = &ref[16 * i];
        const double m11 = std::fabs(b[0]);
        const double d = std::fabs(a[0] - b[0]);
        const double rel = d / std::max(m11, 1e-12);
        r.m11MaxAbs = std::max(r.m11MaxAbs, d);
        r.m11MaxRel = std::max(r.m11MaxRel, rel);
        sumSq += rel * rel;
        if (m11 >= floor)
        {
            r.m11SignificantRel = std::max(r.m11SignificantRel, rel);
            ++r.significantCells;
        }
        const double den = std::max(m11, floor);
        if (den > 0.0)
            for (int k = 0; k < 16; ++k)
                r.elementRel = std::max(r.elementRel, std::fabs(a[k] - b[k]) / den);
    }
    r.cells = (int)n;
    r.m11Rms = std::sqrt(sumSq / std::max<size_t>(1, n));
    return r;
}

// The rows the adaptive option recomputes directly (the reference's
// MBS_FFT_ADAPTIVE_PHI rule, detect_fft_phi_spike_rows :3086-3135): interior
// zenith rows whose azimuth-averaged ratios M_ij/M11 for ij = 23, 24, 32, 34,
// 42, 43, 13, 14 leave the mean of the neighbouring rows' ratios by at least
// `threshold`; the maxRows highest scores, in row order.  rows: the grid's
// azimuth-averaged rows [nZen+1][18] (theta, 2 pi dcos, M11..M44), from the
// device; every row's ratios are formed once.
std::vector<int> RowsToRecompute(const std::vector<double> &rows, int nZen, double threshold,
                                 int maxRows)
{
    static const int kElements[8] = {6, 7, 9, 11, 13, 14, 2, 3};
    std::vector<int> picked;
    if (maxRows <= 0 || nZen < 2)
        return picked;
    std::vector<double> ratio((size_t)(nZen + 1) * 8);
    for (int t = 0; t <= nZen; ++t)
    {
        const double *m = &rows[(size_t)t * 18 + 2];
        const double m11 = std::max(std::fabs(m[0]), 1e-30);
        for (int i = 0; i < 8; ++i)
            ratio[(size_t)t * 8 + i] = m[kElements[i]] / m11;
    }
    std::vector<std::pair<double, int>> scored;
    for (int t = 1; t < nZen; ++t)
    {
        const double *lo = &ratio[(size_t)(t - 1) * 8], *mid = lo + 8, *hi = lo + 16;
        double score = 0.0;
        for (int i = 0; i < 8; ++i)
            score = std::max(score, std::fabs(mid[i] - 0.5 * (lo[i] + hi[i])));
        if (score >= threshold)
            scored.emplace_back(score, t);
    }
    const size_t keep = std::min(scored.size(), (size_t)maxRows);
    std::partial_sort(scored.begin(), scored.begin() + keep, scored.end(),
                      [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                          return a.first > b.first;
                      });
    for (size_t i = 0; i < keep; ++i)
        picked.push_back(scored[i].second);
    std::sort(picked.begin(), picked.end());
    return picked;
}

// A sphere of the handler's azimuths on the listed zenith rows only.
ScatteringRange RowSphere(const ScatteringRange &full, const std::vector<int> &rows,
                          const Light &incident)
{
    ScatteringRange s = full;
    s.isNonUniform = true;
    s.thetaValues.clear();
    for (int t : rows)
        s.thetaValues.push_back(full.GetZenith(t));
    s.nZenith = (int)rows.size() - 1;
    s.zenithStart = s.thetaValues.front();
    s.zenithEnd = s.thetaValues.back();
    s.zenithStep = s.nZenith > 0 ? (s.zenithEnd - s.zenithStart) / s.nZenith : 0.0;
    s.ComputeSphereDirections(incident);
    return s;
}
}  // namespace

// The reference's angular Fourier interpolation (src/cuda/GpuDiffraction.cu:
// 3504-3839): the batch is diffracted on a grid of nLow = nFull / factor
// azimuths (at least 16) -- and, with MBS_FFT_THETA_FACTOR on a uniform
// zenith grid, nZen / factor rows -- and the result is Fourier-upsampled on
// the device to the handler's grid (mbs_upsample_result: zenith, then
// azimuth) and added into localM (and localM_noshadow).  A factor of 1,
// nFull < 32, nLow >= nFull or an azimuth count beyond the device
// upsampler's lines runs the direct path.  MBS_FFT_CHECK=1 or a tolerance:
// the same batch is also diffracted on min(nFull, 2 nLow) azimuths and
// upsampled, the difference is reported, and above the tolerance that grid
// replaces the interpolated one.  MBS_FFT_ADAPTIVE_PHI: zenith rows whose
// polarisation ratios spike against their neighbours are diffracted
// directly on all nFull azimuths and replace the interpolated rows.  The
// batch is packed and uploaded once for all of these runs.
bool HandlerPO::HandleOrientationsToLocalGpuFftPhi(const std::vector<PreparedOrientation> &prepared,
                                                   int start, int count, Arr2D &localM,
                                                   Arr2D &localM_noshadow, double scale,
                                                   double waveIndex)
{
    if (!isCoh)
        return false;
    const int nFull = m_sphere.nAzimuth;
    const int nZen = m_sphere.nZenith;
    const int factor = PhiFactor(nFull, m_fftPhiFactor);
    int nLow = factor > 1 ? nFull / factor : nFull;
    if (nLow < 16)
        nLow = std::min(nFull, 16);
    if (factor <= 1 || nFull < 32 || nLow >= nFull || nFull > kMaxUpsampleLine)
    {
        m_fftStatistics->RecordCall(nFull, nFull, 1);
        return HandleOrientationsToLocalGpu(prepared, start, count, localM, localM_noshadow,
                                            scale, waveIndex);
    }
    if (start < 0 || count < 0 || start > (int)prepared.size())
        return false;
    const int nOrient = std::min(count, (int)prepared.size() - start);
    const char *check = std::getenv("MBS_FFT_CHECK");
    const char *adaptive = std::getenv("MBS_FFT_ADAPTIVE_PHI");
    const int thetaFactor = ThetaFactor(m_sphere);
    const int nZenLow = thetaFactor > 1 ? std::max(2, nZen / thetaFactor) : nZen;
    const bool doCheck = (check && std::strcmp(check, "1") == 0) || m_fftTolerance > 0.0;
    const bool adaptiveRows = adaptive && *adaptive && std::strcmp(adaptive, "0") != 0;
    m_fftStatistics->RecordCall(nFull, nLow, factor);
    static bool announced = false;
    if (!announced)
    {
        std::fprintf(stderr,
                     "GPU FFT angular interpolation: direct Nphi=%d, output Nphi=%d (factor=%d), "
                     "direct Ntheta=%d, output Ntheta=%d (factor=%d). "
                     "This is angular Fourier interpolation, not aperture pFFT/FMM.\n",
                     nLow, nFull, factor, nZenLow + 1, nZen + 1, thetaFactor);
        announced = true;
    }
    if (nOrient == 0)
        return true;
    const bool ns = ComputeNoShadow();
    const PackScale sc(scale, waveIndex, AbsorptionCoefficient());
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    ThreadContexts::Entry *e = nullptr;
    if (!DiffractResident(prepared, start, nOrient,
                          CoarseSphere(m_sphere, nLow, nZenLow, *m_incidentLight), sc, c, ns, e))
        return false;
    if (!e)
        return true;
    std::vector<double> grid, gridNs, rows;
    if (!Upsampled(*e, false, m_sphere, grid, adaptiveRows ? &rows : nullptr)
        || (ns && !Upsampled(*e, true, m_sphere, gridNs, nullptr)))
        return false;
    if (doCheck)
    {
        const int nCheck = std::min(nFull, nLow * 2);
        ScatteringRange checkSphere = m_sphere;
        checkSphere.nAzimuth = nCheck;
        checkSphere.azinuthStep = 2.0 * M_PI / nCheck;
        checkSphere.ComputeSphereDirections(*m_incidentLight);
        std::vector<double> check, checkNs, checkRows;
        if (!Rediffract(*e, checkSphere, c, ns)
            || !Upsampled(*e, false, m_sphere, check, adaptiveRows ? &checkRows : nullptr)
            || (ns && !Upsampled(*e, true, m_sphere, checkNs, nullptr)))
            return false;
        const InterpolationError full = Compare(grid, check);
        full.Print("full");
        double err = full.Estimate();
        if (ns)
        {
            const InterpolationError shadowless = Compare(gridNs, checkNs);
            shadowless.Print("no-shadow");
            err = std::max(err, shadowless.Estimate());
        }
        const bool refine = m_fftTolerance > 0.0 && err > m_fftTolerance;
        m_fftStatistics->RecordValidation(err, refine);
        if (refine)
        {
            std::fprintf(stderr,
                         "GPU FFT refinement: estimated error %.6g above tolerance %.6g, using the "
                         "direct Nphi=%d grid\n",
                         err, m_fftTolerance, nCheck);
            grid.swap(check);
            gridNs.swap(checkNs);
            rows.swap(checkRows);
        }
    }
    if (adaptiveRows)
    {
        const std::vector<int> redo =
            RowsToRecompute(rows, nZen, EnvDouble("MBS_FFT_ADAPTIVE_PHI_THRESHOLD", 0.04),
                            EnvInt("MBS_FFT_ADAPTIVE_PHI_MAX_ROWS", 0, 128, 8));
        if (!redo.empty())
        {
            if (!Rediffract(*e, RowSphere(m_sphere, redo, *m_incidentLight), c, ns))
                return false;
            const int nRows = (int)redo.size();
            const size_t n = (size_t)nFull * nRows * 16;
            std::vector<double> direct(n, 0.0), directNs(ns ? n : 0, 0.0);
            if (!Status(mbs_accumulate_mueller(e->ctx, direct.data(),
                                               ns ? directNs.data() : nullptr),
                        "result copy", e->ctx))
                return false;
            for (int i = 0; i < nRows; ++i)
                for (int p = 0; p < nFull; ++p)
                {
                    const size_t from = ((size_t)p * nRows + i) * 16;
                    const size_t to = ((size_t)p * (nZen + 1) + redo[i]) * 16;
                    std::copy(&direct[from], &direct[from] + 16, &grid[to]);
                    if (ns)
                        std::copy(&directNs[from], &directNs[from] + 16, &gridNs[to]);
                }
            std::fprintf(stderr, "GPU FFT adaptive phi: %d of %d theta rows computed directly\n",
                         nRows, nZen + 1);
        }
    }
    AddGrid(grid, nFull, nZen, localM);
    if (ns)
        AddGrid(gridNs, nFull, nZen, localM_noshadow);
    return true;
}

namespace
{
// The multi-size members' common part: the beam set packed and uploaded
// once at scale 1 with its scaling data, then for every scale s the device
// records re-derived (mbs_rescale_beams, the rule of PackScale and
// ScaledJones), diffracted on `sphere`, and handed to use(s, context).
template <class Use>
bool ForEachScale(const std::vector<PreparedOrientation> &prepared, int start, int nOrient,
                  const ScatteringRange &sphere, const std::vector<double> &scales,
                  double waveIndex, double cAbs, const mbs_consts &c, Use use)
{
    GridSpec g = MakeGrid(sphere);
    bool ok = false;
    ThreadContexts::Entry *e = Context(0, g, ok);
    if (!e)
        return ok;
    e->resident = ThreadContexts::Entry::CacheKey();
    PackedBatch &pk = e->pk;
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
    bsc.c_abs = cAbs;
    bsc.wave_index = waveIndex;
    int rc = mbs_upload_beams(e->ctx, &view, &c);
    if (rc == MBS_OK)
        rc = mbs_upload_beam_scaling(e->ctx, &bsc);
    if (!Status(rc, "upload", e->ctx))
        return false;
    for (size_t s = 0; s < scales.size(); ++s)
    {
        if (!Status(mbs_rescale_beams(e->ctx, scales[s]), "rescaled batch", e->ctx)
            || !DiffractBatch(*e, pk, c, false, s == 0 ? packMs : 0.0, false) || !use(s, *e))
            return false;
    }
    return true;
}
}  // namespace

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
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    return ForEachScale(prepared, start, nOrient, m_sphere, scales, waveIndex,
                        AbsorptionCoefficient(), c,
                        [&](size_t s, ThreadContexts::Entry &e) {
                            return AddResult(e, localMs[s], nullptr);
                        });
}

// The multi-size form of the angular FFT path (src/cuda/GpuDiffraction.cu,
// the member after :3839): only with MBS_GPU_MULTI_K_FULL=1 and full output,
// as in the reference; every scale is diffracted on the coarse grid and
// upsampled on the device into localMs[s].  As in the reference, its
// validation and adaptive options return false.
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
    const int factor = PhiFactor(nFull, m_fftPhiFactor);
    int nLow = factor > 1 ? nFull / factor : nFull;
    if (nLow < 16)
        nLow = std::min(nFull, 16);
    if (factor <= 1 || nFull < 32 || nLow >= nFull || nFull > kMaxUpsampleLine)
    {
        m_fftStatistics->RecordCall(nFull, nFull, 1);
        return HandleOrientationsToLocalGpuMultiK(prepared, start, count, scales, waveIndex,
                                                  localMs);
    }
    if (start < 0 || count < 0 || start > (int)prepared.size() || localMs.size() < scales.size())
        return false;
    m_fftStatistics->RecordCall(nFull, nLow, factor);
    const int nOrient = std::min(count, (int)prepared.size() - start);
    if (nOrient == 0 || scales.empty())
        return true;
    const int thetaFactor = ThetaFactor(m_sphere);
    const int nZenLow = thetaFactor > 1 ? std::max(2, nZen / thetaFactor) : nZen;
    const mbs_consts c = HandlerConsts(m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                                       m_legacySign);
    const ScatteringRange &full = m_sphere;
    return ForEachScale(prepared, start, nOrient,
                        CoarseSphere(m_sphere, nLow, nZenLow, *m_incidentLight), scales,
                        waveIndex, AbsorptionCoefficient(), c,
                        [&](size_t s, ThreadContexts::Entry &e) {
                            std::vector<double> grid;
                            if (!Upsampled(e, false, full, grid, nullptr))
                                return false;
                            AddGrid(grid, nFull, nZen, localMs[s]);
                            return true;
                        });
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
    e->resident = ThreadContexts::Entry::CacheKey();
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
