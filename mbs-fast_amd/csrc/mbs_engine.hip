// mbs_engine.hip -- C ABI of the gfx950 PO diffraction engine (include/mbs_hip.h).
//
// Host side: device buffers, beam upload and packing, work decomposition and
// kernel launches.  Replaces HandlerPO::HandleOrientationsToLocalGpu and its
// helpers in src/cuda/GpuDiffraction.cu:4159-4778 (count/pack, buffer
// management, kernel selection, D2H, host add) with a design of its own:
// packing runs on the device, the orientation reduction is deterministic
// (per-item partials summed in fixed order), and work items are sized so a
// launch always has thousands of workgroups whatever the orientation count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mbs_hip.h"
#include "po_kernels.hip"
#include "orient_kernels.hip"
#include "fft_kernels.hip"
#include "mlp_kernels.hip"
#include "upsample_kernels.hip"
#include "spectrum_kernels.hip"

using namespace mbs;

namespace {

thread_local std::string g_err;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= cap)
            return hipSuccess;
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(n, 1);
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e == hipSuccess)
            cap = want;
        return e;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace

struct mbs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // grid
    bool grid_ok = false;
    int n_az = 0, n_zen1 = 0, cells = 0, tab_mode = -1;
    double az_step = 0;
    DevBuf<double> theta, vf, tab;
    DevBuf<float> tab32;
    // beams
    bool beams_ok = false;
    int n_orient = 0, n_beams = 0;
    int64_t n_vert = 0;
    std::vector<int64_t> h_ooff;
    std::vector<double> h_weights;
    int64_t nv_hist[32] = {0};
    double mean_edges = 0;
    mbs_consts consts{};
    DevBuf<double> bs_in, vx_in, vy_in, syx_in, sxy_in, weights;
    DevBuf<int32_t> nv, ext, vflag;
    DevBuf<int64_t> voff, ooff;
    DevBuf<uint8_t> ef_in;
    DevBuf<double> hdr, vtx, vdd;
    DevBuf<float> hdr32, wb;
    // size rescaling data of the batch (mbs_upload_beam_scaling)
    bool scaling_ok = false;
    double sc_cabs = 0.0, sc_k = 0.0;
    DevBuf<double> sc_raw, sc_path, sc_sign, sc_apath;
    DevBuf<int64_t> sc_aoff;
    // work and results
    DevBuf<WorkItem> items;
    DevBuf<double> part, part_ns, res, res_ns, per_orient, avg_rows, orient;
    // Fourier upsampling of a result (mbs_upsample_result)
    DevBuf<double2> up_roots_in, up_roots_out;
    DevBuf<double> up_a, up_b, up_theta;
    // aperture FFT
    DevBuf<float2> fft_grid, fft_tw;
    DevBuf<double> fft_rec, fft_ab, fft_out;
    DevBuf<int64_t> fft_beams;
    int fft_n = 0, fft_batch = 0;
    float fft_ms[3] = {0, 0, 0};
    hipEvent_t fft_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // near-forward spectrum source (mbs_spectrum_source): the selected
    // beams' slots, records, spectra and F at every cell
    bool spec_ok = false;
    int spec_n = 0;
    double pack_scale = 1.0;  // size scale of the packed beams (mbs_rescale_beams)
    DevBuf<int32_t> spec_sidx, spec_beam, spec_logn, spec_band;
    DevBuf<int64_t> spec_off;
    DevBuf<double> spec_scan, spec_h, spec_rec, spec_deap;
    DevBuf<float2> spec_grid, spec_f, spec_cph;
    std::vector<WorkItem> h_items;
    bool have_result = false, have_ns = false, have_per = false;
    double *pinned = nullptr;
    size_t pinned_cap = 0;
    // page-locked staging of uploads (mbs_stage_beams): two areas used in
    // turn, each with the event that marks the end of its copies
    struct Staging {
        unsigned char *mem = nullptr;
        size_t cap = 0;
        hipEvent_t copied = nullptr;
        bool in_flight = false;
        mbs_beam_staging view{};
    } stage[2];
    int stage_next = 0, stage_cur = -1;
    // ~65 workgroups per slot (256 CUs x 4): items small enough that the
    // static per-XCD share of workgroups carries the same work on every XCD
    int target_blocks = 65536;
    int work_mode = 0;  // 0 auto, 1 whole orientations, 2 beam segments, 3 device queue
    // device work queue: header, and the item count its plan aims for
    DevBuf<int> queue;
    int queue_want = 0;
    // launch shape chosen by mbs_autotune, per precision (0 FP64, 1 mixed)
    struct Tuned {
        bool valid = false;
        int target_blocks = 0, work_mode = 0, mfma = 1, waves = 0;  // waves 0: default
    } tuned[2];
    // graph replay of the diffraction kernels (mbs_set_graph)
    bool use_graph = false;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;
    int64_t graph_captures = 0, graph_replays = 0;
    // stage timing
    bool timing = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool ev_up = false, ev_diff = false;
    int last_items = 0, last_blocks = 0;
};

static int fail(mbs_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
    g_err = buf;
    return code;
}

#define HIP_TRY(ctx, call)                                                        \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail((ctx), MBS_ERR_DEVICE, "%s failed: %s", #call,            \
                        hipGetErrorString(e_));                                   \
    } while (0)

// Scratch memory of mbs_diffract_noalloc (work list and partial sums): one
// fixed array per device, held until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory; the device
// address of each device's copy comes from hipGetSymbolAddress, looked up
// once per device when the first context on it is created (with that device
// current), and only that address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[MBS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
struct ScratchSlot {
    bool looked_up = false;
    hipError_t err = hipSuccess;
    unsigned char *ptr = nullptr;
};
ScratchSlot g_scratch[kMaxDevices];
}  // namespace

// With `device` current: records the address of its copy of `scratch`.
static void scratch_lookup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchSlot &slot = g_scratch[device];
    if (slot.looked_up)
        return;
    void *p = nullptr;
    slot.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (slot.err != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();  // not left for a later launch check to report
    }
    slot.ptr = (unsigned char *)p;
    slot.looked_up = true;
}

static unsigned char *scratch_address(int device, hipError_t *err)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !g_scratch[device].looked_up) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    *err = g_scratch[device].err;
    return g_scratch[device].ptr;
}

extern "C" {

int mbs_abi_version(void) { return MBS_ABI_VERSION; }

const char *mbs_build_info(void)
{
    return "mbs_hip gfx950; FP64 + mixed FP32/FP64-phase kernels; "
           "deterministic orientation reduction";
}

int mbs_visible_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

int mbs_check_runtime(mbs_device_info *info)
{
    if (!info)
        return fail(nullptr, MBS_ERR_INVALID, "info is NULL");
    std::memset(info, 0, sizeof *info);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(nullptr, MBS_ERR_DEVICE, "no HIP device visible (%s)",
                    hipGetErrorString(e));
    int dev = 0;
    HIP_TRY(nullptr, hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(nullptr, hipGetDeviceProperties(&prop, dev));
    size_t fr = 0, tot = 0;
    HIP_TRY(nullptr, hipMemGetInfo(&fr, &tot));
    info->device_id = dev;
    info->visible_devices = n;
    info->total_mem = (int64_t)tot;
    info->free_mem = (int64_t)fr;
    info->compute_units = prop.multiProcessorCount;
    info->clock_khz = prop.clockRate;
    int rv = 0, dv = 0;
    (void)hipRuntimeGetVersion(&rv);
    (void)hipDriverGetVersion(&dv);
    info->runtime_version = rv;
    info->driver_version = dv;
    std::snprintf(info->name, sizeof info->name, "%s", prop.name);
    std::snprintf(info->arch, sizeof info->arch, "%s", prop.gcnArchName);
    return MBS_OK;
}

int mbs_query_memory(int64_t *free_bytes, int64_t *total_bytes)
{
    size_t fr = 0, tot = 0;
    HIP_TRY(nullptr, hipMemGetInfo(&fr, &tot));
    if (free_bytes)
        *free_bytes = (int64_t)fr;
    if (total_bytes)
        *total_bytes = (int64_t)tot;
    return MBS_OK;
}

int mbs_ctx_create(int device, mbs_ctx **out)
{
    if (!out)
        return fail(nullptr, MBS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || device < 0 || device >= n)
        return fail(nullptr, MBS_ERR_DEVICE, "HIP device %d not available (%d visible)",
                    device, n);
    HIP_TRY(nullptr, hipSetDevice(device));
    scratch_lookup(device);
    mbs_ctx *ctx = new mbs_ctx();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, MBS_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return MBS_OK;
}

int mbs_ctx_destroy(mbs_ctx *ctx)
{
    if (!ctx)
        return MBS_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->theta.release();
    ctx->vf.release();
    ctx->tab.release();
    ctx->tab32.release();
    ctx->bs_in.release();
    ctx->sc_raw.release();
    ctx->sc_path.release();
    ctx->sc_sign.release();
    ctx->sc_apath.release();
    ctx->sc_aoff.release();
    ctx->vx_in.release();
    ctx->vy_in.release();
    ctx->syx_in.release();
    ctx->sxy_in.release();
    ctx->weights.release();
    ctx->nv.release();
    ctx->ext.release();
    ctx->vflag.release();
    ctx->voff.release();
    ctx->ooff.release();
    ctx->ef_in.release();
    ctx->hdr.release();
    ctx->vtx.release();
    ctx->vdd.release();
    ctx->wb.release();
    ctx->hdr32.release();
    ctx->items.release();
    ctx->queue.release();
    ctx->part.release();
    ctx->part_ns.release();
    ctx->res.release();
    ctx->res_ns.release();
    ctx->per_orient.release();
    ctx->avg_rows.release();
    ctx->orient.release();
    if (ctx->graph_exec)
        (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->up_roots_in.release();
    ctx->up_roots_out.release();
    ctx->up_a.release();
    ctx->up_b.release();
    ctx->up_theta.release();
    ctx->fft_grid.release();
    ctx->fft_tw.release();
    ctx->fft_rec.release();
    ctx->fft_ab.release();
    ctx->fft_out.release();
    ctx->fft_beams.release();
    ctx->spec_sidx.release();
    ctx->spec_beam.release();
    ctx->spec_logn.release();
    ctx->spec_band.release();
    ctx->spec_off.release();
    ctx->spec_scan.release();
    ctx->spec_h.release();
    ctx->spec_rec.release();
    ctx->spec_deap.release();
    ctx->spec_grid.release();
    ctx->spec_f.release();
    ctx->spec_cph.release();
    for (hipEvent_t &e : ctx->fft_ev)
        if (e)
            (void)hipEventDestroy(e);
    if (ctx->pinned)
        (void)hipHostFree(ctx->pinned);
    for (auto &st : ctx->stage) {
        if (st.mem)
            (void)hipHostFree(st.mem);
        if (st.copied)
            (void)hipEventDestroy(st.copied);
    }
    for (hipEvent_t &e : ctx->ev)
        if (e)
            (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MBS_OK;
}

const char *mbs_last_error(const mbs_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_err.c_str();
}

void *mbs_ctx_stream(mbs_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int mbs_set_grid(mbs_ctx *ctx, const mbs_grid *g)
{
    if (!ctx || !g || !g->theta || !g->vf)
        return fail(ctx, MBS_ERR_INVALID, "mbs_set_grid: NULL argument");
    if (g->n_az <= 0 || g->n_zen < 0)
        return fail(ctx, MBS_ERR_INVALID, "mbs_set_grid: n_az=%d n_zen=%d", g->n_az, g->n_zen);
    const int64_t cells = (int64_t)g->n_az * (g->n_zen + 1);
    if (cells > (1 << 28))
        return fail(ctx, MBS_ERR_CAPACITY, "grid of %lld cells exceeds 2^28", (long long)cells);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    ctx->grid_ok = false;
    ctx->have_result = false;
    ctx->spec_ok = false;
    ctx->n_az = g->n_az;
    ctx->n_zen1 = g->n_zen + 1;
    ctx->cells = (int)cells;
    ctx->az_step = g->az_step;
    HIP_TRY(ctx, ctx->theta.ensure(ctx->n_zen1));
    HIP_TRY(ctx, ctx->vf.ensure((size_t)cells * 3));
    HIP_TRY(ctx, ctx->tab.ensure((size_t)cells * DT_COUNT));
    HIP_TRY(ctx, ctx->tab32.ensure((size_t)cells * DT32_COUNT));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->theta.p, g->theta, sizeof(double) * ctx->n_zen1,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->vf.p, g->vf, sizeof(double) * cells * 3,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->tab_mode = -1;
    ctx->grid_ok = true;
    return MBS_OK;
}

}  // extern "C"

static int ensure_tables(mbs_ctx *ctx)
{
    if (ctx->tab_mode == 0)
        return MBS_OK;
    const int threads = 256;
    const int blocks = (ctx->cells + threads - 1) / threads;
    grid_tables_kernel<<<blocks, threads, 0, ctx->stream>>>(
        ctx->theta.p, ctx->vf.p, ctx->n_az, ctx->n_zen1, ctx->az_step, ctx->tab.p,
        ctx->tab32.p);
    HIP_TRY(ctx, hipGetLastError());
    ctx->tab_mode = 0;
    return MBS_OK;
}

// Device records from the uploaded inputs: pack_beams_kernel on the
// context's stream, for size scale `scale` with the scaling data when
// `rescale`, else as uploaded.
static int launch_pack(mbs_ctx *ctx, int nb, const mbs_consts &c, double scale, bool rescale)
{
    // small-phase area limit: (legacy ? invComplWave : -invComplWave) * area
    const double sg = c.legacy_sign ? 1.0 : -1.0;
    pack_beams_kernel<<<(nb + 127) / 128, 128, 0, ctx->stream>>>(
        nb, ctx->bs_in.p, ctx->nv.p, ctx->ext.p, ctx->voff.p, ctx->vx_in.p, ctx->vy_in.p,
        ctx->syx_in.p, ctx->sxy_in.p, ctx->ef_in.p, c.wave_index, sg * c.inv_compl_wave[0],
        sg * c.inv_compl_wave[1], c.compl_wave[0], c.compl_wave[1], rescale ? scale : 1.0,
        rescale ? ctx->sc_raw.p : nullptr, ctx->sc_path.p, ctx->sc_sign.p, ctx->sc_aoff.p,
        ctx->sc_apath.p, ctx->sc_cabs, ctx->sc_k, ctx->hdr.p, ctx->hdr32.p, ctx->vtx.p,
        ctx->vdd.p, ctx->vflag.p, ctx->wb.p);
    HIP_TRY(ctx, hipGetLastError());
    return MBS_OK;
}

// The upload of a batch: host checks, copies and the device pack on the
// context's stream.  `copied`: when given, the copies end with that event
// and the call returns without waiting (the source is page-locked staging);
// otherwise it waits for the stream.
static int upload_impl(mbs_ctx *ctx, const mbs_beam_batch *bt, const mbs_consts *c,
                       hipEvent_t copied)
{
    if (!ctx || !bt || !c)
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beams: NULL argument");
    ctx->beams_ok = false;
    ctx->spec_ok = false;
    ctx->pack_scale = 1.0;
    ctx->scaling_ok = false;
    ctx->have_result = false;
    if (bt->n_orient < 0 || bt->n_beams < 0 || (bt->n_orient > 0 && !bt->orient_offsets)
        || (bt->n_orient > 0 && !bt->weights))
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beams: bad orientation arrays");
    if (bt->n_beams > 0 && (!bt->beam_scalars || !bt->beam_nv || !bt->beam_external
                            || !bt->vert_offsets || !bt->vx || !bt->vy || !bt->slope_yx
                            || !bt->slope_xy || !bt->edge_flags))
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beams: NULL beam array");
    if (!(c->wave_index > 0) || !(c->eps1 >= 0)
        || !(c->compl_wave[0] * c->compl_wave[0] + c->compl_wave[1] * c->compl_wave[1] > 0))
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beams: bad constants");
    const int no = bt->n_orient, nb = bt->n_beams;
    if (no > 0 && (bt->orient_offsets[0] != 0 || bt->orient_offsets[no] != nb))
        return fail(ctx, MBS_ERR_INVALID, "orient_offsets must run from 0 to n_beams");
    for (int o = 0; o < no; ++o) {
        if (bt->orient_offsets[o + 1] < bt->orient_offsets[o])
            return fail(ctx, MBS_ERR_INVALID, "orient_offsets decrease at %d", o);
        // the MFMA kernel addresses a beam range's operands with 32-bit byte
        // offsets (WB_STRIDE floats per beam)
        if (bt->orient_offsets[o + 1] - bt->orient_offsets[o] > MBS_MAX_BEAMS_PER_ORIENT)
            return fail(ctx, MBS_ERR_CAPACITY,
                        "orientation %d holds %lld beams; at most %d per orientation", o,
                        (long long)(bt->orient_offsets[o + 1] - bt->orient_offsets[o]),
                        MBS_MAX_BEAMS_PER_ORIENT);
    }
    if (nb > 0 && bt->vert_offsets[0] != 0)
        return fail(ctx, MBS_ERR_INVALID, "vert_offsets[0] must be 0");
    std::memset(ctx->nv_hist, 0, sizeof ctx->nv_hist);
    int64_t edges = 0;
    for (int b = 0; b < nb; ++b) {
        const int n = bt->beam_nv[b];
        if (n < MBS_MIN_VERTICES || n > MBS_MAX_VERTICES)
            return fail(ctx, MBS_ERR_BAD_BEAM,
                        "beam %d has %d aperture vertices (engine accepts %d..%d)", b, n,
                        MBS_MIN_VERTICES, MBS_MAX_VERTICES);
        if (bt->vert_offsets[b + 1] - bt->vert_offsets[b] != n)
            return fail(ctx, MBS_ERR_INVALID, "vert_offsets of beam %d disagree with nv", b);
        ctx->nv_hist[n]++;
        for (int64_t v = bt->vert_offsets[b]; v < bt->vert_offsets[b + 1]; ++v)
            edges += ((bt->edge_flags[v] & 1) ? 1 : 0) + ((bt->edge_flags[v] & 2) ? 1 : 0);
    }
    const int64_t nvert = nb > 0 ? bt->vert_offsets[nb] : 0;
    ctx->mean_edges = nb > 0 ? 0.5 * (double)edges / nb : 0.0;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, ctx->bs_in.ensure((size_t)nb * MBS_BEAM_SCALARS));
    HIP_TRY(ctx, ctx->nv.ensure(nb));
    HIP_TRY(ctx, ctx->ext.ensure(nb));
    HIP_TRY(ctx, ctx->voff.ensure(nb + 1));
    HIP_TRY(ctx, ctx->ooff.ensure(no + 1));
    HIP_TRY(ctx, ctx->weights.ensure(no));
    HIP_TRY(ctx, ctx->vx_in.ensure(nvert));
    HIP_TRY(ctx, ctx->vy_in.ensure(nvert));
    HIP_TRY(ctx, ctx->syx_in.ensure(nvert));
    HIP_TRY(ctx, ctx->sxy_in.ensure(nvert));
    HIP_TRY(ctx, ctx->ef_in.ensure(nvert));
    HIP_TRY(ctx, ctx->hdr.ensure((size_t)nb * HDR_STRIDE));
    HIP_TRY(ctx, ctx->hdr32.ensure((size_t)nb * HDR32_STRIDE));
    HIP_TRY(ctx, ctx->vtx.ensure((size_t)nvert * VTX_STRIDE));
    HIP_TRY(ctx, ctx->vdd.ensure((size_t)nvert * VDD_STRIDE));
    HIP_TRY(ctx, ctx->wb.ensure((size_t)std::max(nb, 1) * WB_STRIDE));
    HIP_TRY(ctx, ctx->vflag.ensure(nvert));
    hipStream_t s = ctx->stream;
    if (ctx->timing)
        HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
    if (nb > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->bs_in.p, bt->beam_scalars,
                                    sizeof(double) * nb * MBS_BEAM_SCALARS,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->nv.p, bt->beam_nv, sizeof(int32_t) * nb,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->ext.p, bt->beam_external, sizeof(int32_t) * nb,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->voff.p, bt->vert_offsets, sizeof(int64_t) * (nb + 1),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->vx_in.p, bt->vx, sizeof(double) * nvert,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->vy_in.p, bt->vy, sizeof(double) * nvert,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->syx_in.p, bt->slope_yx, sizeof(double) * nvert,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sxy_in.p, bt->slope_xy, sizeof(double) * nvert,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->ef_in.p, bt->edge_flags, nvert,
                                    hipMemcpyHostToDevice, s));
    }
    if (no > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->ooff.p, bt->orient_offsets, sizeof(int64_t) * (no + 1),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->weights.p, bt->weights, sizeof(double) * no,
                                    hipMemcpyHostToDevice, s));
    }
    if (copied)
        HIP_TRY(ctx, hipEventRecord(copied, s));
    if (nb > 0) {
        const int rc = launch_pack(ctx, nb, *c, 1.0, false);
        if (rc)
            return rc;
    }
    if (ctx->timing) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
        ctx->ev_up = true;
    }
    if (!copied)
        HIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->h_ooff.assign(bt->orient_offsets, bt->orient_offsets + no + 1);
    if (no == 0)
        ctx->h_ooff.assign(1, 0);
    ctx->h_weights.assign(bt->weights, bt->weights + no);
    ctx->n_orient = no;
    ctx->n_beams = nb;
    ctx->n_vert = nvert;
    ctx->consts = *c;
    ctx->beams_ok = true;
    return MBS_OK;
}

extern "C" int mbs_upload_beams(mbs_ctx *ctx, const mbs_beam_batch *bt, const mbs_consts *c)
{
    return upload_impl(ctx, bt, c, nullptr);
}

// Byte layout of one staging area: every array 256-byte aligned.
static size_t stage_layout(int32_t no, int32_t nb, int64_t nv, size_t off[11])
{
    const size_t sizes[11] = {sizeof(int64_t) * (size_t)(no + 1), sizeof(double) * (size_t)no,
                              sizeof(double) * (size_t)nb * MBS_BEAM_SCALARS,
                              sizeof(int32_t) * (size_t)nb, sizeof(int32_t) * (size_t)nb,
                              sizeof(int64_t) * (size_t)(nb + 1), sizeof(double) * (size_t)nv,
                              sizeof(double) * (size_t)nv, sizeof(double) * (size_t)nv,
                              sizeof(double) * (size_t)nv, (size_t)nv};
    size_t total = 0;
    for (int i = 0; i < 11; ++i) {
        off[i] = total;
        total += (sizes[i] + 255) / 256 * 256;
    }
    return total;
}

extern "C" int mbs_stage_beams(mbs_ctx *ctx, int32_t n_orient, int32_t n_beams, int64_t n_vert,
                               mbs_beam_staging *out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_stage_beams: NULL argument");
    if (n_orient < 0 || n_beams < 0 || n_vert < 0)
        return fail(ctx, MBS_ERR_INVALID, "mbs_stage_beams: negative size");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int k = ctx->stage_next;
    mbs_ctx::Staging &st = ctx->stage[k];
    if (st.in_flight) {
        HIP_TRY(ctx, hipEventSynchronize(st.copied));
        st.in_flight = false;
    }
    if (!st.copied)
        HIP_TRY(ctx, hipEventCreateWithFlags(&st.copied, hipEventDisableTiming));
    size_t off[11];
    const size_t need = stage_layout(n_orient, n_beams, n_vert, off);
    if (st.cap < need) {
        if (st.mem)
            (void)hipHostFree(st.mem);
        st.mem = nullptr;
        st.cap = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&st.mem, std::max<size_t>(need, 256)));
        st.cap = std::max<size_t>(need, 256);
    }
    unsigned char *m = st.mem;
    mbs_beam_staging &v = st.view;
    v.n_orient = n_orient;
    v.n_beams = n_beams;
    v.n_vert = n_vert;
    v.orient_offsets = (int64_t *)(m + off[0]);
    v.weights = (double *)(m + off[1]);
    v.beam_scalars = (double *)(m + off[2]);
    v.beam_nv = (int32_t *)(m + off[3]);
    v.beam_external = (int32_t *)(m + off[4]);
    v.vert_offsets = (int64_t *)(m + off[5]);
    v.vx = (double *)(m + off[6]);
    v.vy = (double *)(m + off[7]);
    v.slope_yx = (double *)(m + off[8]);
    v.slope_xy = (double *)(m + off[9]);
    v.edge_flags = (uint8_t *)(m + off[10]);
    *out = v;
    ctx->stage_cur = k;
    ctx->stage_next = 1 - k;
    return MBS_OK;
}

extern "C" int mbs_upload_staged(mbs_ctx *ctx, const mbs_consts *consts)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (ctx->stage_cur < 0)
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_staged: nothing staged (mbs_stage_beams)");
    mbs_ctx::Staging &st = ctx->stage[ctx->stage_cur];
    const mbs_beam_staging &v = st.view;
    mbs_beam_batch bt;
    bt.n_orient = v.n_orient;
    bt.n_beams = v.n_beams;
    bt.orient_offsets = v.orient_offsets;
    bt.weights = v.weights;
    bt.beam_scalars = v.beam_scalars;
    bt.beam_nv = v.beam_nv;
    bt.beam_external = v.beam_external;
    bt.vert_offsets = v.vert_offsets;
    bt.vx = v.vx;
    bt.vy = v.vy;
    bt.slope_yx = v.slope_yx;
    bt.slope_xy = v.slope_xy;
    bt.edge_flags = v.edge_flags;
    if ((int64_t)(v.n_beams > 0 ? v.vert_offsets[v.n_beams] : 0) != v.n_vert)
        return fail(ctx, MBS_ERR_INVALID, "staged vert_offsets end at %lld, staged for %lld",
                    (long long)(v.n_beams > 0 ? v.vert_offsets[v.n_beams] : 0),
                    (long long)v.n_vert);
    ctx->stage_cur = -1;
    const int rc = upload_impl(ctx, &bt, consts, st.copied);
    if (rc == MBS_OK)
        st.in_flight = true;
    return rc;
}

extern "C" int mbs_upload_beam_scaling(mbs_ctx *ctx, const mbs_beam_scaling *sc)
{
    if (!ctx || !sc)
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beam_scaling: NULL argument");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    ctx->scaling_ok = false;
    const int nb = ctx->n_beams;
    if (nb > 0 && (!sc->raw_jones || !sc->path || !sc->sign || !sc->abs_offsets))
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beam_scaling: NULL array");
    if (!std::isfinite(sc->c_abs) || !std::isfinite(sc->wave_index))
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beam_scaling: bad constants");
    const int64_t na = nb > 0 ? sc->abs_offsets[nb] : 0;
    if (nb > 0 && (sc->abs_offsets[0] != 0 || (na > 0 && !sc->abs_paths)))
        return fail(ctx, MBS_ERR_INVALID, "abs_offsets must start at 0 (and abs_paths be set)");
    for (int b = 0; b < nb; ++b)
        if (sc->abs_offsets[b + 1] < sc->abs_offsets[b])
            return fail(ctx, MBS_ERR_INVALID, "abs_offsets decrease at %d", b);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, ctx->sc_raw.ensure((size_t)nb * 8));
    HIP_TRY(ctx, ctx->sc_path.ensure(nb));
    HIP_TRY(ctx, ctx->sc_sign.ensure(nb));
    HIP_TRY(ctx, ctx->sc_aoff.ensure((size_t)nb + 1));
    HIP_TRY(ctx, ctx->sc_apath.ensure((size_t)na));
    hipStream_t s = ctx->stream;
    if (nb > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sc_raw.p, sc->raw_jones, sizeof(double) * 8 * nb,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sc_path.p, sc->path, sizeof(double) * nb,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sc_sign.p, sc->sign, sizeof(double) * nb,
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sc_aoff.p, sc->abs_offsets, sizeof(int64_t) * (nb + 1),
                                    hipMemcpyHostToDevice, s));
        if (na > 0)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->sc_apath.p, sc->abs_paths, sizeof(double) * na,
                                        hipMemcpyHostToDevice, s));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->sc_cabs = sc->c_abs;
    ctx->sc_k = sc->wave_index;
    ctx->scaling_ok = true;
    return MBS_OK;
}

extern "C" int mbs_rescale_beams(mbs_ctx *ctx, double scale)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    if (!ctx->scaling_ok)
        return fail(ctx, MBS_ERR_INVALID, "mbs_upload_beam_scaling has not been called");
    if (!(scale > 0) || !std::isfinite(scale))
        return fail(ctx, MBS_ERR_INVALID, "scale must be positive");
    ctx->have_result = false;
    ctx->spec_ok = false;
    if (ctx->n_beams == 0)
        return MBS_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int rc = launch_pack(ctx, ctx->n_beams, ctx->consts, scale,
                               std::fabs(scale - 1.0) > 1e-15);
    ctx->pack_scale = std::fabs(scale - 1.0) > 1e-15 ? scale : 1.0;
    if (rc)
        return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

// Splits the batch into work items.
//  * Segment mode (default): every orientation is cut into beam segments of
//    about equal size, so items carry equal work whatever the spread of beam
//    counts between orientations (50 to 40k beams in an aggregate average);
//    each item writes a partial Jones grid and a second kernel adds the
//    segments of each orientation in order and converts to Mueller.
//  * Orientation mode: runs of whole orientations with about equal beam
//    totals; each item keeps the weighted Mueller sum of its run.  Used when
//    the partial Jones grids of segment mode would exceed the memory budget.
static const size_t kPartialBudget = (size_t)8 << 30;  // bytes
// Workgroups of a persistent (work-queue) launch: 256 CUs x 8 slots.
static const int kQueueBlocks = 2048;

static bool plan_items(mbs_ctx *ctx, int n_tiles, int target, int work_mode,
                       std::vector<WorkItem> &items)
{
    items.clear();
    const int no = ctx->n_orient;
    const int64_t nb = ctx->n_beams;
    const size_t cells = (size_t)ctx->cells;
    const int64_t want_items = std::max<int64_t>(1, (target + n_tiles - 1) / n_tiles);
    // segment size in beams: enough items to fill the device, but at least
    // kMinSeg beams each so the per-item writes stay negligible
    const int64_t kMinSeg = 32;
    int64_t seg = std::max<int64_t>(kMinSeg, (nb + want_items - 1) / std::max<int64_t>(want_items, 1));
    int64_t n_seg = 0;
    for (int o = 0; o < no; ++o) {
        const int64_t n = ctx->h_ooff[o + 1] - ctx->h_ooff[o];
        n_seg += std::max<int64_t>(1, (n + seg - 1) / seg);
    }
    const bool segments_fit = (size_t)n_seg * cells * 8 * sizeof(double) <= kPartialBudget;
    const bool jones = work_mode == 2 || (work_mode == 0 && segments_fit && nb > 0);
    if (jones) {
        for (int o = 0; o < no; ++o) {
            const int64_t b0 = ctx->h_ooff[o], b1 = ctx->h_ooff[o + 1];
            const int64_t n = b1 - b0;
            const int64_t k = std::max<int64_t>(1, (n + seg - 1) / seg);
            for (int64_t i = 0; i < k; ++i)
                items.push_back(WorkItem{o, o + 1, b0 + n * i / k, b0 + n * (i + 1) / k});
        }
        return true;
    }
    int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(want_items, no));
    const double per = (double)nb / chunks;
    int o = 0;
    for (int ci = 0; ci < chunks && o < no; ++ci) {
        const double goal = per * (ci + 1);
        int e = o + 1;
        while (e < no && (double)ctx->h_ooff[e] < goal && (no - e) > (chunks - ci - 1))
            ++e;
        if (ci == chunks - 1)
            e = no;
        items.push_back(WorkItem{o, e, 0, 0});
        o = e;
    }
    return false;
}

template <int PREC, int NS, int JONES, int PER, int Q>
static void launch_one(const KernelArgs &a, int blocks, hipStream_t s)
{
    if constexpr (PREC == 4)
        diffract_mfma_kernel<NS, JONES, PER, Q, 1><<<blocks, 256, 0, s>>>(
            a.items, a.n_items, a.n_tiles, a.cells, a.n_az, a.tab, a.hdr, a.wb, a.vtx, a.vdd,
            a.vflag, a.nv, a.voff, a.ext, a.ooff, a.weights, a.k, a.eps1, a.out, a.out_ns,
            a.per_orient, a.queue, a.tab32, a.sidx, a.fbuf);
    else if constexpr (PREC == 3)
        diffract_incoherent_kernel<<<blocks, 256, 0, s>>>(
            a.items, a.n_items, a.n_tiles, a.cells, a.tab, a.hdr, a.vtx, a.vdd, a.vflag, a.nv,
            a.voff, a.ooff, a.weights, a.k, a.eps1, a.out);
    else if constexpr (PREC == 2)
        diffract_mfma_kernel<NS, JONES, PER, Q><<<blocks, 256, 0, s>>>(
            a.items, a.n_items, a.n_tiles, a.cells, a.n_az, a.tab, a.hdr, a.wb, a.vtx, a.vdd,
            a.vflag,
            a.nv, a.voff, a.ext, a.ooff, a.weights, a.k, a.eps1, a.out, a.out_ns,
            a.per_orient, a.queue, a.tab32);
    else
        diffract_kernel<PREC, NS, JONES, PER, Q><<<blocks, 256, 0, s>>>(
            a.items, a.n_items, a.n_tiles, a.cells, a.n_az, a.tab, a.hdr, a.hdr32, a.vtx, a.vdd,
            a.vflag, a.nv, a.voff, a.ext, a.ooff, a.weights, a.k, a.eps1, a.out, a.out_ns,
            a.per_orient, a.queue, a.tab32);
}

// queue: the persistent work-queue kernels (beam segments, no per-orientation
// grids), which take their units from a.queue.  a.waves selects an occupancy
// variant of the beam-segment MFMA kernel (4 or 6 waves per SIMD).
static void launch_diffract(int prec, int ns, int jones, int per, bool queue,
                            const KernelArgs &a, int blocks, hipStream_t s)
{
    if (prec == 2 && !ns && jones && !per && !queue && (a.waves == 4 || a.waves == 6)) {
        if (a.waves == 4)
            diffract_mfma_kernel<0, 1, 0, 0, 0, 4><<<blocks, 256, 0, s>>>(
                a.items, a.n_items, a.n_tiles, a.cells, a.n_az, a.tab, a.hdr, a.wb, a.vtx,
                a.vdd, a.vflag, a.nv, a.voff, a.ext, a.ooff, a.weights, a.k, a.eps1, a.out,
                a.out_ns, a.per_orient, a.queue, a.tab32);
        else
            diffract_mfma_kernel<0, 1, 0, 0, 0, 6><<<blocks, 256, 0, s>>>(
                a.items, a.n_items, a.n_tiles, a.cells, a.n_az, a.tab, a.hdr, a.wb, a.vtx,
                a.vdd, a.vflag, a.nv, a.voff, a.ext, a.ooff, a.weights, a.k, a.eps1, a.out,
                a.out_ns, a.per_orient, a.queue, a.tab32);
        return;
    }
#define MBS_L(P, N, J, R, Q)                                                 \
    if (prec == P && ns == N && jones == J && per == R && (int)queue == Q) { \
        launch_one<P, N, J, R, Q>(a, blocks, s);                             \
        return;                                                              \
    }
    MBS_L(0, 0, 0, 0, 0) MBS_L(0, 1, 0, 0, 0) MBS_L(0, 0, 1, 0, 0) MBS_L(0, 1, 1, 0, 0)
    MBS_L(0, 0, 0, 1, 0) MBS_L(0, 1, 0, 1, 0)
    MBS_L(1, 0, 0, 0, 0) MBS_L(1, 1, 0, 0, 0) MBS_L(1, 0, 1, 0, 0) MBS_L(1, 1, 1, 0, 0)
    MBS_L(1, 0, 0, 1, 0) MBS_L(1, 1, 0, 1, 0)
    MBS_L(2, 0, 0, 0, 0) MBS_L(2, 1, 0, 0, 0) MBS_L(2, 0, 1, 0, 0) MBS_L(2, 1, 1, 0, 0)
    MBS_L(2, 0, 0, 1, 0) MBS_L(2, 1, 0, 1, 0)
    MBS_L(3, 0, 0, 0, 0)
    MBS_L(0, 0, 1, 0, 1) MBS_L(0, 1, 1, 0, 1) MBS_L(1, 0, 1, 0, 1) MBS_L(1, 1, 1, 0, 1)
    MBS_L(2, 0, 1, 0, 1) MBS_L(2, 1, 1, 0, 1)
    MBS_L(4, 0, 0, 0, 0) MBS_L(4, 1, 0, 0, 0) MBS_L(4, 0, 1, 0, 0) MBS_L(4, 1, 1, 0, 0)
#undef MBS_L
}

// Launches the diffraction kernel and its reduction on the context's stream:
// kernels only (no copies, no memsets), so the sequence can be captured in a
// graph.  The partial sums and the work list are a.out / a.out_ns / a.items;
// the reduction writes res (res_ns with NS; per_res, the per-orientation
// grids, in Jones mode with PER).  With `events`, ev[2..3] bracket the main
// kernel.
static int issue_diffraction(mbs_ctx *ctx, int prec, int ns, bool jones, int per_k,
                             const KernelArgs &a, int blocks, int n_items, double *res,
                             double *res_ns, double *per_res, bool events)
{
    hipStream_t s = ctx->stream;
    const int cells = ctx->cells;
    const size_t mcount = (size_t)cells * 16;
    const bool queue = a.queue != nullptr;
    if (events)
        HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
    if (queue) {
        // the segment plan and the queue header, made on the device (timed
        // with the main kernel)
        plan_segments_kernel<<<1, 1024, 0, s>>>(a.ooff, ctx->n_orient, ctx->queue_want,
                                                a.n_tiles, const_cast<WorkItem *>(a.items),
                                                a.queue);
        HIP_TRY(ctx, hipGetLastError());
    }
    launch_diffract(prec, ns, jones ? 1 : 0, per_k, queue, a, blocks, s);
    HIP_TRY(ctx, hipGetLastError());
    if (events)
        HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    const int rt = 256;
    if (jones) {
        // the mixed kernels (prec 1, 2, 4) leave FP32 Jones partials
        const bool f32 = prec == 1 || prec == 2 || prec == 4;
        const dim3 jb((cells + rt - 1) / rt);
        for (int q = 0; q <= ns; ++q) {
            const double *src = q ? a.out_ns : a.out;
            double *dst = q ? res_ns : res;
            double *po = q ? nullptr : per_res;
            if (f32)
                jones_to_mueller_kernel<float><<<jb, rt, 0, s>>>(
                    reinterpret_cast<const float *>(src), a.items, n_items, cells, ctx->n_az,
                    ctx->weights.p, dst, po, a.queue);
            else
                jones_to_mueller_kernel<double><<<jb, rt, 0, s>>>(
                    src, a.items, n_items, cells, ctx->n_az, ctx->weights.p, dst, po, a.queue);
            HIP_TRY(ctx, hipGetLastError());
        }
    } else {
        const int rb = (int)((mcount + rt - 1) / rt);
        reduce_mueller_kernel<<<rb, rt, 0, s>>>(a.out, n_items, cells, ctx->n_az, res);
        HIP_TRY(ctx, hipGetLastError());
        if (ns) {
            reduce_mueller_kernel<<<rb, rt, 0, s>>>(a.out_ns, n_items, cells, ctx->n_az, res_ns);
            HIP_TRY(ctx, hipGetLastError());
        }
    }
    return MBS_OK;
}

// The kernel arguments that come from the grid and the uploaded batch.
static void batch_args(const mbs_ctx *ctx, int prec, int n_items, int n_tiles, KernelArgs &a)
{
    std::memset(&a, 0, sizeof a);  // padding too: the struct is a graph-cache key
    a.n_items = n_items;
    a.n_tiles = n_tiles;
    a.cells = ctx->cells;
    a.n_az = ctx->n_az;
    a.tab = ctx->tab.p;
    a.tab32 = ctx->tab32.p;
    a.hdr = ctx->hdr.p;
    a.hdr32 = ctx->hdr32.p;
    a.wb = ctx->wb.p;
    a.vtx = ctx->vtx.p;
    a.vdd = ctx->vdd.p;
    a.vflag = ctx->vflag.p;
    a.nv = ctx->nv.p;
    a.voff = ctx->voff.p;
    a.ext = ctx->ext.p;
    a.ooff = ctx->ooff.p;
    a.weights = ctx->weights.p;
    a.k = ctx->consts.wave_index;
    // The mixed kernel forms edge phasor differences in FP32, so its
    // switch to the exact half-angle edge form happens at |C| <= 3e-5, the
    // floor the reference applies to its FP32 build (GpuDiffraction.cu:498-512).
    // The FP64 kernels (prec 0, and 3 = the FP64 incoherent sum) keep the
    // handler's eps1.  (prec 4: the mixed MFMA kernel with the spectrum
    // source.)
    a.eps1 = (prec == 1 || prec == 2 || prec == 4) ? std::max(ctx->consts.eps1, 3e-5)
                                                   : ctx->consts.eps1;
    if (prec == 4) {
        a.sidx = ctx->spec_sidx.p;
        a.fbuf = ctx->spec_f.p;
    }
}

extern "C" {

int mbs_diffract(mbs_ctx *ctx, uint32_t flags)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (!ctx->grid_ok)
        return fail(ctx, MBS_ERR_NO_GRID, "mbs_set_grid has not been called");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    // mixed precision runs the coherent sum on MFMA unless MBS_NO_MFMA, or
    // unless the autotuner found the vector-ALU form faster on this batch
    const int base = (flags & MBS_PREC_MIXED) ? 1 : 0;
    const mbs_ctx::Tuned &tn = ctx->tuned[base];
    const bool mfma = !(flags & MBS_NO_MFMA) && !(tn.valid && !tn.mfma);
    const bool incoh = (flags & MBS_INCOHERENT) != 0;
    const bool spec = (flags & MBS_SPECTRUM) != 0;
    const int prec = incoh ? 3 : spec ? 4 : base ? (mfma ? 2 : 1) : 0;
    const int ns = (flags & MBS_NOSHADOW) ? 1 : 0;
    const int per = (flags & MBS_PER_ORIENT) ? 1 : 0;
    if (flags & ~(uint32_t)(MBS_PREC_MIXED | MBS_NOSHADOW | MBS_PER_ORIENT | MBS_NO_MFMA |
                            MBS_INCOHERENT | MBS_SPECTRUM))
        return fail(ctx, MBS_ERR_INVALID, "unknown flag bits 0x%x", flags);
    if (incoh && (base || ns || per))
        return fail(ctx, MBS_ERR_INVALID,
                    "MBS_INCOHERENT runs in FP64 without MBS_NOSHADOW or MBS_PER_ORIENT");
    if (spec && (!base || per || incoh || (flags & MBS_NO_MFMA)))
        return fail(ctx, MBS_ERR_INVALID,
                    "MBS_SPECTRUM runs in mixed precision on the MFMA contraction, without "
                    "MBS_PER_ORIENT");
    if (spec && !ctx->spec_ok)
        return fail(ctx, MBS_ERR_INVALID,
                    "MBS_SPECTRUM: no spectrum source for this batch and grid "
                    "(mbs_spectrum_source)");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_tables(ctx);
    if (rc)
        return rc;
    const int cells = ctx->cells;
    const int n_tiles = (cells + 255) / 256;
    const int target = tn.valid ? tn.target_blocks : ctx->target_blocks;
    const int mode = tn.valid ? tn.work_mode : ctx->work_mode;
    // the device queue plans beam segments on the device; the incoherent sum
    // and per-orientation grids keep the host plan
    const bool queue = mode == MBS_WORK_QUEUE && !incoh && !per && !spec && ctx->n_beams > 0;
    bool jones = true;
    int n_items;
    if (queue) {
        ctx->queue_want = (int)std::max<int64_t>(1, (target + n_tiles - 1) / n_tiles);
        n_items = ctx->queue_want + ctx->n_orient;  // the most the device plan makes
        ctx->h_items.clear();
    } else {
        // the incoherent sum keeps per-item Mueller sums (orientation items)
        jones = plan_items(ctx, n_tiles, target,
                           incoh ? MBS_WORK_ORIENTATIONS
                                 : mode == MBS_WORK_QUEUE ? MBS_WORK_SEGMENTS : mode,
                           ctx->h_items);
        n_items = (int)ctx->h_items.size();
    }
    const size_t mcount = (size_t)cells * 16;
    HIP_TRY(ctx, ctx->res.ensure(mcount));
    if (ns)
        HIP_TRY(ctx, ctx->res_ns.ensure(mcount));
    if (per)
        HIP_TRY(ctx, ctx->per_orient.ensure((size_t)ctx->n_orient * mcount));
    hipStream_t s = ctx->stream;
    if (n_items == 0) {
        const int zb = (int)((mcount + 255) / 256);
        zero_kernel<<<zb, 256, 0, s>>>(ctx->res.p, mcount);
        if (ns)
            zero_kernel<<<zb, 256, 0, s>>>(ctx->res_ns.p, mcount);
        HIP_TRY(ctx, hipGetLastError());
        ctx->have_result = true;
        ctx->have_ns = ns;
        ctx->have_per = per;
        return MBS_OK;
    }
    const size_t pstride = (size_t)cells * (jones ? 8 : 16);
    HIP_TRY(ctx, ctx->part.ensure((size_t)n_items * pstride));
    if (ns)
        HIP_TRY(ctx, ctx->part_ns.ensure((size_t)n_items * pstride));
    HIP_TRY(ctx, ctx->items.ensure(n_items));
    if (queue)
        HIP_TRY(ctx, ctx->queue.ensure(Q_WORDS));
    else
        HIP_TRY(ctx, hipMemcpyAsync(ctx->items.p, ctx->h_items.data(),
                                    sizeof(WorkItem) * n_items, hipMemcpyHostToDevice, s));
    KernelArgs a;
    batch_args(ctx, prec, n_items, n_tiles, a);
    a.waves = tn.valid ? tn.waves : 0;
    a.items = ctx->items.p;
    a.out = ctx->part.p;
    a.out_ns = ns ? ctx->part_ns.p : nullptr;
    a.per_orient = (per && !jones) ? ctx->per_orient.p : nullptr;
    a.queue = queue ? ctx->queue.p : nullptr;
    const int64_t total = (int64_t)n_items * n_tiles;
    if (total > (int64_t)1 << 30)
        return fail(ctx, MBS_ERR_CAPACITY, "launch of %lld workgroups", (long long)total);
    // the queue kernels are persistent: workgroups enough to fill every
    // CU's slots, each taking units until the queue is empty
    const int blocks = queue ? (int)std::min<int64_t>((total + 7) / 8 * 8, kQueueBlocks)
                             : (int)((total + 7) / 8 * 8);
    ctx->last_items = n_items;
    ctx->last_blocks = blocks;
    const int per_k = (per && !jones) ? 1 : 0;
    if (ctx->use_graph) {
        // Replay the captured kernel sequence when nothing it depends on
        // (launch shape, buffers, sizes, constants) has changed.
        const int tail[8] = {prec, ns, per, jones ? 1 : 0, blocks, n_items, cells, per_k};
        const void *outs[3] = {ctx->res.p, ctx->res_ns.p, ctx->per_orient.p};
        std::vector<unsigned char> key(sizeof(KernelArgs) + sizeof tail + sizeof outs);
        std::memcpy(key.data(), &a, sizeof(KernelArgs));
        std::memcpy(key.data() + sizeof(KernelArgs), tail, sizeof tail);
        std::memcpy(key.data() + sizeof(KernelArgs) + sizeof tail, outs, sizeof outs);
        if (!ctx->graph_exec || key != ctx->graph_key) {
            if (ctx->graph_exec) {
                (void)hipGraphExecDestroy(ctx->graph_exec);
                ctx->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc2 = issue_diffraction(ctx, prec, ns, jones, per_k, a, blocks, n_items,
                                              ctx->res.p, ctx->res_ns.p,
                                              per ? ctx->per_orient.p : nullptr, false);
            hipError_t ce = hipStreamEndCapture(s, &g);
            if (rc2)
                return rc2;
            if (ce != hipSuccess)
                return fail(ctx, MBS_ERR_DEVICE, "graph capture: %s", hipGetErrorString(ce));
            ce = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ce != hipSuccess) {
                ctx->graph_exec = nullptr;
                return fail(ctx, MBS_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ce));
            }
            ctx->graph_key.swap(key);
            ++ctx->graph_captures;
        }
        ++ctx->graph_replays;
        if (ctx->timing)
            HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
        HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, s));
        if (ctx->timing)
            HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    } else {
        const int rc2 = issue_diffraction(ctx, prec, ns, jones, per_k, a, blocks, n_items,
                                          ctx->res.p, ctx->res_ns.p,
                                          per ? ctx->per_orient.p : nullptr, ctx->timing);
        if (rc2)
            return rc2;
    }
    if (ctx->timing) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
        ctx->ev_diff = true;
    }
    ctx->have_result = true;
    ctx->have_ns = ns;
    ctx->have_per = per;
    return MBS_OK;
}

// Orientation reduction into caller-owned device memory, with the work list
// and the partial sums in `scratch` (see scratch_lookup): the launch is the
// one mbs_diffract makes, with the item count capped so that everything fits
// the array.
int mbs_diffract_noalloc(mbs_ctx *ctx, uint32_t flags, double *d_m_full, double *d_m_noshadow)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (!ctx->grid_ok)
        return fail(ctx, MBS_ERR_NO_GRID, "mbs_set_grid has not been called");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    if (flags & ~(uint32_t)(MBS_PREC_MIXED | MBS_NOSHADOW | MBS_NO_MFMA))
        return fail(ctx, MBS_ERR_INVALID,
                    "mbs_diffract_noalloc takes MBS_PREC_MIXED, MBS_NO_MFMA and MBS_NOSHADOW");
    const int ns = (flags & MBS_NOSHADOW) ? 1 : 0;
    if (!d_m_full || (ns != 0) != (d_m_noshadow != nullptr))
        return fail(ctx, MBS_ERR_INVALID,
                    "mbs_diffract_noalloc: d_m_full is required, d_m_noshadow exactly with "
                    "MBS_NOSHADOW");
    hipError_t serr = hipSuccess;
    unsigned char *base_ptr = scratch_address(ctx->device, &serr);
    if (!base_ptr)
        return fail(ctx, MBS_ERR_DEVICE, "no scratch address for device %d: %s", ctx->device,
                    hipGetErrorString(serr));
    const int base = (flags & MBS_PREC_MIXED) ? 1 : 0;
    const int prec = base ? ((flags & MBS_NO_MFMA) ? 1 : 2) : 0;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_tables(ctx);
    if (rc)
        return rc;
    const int cells = ctx->cells;
    const int n_tiles = (cells + 255) / 256;
    const size_t mcount = (size_t)cells * 16;
    hipStream_t s = ctx->stream;
    // Items that fit: every item needs its partial grid(s) and a list entry
    // (one alignment gap of at most 256 bytes between list and partials).
    const size_t cap = MBS_NOALLOC_SCRATCH_BYTES - 256;
    const size_t per_jones = (size_t)cells * 8 * sizeof(double) * (1 + ns) + sizeof(WorkItem);
    const size_t per_mueller = (size_t)cells * 16 * sizeof(double) * (1 + ns) + sizeof(WorkItem);
    if (per_mueller > cap)
        return fail(ctx, MBS_ERR_CAPACITY,
                    "mbs_diffract_noalloc: a grid of %d cells needs %zu bytes of scratch per "
                    "work item, above the %zu of MBS_NOALLOC_SCRATCH_BYTES",
                    cells, per_mueller, (size_t)MBS_NOALLOC_SCRATCH_BYTES);
    const int64_t max_jones = (int64_t)(cap / per_jones);
    const int64_t max_mueller = (int64_t)(cap / per_mueller);
    std::vector<WorkItem> &items = ctx->h_items;
    bool jones = false;
    if (ctx->n_orient <= max_jones) {
        // beam segments, halving the target until the segments fit
        int64_t target = std::min<int64_t>(ctx->target_blocks, max_jones * n_tiles);
        for (;;) {
            jones = plan_items(ctx, n_tiles, (int)target, MBS_WORK_SEGMENTS, items);
            if ((int64_t)items.size() <= max_jones || target <= n_tiles)
                break;
            target = std::max<int64_t>(n_tiles, target / 2);
        }
        if ((int64_t)items.size() > max_jones)
            jones = false;
    }
    if (!jones) {
        const int64_t target = std::min<int64_t>(ctx->target_blocks, max_mueller * n_tiles);
        plan_items(ctx, n_tiles, (int)target, MBS_WORK_ORIENTATIONS, items);
        if ((int64_t)items.size() > max_mueller)
            return fail(ctx, MBS_ERR_CAPACITY, "mbs_diffract_noalloc: %zu work items exceed %lld",
                        items.size(), (long long)max_mueller);
    }
    const int n_items = (int)items.size();
    if (n_items == 0) {
        const int zb = (int)((mcount + 255) / 256);
        zero_kernel<<<zb, 256, 0, s>>>(d_m_full, mcount);
        if (ns)
            zero_kernel<<<zb, 256, 0, s>>>(d_m_noshadow, mcount);
        HIP_TRY(ctx, hipGetLastError());
        return MBS_OK;
    }
    const size_t list_bytes = ((sizeof(WorkItem) * n_items + 255) / 256) * 256;
    const size_t pstride = (size_t)cells * (jones ? 8 : 16);
    const size_t part_bytes = (size_t)n_items * pstride * sizeof(double);
    if (list_bytes + part_bytes * (1 + ns) > MBS_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, MBS_ERR_CAPACITY, "mbs_diffract_noalloc: %zu bytes of scratch needed",
                    list_bytes + part_bytes * (1 + ns));
    WorkItem *d_items = (WorkItem *)base_ptr;
    double *d_part = (double *)(base_ptr + list_bytes);
    double *d_part_ns = ns ? (double *)(base_ptr + list_bytes + part_bytes) : nullptr;
    HIP_TRY(ctx, hipMemcpyAsync(d_items, items.data(), sizeof(WorkItem) * n_items,
                                hipMemcpyHostToDevice, s));
    KernelArgs a;
    batch_args(ctx, prec, n_items, n_tiles, a);
    a.items = d_items;
    a.out = d_part;
    a.out_ns = d_part_ns;
    a.per_orient = nullptr;
    const int64_t total = (int64_t)n_items * n_tiles;
    if (total > (int64_t)1 << 30)
        return fail(ctx, MBS_ERR_CAPACITY, "launch of %lld workgroups", (long long)total);
    const int blocks = (int)((total + 7) / 8 * 8);
    ctx->last_items = n_items;
    ctx->last_blocks = blocks;
    rc = issue_diffraction(ctx, prec, ns, jones, 0, a, blocks, n_items, d_m_full, d_m_noshadow,
                           nullptr, ctx->timing);
    if (rc)
        return rc;
    if (ctx->timing) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
        ctx->ev_diff = true;
    }
    return MBS_OK;
}

int mbs_set_timing(mbs_ctx *ctx, int enable)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (enable)
        for (hipEvent_t &e : ctx->ev)
            if (!e)
                HIP_TRY(ctx, hipEventCreate(&e));
    ctx->timing = enable != 0;
    ctx->ev_up = ctx->ev_diff = false;
    return MBS_OK;
}

int mbs_get_timing(mbs_ctx *ctx, mbs_timing *out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_get_timing: NULL argument");
    if (!ctx->timing)
        return fail(ctx, MBS_ERR_INVALID, "timing is not enabled (mbs_set_timing)");
    std::memset(out, 0, sizeof *out);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->ev_up) {
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
        HIP_TRY(ctx, hipEventElapsedTime(&out->upload_ms, ctx->ev[0], ctx->ev[1]));
    }
    if (ctx->ev_diff) {
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev[4]));
        HIP_TRY(ctx, hipEventElapsedTime(&out->kernel_ms, ctx->ev[2], ctx->ev[3]));
        HIP_TRY(ctx, hipEventElapsedTime(&out->reduce_ms, ctx->ev[3], ctx->ev[4]));
    }
    out->n_items = ctx->last_items;
    out->n_blocks = ctx->last_blocks;
    return MBS_OK;
}

}  // extern "C"

// LDS of the aperture FFT's row and column workgroups for a window of
// 2^logl.
static size_t fft_rows_lds(int logl)
{
    return (size_t)(fft_ld(1 << logl) + (1 << logl)) * sizeof(float2) +
           (size_t)MBS_MAX_VERTICES * sizeof(double);
}
static size_t fft_cols_lds(int logl) { return (size_t)(FFT_WCOLS + 1) * ((size_t)1 << logl) * sizeof(float2); }

// Row and column passes over the window [j0, j0 + 2^logl): the instance of
// the window length, fft_min_logl(LOGN) <= logl <= LOGN.
template <int LOGN, int LOGL>
static int launch_fft(mbs_ctx *ctx, int nb, int j0, int logl)
{
    if constexpr (LOGL < LOGN)
        if (logl != LOGL)
            return launch_fft<LOGN, LOGL + 1>(ctx, nb, j0, logl);
    constexpr int N = 1 << LOGN, L = 1 << LOGL;
    const size_t lds_rows = fft_rows_lds(LOGL), lds_cols = fft_cols_lds(LOGL);
    HIP_TRY(ctx, hipFuncSetAttribute((const void *)fft_rows_kernel<LOGN, LOGL>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rows));
    HIP_TRY(ctx, hipFuncSetAttribute((const void *)fft_cols_kernel<LOGN, LOGL>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cols));
    hipStream_t s = ctx->stream;
    const int rb = (nb * L + 7) / 8 * 8;
    const int cb = (nb * (N / FFT_WCOLS) + 7) / 8 * 8;
    if (ctx->timing)
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[0], s));
    fft_rows_kernel<LOGN, LOGL><<<rb, fft_wrow_threads(LOGL), lds_rows, s>>>(
        nb, j0, ctx->fft_rec.p, ctx->fft_tw.p, ctx->fft_grid.p);
    HIP_TRY(ctx, hipGetLastError());
    if (ctx->timing)
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[1], s));
    fft_cols_kernel<LOGN, LOGL><<<cb, FFT_WCOL_THREADS, lds_cols, s>>>(
        nb, j0, ctx->fft_rec.p, ctx->fft_tw.p, ctx->fft_grid.p);
    HIP_TRY(ctx, hipGetLastError());
    if (ctx->timing)
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[2], s));
    return MBS_OK;
}

static int fetch(mbs_ctx *ctx, const double *src, size_t n, double *dst_add, bool add)
{
    if (ctx->pinned_cap < n) {
        if (ctx->pinned)
            (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_cap = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&ctx->pinned, n * sizeof(double)));
        ctx->pinned_cap = n;
    }
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, src, n * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (add)
        for (size_t i = 0; i < n; ++i)
            dst_add[i] += ctx->pinned[i];
    else
        std::memcpy(dst_add, ctx->pinned, n * sizeof(double));
    return MBS_OK;
}

extern "C" {

int mbs_accumulate_mueller(mbs_ctx *ctx, double *m_full, double *m_noshadow)
{
    if (!ctx || !m_full)
        return fail(ctx, MBS_ERR_INVALID, "mbs_accumulate_mueller: NULL argument");
    if (!ctx->have_result)
        return fail(ctx, MBS_ERR_INVALID, "no result: call mbs_diffract first");
    if (m_noshadow && !ctx->have_ns)
        return fail(ctx, MBS_ERR_INVALID, "no-shadow result was not requested (MBS_NOSHADOW)");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->cells * 16;
    int rc = fetch(ctx, ctx->res.p, n, m_full, true);
    if (rc)
        return rc;
    if (m_noshadow)
        return fetch(ctx, ctx->res_ns.p, n, m_noshadow, true);
    return MBS_OK;
}

int mbs_fetch_per_orient(mbs_ctx *ctx, double *out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_fetch_per_orient: NULL argument");
    if (!ctx->have_result || !ctx->have_per)
        return fail(ctx, MBS_ERR_INVALID, "per-orientation result was not requested");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return fetch(ctx, ctx->per_orient.p, (size_t)ctx->n_orient * ctx->cells * 16, out, false);
}

int mbs_diffract_batch(mbs_ctx *ctx, const mbs_grid *grid, const mbs_beam_batch *batch,
                       const mbs_consts *consts, uint32_t flags, double *m_full,
                       double *m_noshadow)
{
    int rc = mbs_set_grid(ctx, grid);
    if (rc)
        return rc;
    rc = mbs_upload_beams(ctx, batch, consts);
    if (rc)
        return rc;
    rc = mbs_diffract(ctx, flags | (m_noshadow ? MBS_NOSHADOW : 0u));
    if (rc)
        return rc;
    return mbs_accumulate_mueller(ctx, m_full, m_noshadow);
}

int mbs_azimuth_average(mbs_ctx *ctx, int noshadow, double *rows)
{
    if (!ctx || !rows)
        return fail(ctx, MBS_ERR_INVALID, "mbs_azimuth_average: NULL argument");
    if (!ctx->have_result)
        return fail(ctx, MBS_ERR_INVALID, "no result: call mbs_diffract first");
    if (noshadow && !ctx->have_ns)
        return fail(ctx, MBS_ERR_INVALID, "no-shadow result was not requested (MBS_NOSHADOW)");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->n_zen1 * 18;
    HIP_TRY(ctx, ctx->avg_rows.ensure(n));
    azimuth_average_kernel<<<ctx->n_zen1, 64, 0, ctx->stream>>>(
        noshadow ? ctx->res_ns.p : ctx->res.p, ctx->theta.p, ctx->n_az, ctx->n_zen1 - 1,
        ctx->az_step, ctx->avg_rows.p);
    HIP_TRY(ctx, hipGetLastError());
    return fetch(ctx, ctx->avg_rows.p, n, rows, false);
}

int mbs_azimuth_average_per_orient(mbs_ctx *ctx, double *rows)
{
    if (!ctx || !rows)
        return fail(ctx, MBS_ERR_INVALID, "mbs_azimuth_average_per_orient: NULL argument");
    if (!ctx->have_result || !ctx->have_per)
        return fail(ctx, MBS_ERR_INVALID, "per-orientation result was not requested");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->n_orient * ctx->n_zen1 * 18;
    HIP_TRY(ctx, ctx->avg_rows.ensure(n));
    azimuth_average_kernel<<<dim3(ctx->n_zen1, ctx->n_orient), 64, 0, ctx->stream>>>(
        ctx->per_orient.p, ctx->theta.p, ctx->n_az, ctx->n_zen1 - 1, ctx->az_step,
        ctx->avg_rows.p);
    HIP_TRY(ctx, hipGetLastError());
    return fetch(ctx, ctx->avg_rows.p, n, rows, false);
}

}  // extern "C"

// Radices of an FFT of length n: 4, 2, 3, 5, 7, then the other primes.
static bool up_factors(int n, UpFactors &f)
{
    f.n = n;
    f.count = 0;
    int m = n;
    for (int r : {4, 2, 3, 5, 7})
        while (m > 1 && m % r == 0) {
            if (f.count == kUpMaxFactors)
                return false;
            f.r[f.count++] = r;
            m /= r;
        }
    for (int p = 11; m > 1; p += 2)
        while (m % p == 0) {
            if (f.count == kUpMaxFactors)
                return false;
            f.r[f.count++] = p;
            m /= p;
        }
    return true;
}

// One pass of upsample_lines_kernel over `lines` lines.
static int upsample_pass(mbs_ctx *ctx, const UpLines &g, int lines, const double *src,
                         double *dst)
{
    UpFactors fin, fout;
    if (!up_factors(g.n_in, fin) || !up_factors(g.n_out, fout))
        return fail(ctx, MBS_ERR_CAPACITY, "upsampling: too many factors (%d -> %d)", g.n_in,
                    g.n_out);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, ctx->up_roots_in.ensure(g.n_in));
    HIP_TRY(ctx, ctx->up_roots_out.ensure(g.n_out));
    up_roots_kernel<<<(g.n_in + 255) / 256, 256, 0, s>>>(g.n_in, ctx->up_roots_in.p);
    up_roots_kernel<<<(g.n_out + 255) / 256, 256, 0, s>>>(g.n_out, ctx->up_roots_out.p);
    HIP_TRY(ctx, hipGetLastError());
    const size_t lds = (size_t)2 * g.nbuf * sizeof(double2);
    HIP_TRY(ctx, hipFuncSetAttribute((const void *)upsample_lines_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    upsample_lines_kernel<<<lines, kUpThreads, lds, s>>>(g, fin, fout, ctx->up_roots_in.p,
                                                         ctx->up_roots_out.p, src, dst);
    HIP_TRY(ctx, hipGetLastError());
    return MBS_OK;
}

extern "C" {

int mbs_upsample_result(mbs_ctx *ctx, int noshadow, int32_t n_az_out, int32_t n_zen_out,
                        const double *theta_out, int accumulate, double *grid_out,
                        double *rows_out)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (!ctx->have_result)
        return fail(ctx, MBS_ERR_INVALID, "no result: call mbs_diffract first");
    if (noshadow && !ctx->have_ns)
        return fail(ctx, MBS_ERR_INVALID, "no-shadow result was not requested (MBS_NOSHADOW)");
    const int n_az = ctx->n_az, n_zen = ctx->n_zen1 - 1;
    if (n_az_out < n_az || n_zen_out < n_zen)
        return fail(ctx, MBS_ERR_INVALID, "upsampling %d x %d to %d x %d: sizes may only grow",
                    n_az, n_zen + 1, n_az_out, n_zen_out + 1);
    if (n_zen_out > n_zen && n_zen < 1)
        return fail(ctx, MBS_ERR_INVALID, "zenith upsampling needs at least two rows");
    if (rows_out && !theta_out)
        return fail(ctx, MBS_ERR_INVALID, "rows_out needs theta_out");
    const bool up_az = n_az_out > n_az, up_zen = n_zen_out > n_zen;
    if ((up_az && n_az_out > kUpMaxN) || (up_zen && 2 * n_zen_out > kUpMaxN))
        return fail(ctx, MBS_ERR_CAPACITY,
                    "upsampling lines are limited to %d samples (%d azimuths, 2 x %d rows "
                    "asked)", kUpMaxN, n_az_out, n_zen_out);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const double *cur = noshadow ? ctx->res_ns.p : ctx->res.p;
    const size_t mid = (size_t)n_az * (n_zen_out + 1) * 16;
    const size_t out = (size_t)n_az_out * (n_zen_out + 1) * 16;
    if (up_zen) {
        HIP_TRY(ctx, ctx->up_a.ensure(mid));
        UpLines g{1, 2 * n_zen, 2 * n_zen_out, 2 * n_zen_out, n_zen + 1, n_zen, n_zen_out};
        int rc = upsample_pass(ctx, g, n_az * 8, cur, ctx->up_a.p);
        if (rc)
            return rc;
        cur = ctx->up_a.p;
    }
    if (up_az) {
        HIP_TRY(ctx, ctx->up_b.ensure(out));
        UpLines g{0, n_az, n_az_out, n_az_out, n_zen_out + 1, 0, 0};
        int rc = upsample_pass(ctx, g, (n_zen_out + 1) * 8, cur, ctx->up_b.p);
        if (rc)
            return rc;
        cur = ctx->up_b.p;
    }
    if (rows_out) {
        const size_t n = (size_t)(n_zen_out + 1) * 18;
        HIP_TRY(ctx, ctx->up_theta.ensure(n_zen_out + 1));
        HIP_TRY(ctx, ctx->avg_rows.ensure(n));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->up_theta.p, theta_out, sizeof(double) * (n_zen_out + 1),
                                    hipMemcpyHostToDevice, ctx->stream));
        azimuth_average_kernel<<<n_zen_out + 1, 64, 0, ctx->stream>>>(
            cur, ctx->up_theta.p, n_az_out, n_zen_out, ctx->az_step * n_az / n_az_out,
            ctx->avg_rows.p);
        HIP_TRY(ctx, hipGetLastError());
        int rc = fetch(ctx, ctx->avg_rows.p, n, rows_out, false);
        if (rc)
            return rc;
    }
    if (grid_out)
        return fetch(ctx, cur, out, grid_out, accumulate != 0);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_orientation_set(mbs_ctx *ctx, int kind, int32_t n, uint32_t seed, double beta_sym,
                        double gamma_sym, double *out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_orientation_set: NULL argument");
    if (kind != MBS_ORIENT_SOBOL && kind != MBS_ORIENT_HAMMERSLEY && kind != MBS_ORIENT_SHOEMAKE)
        return fail(ctx, MBS_ERR_INVALID, "mbs_orientation_set: unknown kind %d", kind);
    if (n <= 0)
        return fail(ctx, MBS_ERR_INVALID, "orientation count must be positive");
    if (n > MBS_MAX_ORIENT_SET)
        return fail(ctx, MBS_ERR_CAPACITY, "orientation count %d exceeds %d", n,
                    MBS_MAX_ORIENT_SET);
    if (kind != MBS_ORIENT_SHOEMAKE &&
        !(beta_sym > 0.0 && beta_sym <= M_PI && gamma_sym > 0.0 && gamma_sym <= 2.0 * M_PI))
        return fail(ctx, MBS_ERR_INVALID,
                    "symmetry domain must satisfy 0 < beta <= pi and 0 < gamma <= 2*pi");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t count = (size_t)n * MBS_ORIENT_STRIDE;
    HIP_TRY(ctx, ctx->orient.ensure(count));
    orientation_set_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(
        kind, n, seed, std::cos(beta_sym), gamma_sym, ctx->orient.p);
    HIP_TRY(ctx, hipGetLastError());
    return fetch(ctx, ctx->orient.p, count, out, false);
}

int mbs_aperture_fft(mbs_ctx *ctx, const int64_t *beams, int32_t n_beams, int32_t n,
                     double pad, int32_t n_points, const double *ab, int32_t radius,
                     double *out)
{
    if (!ctx || !beams || (n_points > 0 && (!ab || !out)))
        return fail(ctx, MBS_ERR_INVALID, "mbs_aperture_fft: NULL argument");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    if (n_beams < 1 || n_beams > MBS_FFT_MAX_BATCH)
        return fail(ctx, MBS_ERR_CAPACITY, "FFT batch of %d beams (1..%d)", n_beams,
                    MBS_FFT_MAX_BATCH);
    if (n < MBS_FFT_MIN_N || n > MBS_FFT_MAX_N || (n & (n - 1)))
        return fail(ctx, MBS_ERR_INVALID, "FFT size %d: a power of two in %d..%d", n,
                    MBS_FFT_MIN_N, MBS_FFT_MAX_N);
    if (!(pad >= 1.0) || radius < 1 || radius > 8 || n_points < 0)
        return fail(ctx, MBS_ERR_INVALID, "pad >= 1, 1 <= radius <= 8, n_points >= 0");
    for (int i = 0; i < n_beams; ++i)
        if (beams[i] < 0 || beams[i] >= ctx->n_beams)
            return fail(ctx, MBS_ERR_INVALID, "beam index %lld outside the uploaded batch",
                        (long long)beams[i]);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int limit = 0;
    HIP_TRY(ctx, hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                       ctx->device));
    int j0 = 0, logl = 0;
    fft_window(n, pad, &j0, &logl);
    const size_t lds_need = std::max(fft_rows_lds(logl), fft_cols_lds(logl));
    if ((size_t)limit < lds_need)
        return fail(ctx, MBS_ERR_CAPACITY, "FFT size %d needs %zu bytes of LDS per workgroup "
                    "(device limit %d)", n, lds_need, limit);
    int logn = 0;
    while ((1 << logn) < n)
        ++logn;
    HIP_TRY(ctx, ctx->fft_grid.ensure((size_t)n_beams * n * n));
    HIP_TRY(ctx, ctx->fft_rec.ensure((size_t)n_beams * FB_STRIDE));
    HIP_TRY(ctx, ctx->fft_beams.ensure(n_beams));
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->fft_beams.p, beams, sizeof(int64_t) * n_beams,
                                hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, ctx->fft_tw.ensure(n));
    fft_twiddle_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, ctx->fft_tw.p);
    fft_setup_kernel<<<(n_beams + 63) / 64, 64, 0, s>>>(n_beams, ctx->fft_beams.p, ctx->nv.p,
                                                       ctx->voff.p, ctx->vx_in.p, ctx->vy_in.p,
                                                       n, pad, ctx->fft_rec.p);
    HIP_TRY(ctx, hipGetLastError());
    if (ctx->timing)
        for (hipEvent_t &e : ctx->fft_ev)
            if (!e)
                HIP_TRY(ctx, hipEventCreate(&e));
    int rc = MBS_OK;
    switch (logn) {
    case 6: rc = launch_fft<6, fft_min_logl(6)>(ctx, n_beams, j0, logl); break;
    case 7: rc = launch_fft<7, fft_min_logl(7)>(ctx, n_beams, j0, logl); break;
    case 8: rc = launch_fft<8, fft_min_logl(8)>(ctx, n_beams, j0, logl); break;
    case 9: rc = launch_fft<9, fft_min_logl(9)>(ctx, n_beams, j0, logl); break;
    case 10: rc = launch_fft<10, fft_min_logl(10)>(ctx, n_beams, j0, logl); break;
    case 11: rc = launch_fft<11, fft_min_logl(11)>(ctx, n_beams, j0, logl); break;
    default: return fail(ctx, MBS_ERR_INVALID, "FFT size %d", n);
    }
    if (rc)
        return rc;
    ctx->fft_n = n;
    ctx->fft_batch = n_beams;
    if (n_points > 0) {
        const size_t np = (size_t)n_beams * n_points;
        HIP_TRY(ctx, ctx->fft_ab.ensure((size_t)n_points * 2));
        HIP_TRY(ctx, ctx->fft_out.ensure(np * 2));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->fft_ab.p, ab, sizeof(double) * 2 * n_points,
                                    hipMemcpyHostToDevice, s));
        fft_sample_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(
            n_beams, n_points, n, ctx->fft_rec.p, ctx->fft_grid.p, ctx->fft_ab.p,
            ctx->consts.wave_index, radius, ctx->fft_out.p);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (ctx->timing) {
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[3], s));
        HIP_TRY(ctx, hipEventSynchronize(ctx->fft_ev[3]));
        HIP_TRY(ctx, hipEventElapsedTime(&ctx->fft_ms[0], ctx->fft_ev[0], ctx->fft_ev[1]));
        HIP_TRY(ctx, hipEventElapsedTime(&ctx->fft_ms[1], ctx->fft_ev[1], ctx->fft_ev[2]));
        HIP_TRY(ctx, hipEventElapsedTime(&ctx->fft_ms[2], ctx->fft_ev[2], ctx->fft_ev[3]));
    }
    if (n_points > 0)
        return fetch(ctx, ctx->fft_out.p, (size_t)n_beams * n_points * 2, out, false);
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return MBS_OK;
}

}  // extern "C"

// Raster + row pass and column pass of one group of spectrum beams (all of
// grid side 2^LOGN).
template <int LOGN>
static int launch_spectrum_fft(mbs_ctx *ctx, int nb, const double *rec, const int64_t *off,
                               int pmax)
{
    constexpr int N = 1 << LOGN;
    // the row kernel's LDS: the rows, then a queue slot per pixel
    const size_t lds_rows =
        (size_t)FFT_ROWS * fft_ld(N) * sizeof(float2) + (size_t)FFT_ROWS * N * sizeof(int);
    const size_t lds_cols = (size_t)FFT_COLS * fft_ld(N) * sizeof(float2);
    HIP_TRY(ctx, hipFuncSetAttribute((const void *)spectrum_rows_kernel<LOGN>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rows));
    HIP_TRY(ctx, hipFuncSetAttribute((const void *)spectrum_cols_kernel<LOGN>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cols));
    hipStream_t s = ctx->stream;
    // band columns per beam: the group's widest band, in units of FFT_COLS
    const int units = (pmax + FFT_COLS - 1) / FFT_COLS;
    const int rb = (nb * (N / FFT_ROWS) + 7) / 8 * 8;
    const int cb = (nb * units + 7) / 8 * 8;
    spectrum_rows_kernel<LOGN><<<rb, FFT_ROW_THREADS, lds_rows, s>>>(
        nb, rec, ctx->spec_deap.p, ctx->fft_tw.p, ctx->spec_cph.p, off, ctx->spec_grid.p);
    HIP_TRY(ctx, hipGetLastError());
    spectrum_cols_kernel<LOGN><<<cb, FFT_COL_THREADS, lds_cols, s>>>(
        nb, units, rec, ctx->fft_tw.p, ctx->spec_cph.p, off, ctx->spec_grid.p);
    HIP_TRY(ctx, hipGetLastError());
    return MBS_OK;
}

extern "C" {

int mbs_spectrum_source(mbs_ctx *ctx, const mbs_spectrum_params *p, mbs_spectrum_info *info)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    ctx->spec_ok = false;
    ctx->have_result = false;
    if (!p)
        return MBS_OK;
    if (!ctx->grid_ok)
        return fail(ctx, MBS_ERR_NO_GRID, "mbs_set_grid has not been called");
    if (!ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "mbs_upload_beams has not been called");
    const auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!(p->a_max > 0.0) || !(p->oversample >= 2.0) || !pow2(p->min_n) || !pow2(p->max_n) ||
        p->min_n < MBS_FFT_MIN_N || p->max_n > MBS_FFT_MAX_N || p->min_n > p->max_n)
        return fail(ctx, MBS_ERR_INVALID,
                    "spectrum source: a_max > 0, oversample >= 2, min_n <= max_n powers of two "
                    "in %d..%d", MBS_FFT_MIN_N, MBS_FFT_MAX_N);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int limit = 0;
    HIP_TRY(ctx, hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                       ctx->device));
    if ((size_t)limit < (size_t)FFT_COLS * fft_ld(p->max_n) * sizeof(float2))
        return fail(ctx, MBS_ERR_CAPACITY, "FFT side %d needs more LDS than the device has",
                    p->max_n);
    int rc = ensure_tables(ctx);
    if (rc)
        return rc;
    mbs_spectrum_info inf{};
    const int nb = ctx->n_beams, cells = ctx->cells;
    const double k = ctx->consts.wave_index;
    hipStream_t s = ctx->stream;
    for (hipEvent_t &e : ctx->fft_ev)
        if (!e)
            HIP_TRY(ctx, hipEventCreate(&e));
    HIP_TRY(ctx, ctx->spec_sidx.ensure(nb));
    std::vector<int32_t> sidx(nb, -1);
    std::vector<double> scan((size_t)nb * 5);
    std::vector<int32_t> nv(nb);
    std::vector<double> area(nb);
    if (nb > 0) {
        HIP_TRY(ctx, ctx->spec_scan.ensure((size_t)nb * 5));
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[0], s));
        spectrum_scan_kernel<<<nb, 256, 0, s>>>(nb, ctx->hdr.p, ctx->tab.p, cells, ctx->nv.p,
                                                ctx->voff.p, ctx->vx_in.p, ctx->vy_in.p,
                                                ctx->pack_scale, ctx->spec_scan.p);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[1], s));
        HIP_TRY(ctx, hipMemcpyAsync(scan.data(), ctx->spec_scan.p, scan.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(nv.data(), ctx->nv.p, nb * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
        // the aperture areas (field 12 of the reference's beam scalars)
        HIP_TRY(ctx, hipMemcpy2DAsync(area.data(), sizeof(double), ctx->bs_in.p + 12,
                                      MBS_BEAM_SCALARS * sizeof(double), sizeof(double), nb,
                                      hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        HIP_TRY(ctx, hipEventElapsedTime(&inf.scan_ms, ctx->fft_ev[0], ctx->fft_ev[1]));
    }
    // selection, grouped by FFT side (slots in ascending side order)
    struct Pick {
        int beam, logn, band;
        double h;
    };
    std::vector<Pick> picks;
    for (int b = 0; b < nb; ++b) {
        const double *sc = &scan[(size_t)b * 5];
        const double amax = sc[0], ext = sc[1], width = sc[4];
        area[b] *= ctx->pack_scale * ctx->pack_scale;
        inf.total_area += std::fabs(area[b]);
        if (sc[2] == 0.0 || nv[b] > kSpecMaxNv || !(amax <= p->a_max) || !(amax > 0.0) ||
            !(width > 0.0))
            continue;
        for (double os = p->oversample; os >= 2.0; os *= 0.5) {
            // enough pixels across the polygon in every direction (see
            // kSpecPixelsAcross)
            const double h = std::min({kPi / (k * amax * os), ext / kSpecPixelsAcross,
                                       width / kSpecPixelsWidth});
            const double need = 2.0 * (ext + 4.0 * h) / h;
            int logn = 0;
            while ((1 << logn) < need || (1 << logn) < p->min_n)
                ++logn;
            if ((1 << logn) <= p->max_n) {
                // the band of bins the taps reach (see spectrum_kernels.hip)
                const double umax = k * amax * (double)(1 << logn) * h / (2.0 * kPi);
                const int band = (int)std::ceil(umax + 0.5 * kSpecW) + 1;
                picks.push_back(Pick{b, logn, band, h});
                inf.selected_area += std::fabs(area[b]);
                break;
            }
        }
    }
    std::stable_sort(picks.begin(), picks.end(),
                     [](const Pick &x, const Pick &y) { return x.logn < y.logn; });
    const int ns = (int)picks.size();
    std::vector<int32_t> sb(ns), sl(ns), sm(ns), sp(ns);
    std::vector<double> sh(ns);
    std::vector<int64_t> so(ns);
    int64_t total = 0;
    int max_n = 0;
    for (int i = 0; i < ns; ++i) {
        const int n = 1 << picks[i].logn;
        sb[i] = picks[i].beam;
        sl[i] = picks[i].logn;
        sh[i] = picks[i].h;
        sm[i] = picks[i].band;
        sp[i] = 2 * sm[i] + 1 < n ? 2 * sm[i] + 1 : n;  // as spectrum_setup_kernel
        so[i] = total;
        sidx[picks[i].beam] = i;
        total += (int64_t)n * sp[i];  // [N][P] for the row pass, [P][P] kept
        max_n = std::max(max_n, 1 << picks[i].logn);
        ++inf.n_by_size[picks[i].logn - 6];
    }
    inf.n_selected = ns;
    inf.spectrum_bytes = total * (int64_t)sizeof(float2);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_sidx.p, sidx.data(), nb * sizeof(int32_t),
                                hipMemcpyHostToDevice, s));
    if (ns > 0) {
        HIP_TRY(ctx, ctx->spec_beam.ensure(ns));
        HIP_TRY(ctx, ctx->spec_logn.ensure(ns));
        HIP_TRY(ctx, ctx->spec_band.ensure(ns));
        HIP_TRY(ctx, ctx->spec_cph.ensure(max_n + 1));
        HIP_TRY(ctx, ctx->spec_h.ensure(ns));
        HIP_TRY(ctx, ctx->spec_off.ensure(ns));
        HIP_TRY(ctx, ctx->spec_rec.ensure((size_t)ns * SP_STRIDE));
        HIP_TRY(ctx, ctx->spec_grid.ensure((size_t)total));
        HIP_TRY(ctx, ctx->spec_f.ensure((size_t)ns * cells));
        HIP_TRY(ctx, ctx->fft_tw.ensure(max_n));
        HIP_TRY(ctx, ctx->spec_deap.ensure(max_n));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_beam.p, sb.data(), ns * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_logn.p, sl.data(), ns * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_h.p, sh.data(), ns * sizeof(double),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_band.p, sm.data(), ns * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->spec_off.p, so.data(), ns * sizeof(int64_t),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[1], s));
        spectrum_setup_kernel<<<(ns + 63) / 64, 64, 0, s>>>(ns, ctx->spec_beam.p, ctx->spec_h.p,
                                                            ctx->spec_logn.p, ctx->spec_band.p,
                                                            ctx->nv.p, ctx->voff.p,
                                                            ctx->vx_in.p, ctx->vy_in.p,
                                                            ctx->pack_scale, ctx->spec_rec.p);
        HIP_TRY(ctx, hipGetLastError());
        for (int i0 = 0; i0 < ns;) {
            int i1 = i0;
            while (i1 < ns && sl[i1] == sl[i0])
                ++i1;
            const int n = 1 << sl[i0];
            int pmax = 0;
            for (int i = i0; i < i1; ++i)
                pmax = std::max(pmax, sp[i]);
            fft_twiddle_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, ctx->fft_tw.p);
            spectrum_deap_kernel<<<(n + 3) / 4, 256, 0, s>>>(n, ctx->spec_deap.p);
            spectrum_phase_kernel<<<(n + 256) / 256, 256, 0, s>>>(n, ctx->spec_cph.p);
            HIP_TRY(ctx, hipGetLastError());
            const double *rec = ctx->spec_rec.p + (size_t)i0 * SP_STRIDE;
            const int64_t *off = ctx->spec_off.p + i0;
            const int g = i1 - i0;
            switch (sl[i0]) {
            case 6: rc = launch_spectrum_fft<6>(ctx, g, rec, off, pmax); break;
            case 7: rc = launch_spectrum_fft<7>(ctx, g, rec, off, pmax); break;
            case 8: rc = launch_spectrum_fft<8>(ctx, g, rec, off, pmax); break;
            case 9: rc = launch_spectrum_fft<9>(ctx, g, rec, off, pmax); break;
            case 10: rc = launch_spectrum_fft<10>(ctx, g, rec, off, pmax); break;
            case 11: rc = launch_spectrum_fft<11>(ctx, g, rec, off, pmax); break;
            default: rc = fail(ctx, MBS_ERR_INVALID, "FFT side %d", n);
            }
            if (rc)
                return rc;
            i0 = i1;
        }
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[2], s));
        const size_t units = (size_t)ns * cells;
        spectrum_sample_kernel<<<(unsigned)((units + 255) / 256), 256, 0, s>>>(
            ns, cells, ctx->spec_beam.p, ctx->spec_off.p, ctx->spec_logn.p, ctx->spec_rec.p,
            ctx->spec_grid.p, ctx->hdr.p, ctx->tab.p, k, ctx->spec_f.p);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->fft_ev[3], s));
        HIP_TRY(ctx, hipEventSynchronize(ctx->fft_ev[3]));
        HIP_TRY(ctx, hipEventElapsedTime(&inf.fft_ms, ctx->fft_ev[1], ctx->fft_ev[2]));
        HIP_TRY(ctx, hipEventElapsedTime(&inf.sample_ms, ctx->fft_ev[2], ctx->fft_ev[3]));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->spec_n = ns;
    ctx->spec_ok = true;
    if (info)
        *info = inf;
    return MBS_OK;
}

int mbs_spectrum_values(mbs_ctx *ctx, int32_t *beams, float *f)
{
    if (!ctx || !beams)
        return fail(ctx, MBS_ERR_INVALID, "mbs_spectrum_values: NULL argument");
    if (!ctx->spec_ok)
        return fail(ctx, MBS_ERR_INVALID, "no spectrum source (mbs_spectrum_source)");
    const int ns = ctx->spec_n;
    if (ns == 0)
        return MBS_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(beams, ctx->spec_beam.p, ns * sizeof(int32_t),
                                hipMemcpyDeviceToHost, ctx->stream));
    if (f)
        HIP_TRY(ctx, hipMemcpyAsync(f, ctx->spec_f.p, (size_t)ns * ctx->cells * sizeof(float2),
                                    hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_aperture_fft_spectrum(mbs_ctx *ctx, int32_t index, float *out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_aperture_fft_spectrum: NULL argument");
    if (index < 0 || index >= ctx->fft_batch)
        return fail(ctx, MBS_ERR_INVALID, "no spectrum %d (last batch holds %d)", index,
                    ctx->fft_batch);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n2 = (size_t)ctx->fft_n * ctx->fft_n;
    HIP_TRY(ctx, hipMemcpyAsync(out, ctx->fft_grid.p + (size_t)index * n2, n2 * sizeof(float2),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_aperture_fft_timing(mbs_ctx *ctx, float *rows_ms, float *cols_ms, float *sample_ms)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (!ctx->timing)
        return fail(ctx, MBS_ERR_INVALID, "timing is not enabled (mbs_set_timing)");
    if (rows_ms)
        *rows_ms = ctx->fft_ms[0];
    if (cols_ms)
        *cols_ms = ctx->fft_ms[1];
    if (sample_ms)
        *sample_ms = ctx->fft_ms[2];
    return MBS_OK;
}

int mbs_device_result(mbs_ctx *ctx, const double **m_full, const double **m_noshadow)
{
    if (!ctx || !ctx->have_result)
        return fail(ctx, MBS_ERR_INVALID, "no result on the device");
    if (m_full)
        *m_full = ctx->res.p;
    if (m_noshadow)
        *m_noshadow = ctx->have_ns ? ctx->res_ns.p : nullptr;
    return MBS_OK;
}

int mbs_batch_stats(mbs_ctx *ctx, int64_t *n_beams, double *units, int64_t *nv_hist32,
                    double *mean_edges)
{
    if (!ctx || !ctx->beams_ok)
        return fail(ctx, MBS_ERR_NO_BEAMS, "no beams uploaded");
    if (n_beams)
        *n_beams = ctx->n_beams;
    if (units)
        *units = (double)ctx->n_beams * (ctx->grid_ok ? ctx->cells : 0);
    if (nv_hist32)
        std::memcpy(nv_hist32, ctx->nv_hist, sizeof ctx->nv_hist);
    if (mean_edges)
        *mean_edges = ctx->mean_edges;
    return MBS_OK;
}

int mbs_set_work_mode(mbs_ctx *ctx, int mode)
{
    if (!ctx || mode < 0 || mode > 3)
        return fail(ctx, MBS_ERR_INVALID, "work mode must be 0, 1, 2 or 3");
    ctx->work_mode = mode;
    ctx->tuned[0].valid = ctx->tuned[1].valid = false;
    return MBS_OK;
}

int mbs_set_target_blocks(mbs_ctx *ctx, int blocks)
{
    if (!ctx || blocks < 1)
        return fail(ctx, MBS_ERR_INVALID, "bad target block count");
    ctx->target_blocks = blocks;
    ctx->tuned[0].valid = ctx->tuned[1].valid = false;
    return MBS_OK;
}

int mbs_set_graph(mbs_ctx *ctx, int enable)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    ctx->use_graph = enable != 0;
    if (!ctx->use_graph && ctx->graph_exec) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
        ctx->graph_key.clear();
    }
    return MBS_OK;
}

int mbs_graph_stats(mbs_ctx *ctx, int64_t *captures, int64_t *launches)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (captures)
        *captures = ctx->graph_captures;
    if (launches)
        *launches = ctx->graph_replays;
    return MBS_OK;
}

int mbs_autotune(mbs_ctx *ctx, uint32_t flags, mbs_tune_entry *entries, int32_t max_entries,
                 int32_t *n_entries)
{
    if (!ctx)
        return fail(nullptr, MBS_ERR_INVALID, "ctx is NULL");
    if (flags & ~(uint32_t)(MBS_PREC_MIXED | MBS_NOSHADOW))
        return fail(ctx, MBS_ERR_INVALID, "mbs_autotune takes MBS_PREC_MIXED and MBS_NOSHADOW");
    if (!ctx->grid_ok || !ctx->beams_ok)
        return fail(ctx, ctx->grid_ok ? MBS_ERR_NO_BEAMS : MBS_ERR_NO_GRID,
                    "mbs_autotune needs a grid and an uploaded batch");
    const int base = (flags & MBS_PREC_MIXED) ? 1 : 0;
    const bool had_timing = ctx->timing;
    int rc = mbs_set_timing(ctx, 1);
    if (rc)
        return rc;
    static const int kBlocks[] = {4096, 8192, 16384, 32768, 65536, 131072};
    const int n_tiles = (ctx->cells + 255) / 256;
    std::vector<mbs_tune_entry> tried;
    mbs_ctx::Tuned best;
    float best_ms = 0.0f;
    // occupancy variants of the MFMA kernel with beam segments
    struct Cand {
        int mfma, mode, waves;
    };
    std::vector<Cand> cands;
    for (int mfma = (base ? 1 : 0); mfma >= 0; --mfma)
        for (int mode = 3; mode >= 1; --mode)
            for (int w : {0, 4, 6})
                if (w == 0 || (mfma && mode == MBS_WORK_SEGMENTS && !(flags & MBS_NOSHADOW)))
                    cands.push_back(Cand{mfma, mode, w});
    for (const Cand &cd : cands) {
        const int mfma = cd.mfma, mode = cd.mode;
            for (int tb : kBlocks) {
                std::vector<WorkItem> items;
                const bool jones = plan_items(ctx, n_tiles, tb,
                                              mode == MBS_WORK_QUEUE ? MBS_WORK_SEGMENTS : mode,
                                              items);
                const size_t n_it = mode == MBS_WORK_QUEUE
                                        ? (size_t)((tb + n_tiles - 1) / n_tiles + ctx->n_orient)
                                        : items.size();
                if (jones && n_it * (size_t)ctx->cells * 8 * sizeof(double) > kPartialBudget)
                    continue;
                mbs_ctx::Tuned t;
                t.valid = true;
                t.target_blocks = tb;
                t.work_mode = mode;
                t.mfma = mfma;
                t.waves = cd.waves;
                ctx->tuned[base] = t;
                // the first run allocates and warms the caches; the shape's
                // time is the least of three more (the shapes within a few
                // per cent of each other are told apart by one run only by
                // its noise)
                float ms = 0.0f;
                for (int rep = 0; rep < 4 && !rc; ++rep) {
                    rc = mbs_diffract(ctx, flags);
                    mbs_timing tm;
                    if (!rc)
                        rc = mbs_get_timing(ctx, &tm);
                    const float t = tm.kernel_ms + tm.reduce_ms;
                    if (rep == 1 || (rep > 1 && t < ms))
                        ms = t;
                }
                if (rc)
                    break;
                mbs_tune_entry e;
                std::memset(&e, 0, sizeof e);
                e.target_blocks = tb;
                e.work_mode = mode;
                e.mfma = mfma;
                e.ms = ms;
                e.waves = cd.waves ? cd.waves : (mfma ? MBS_MFMA_WAVES : 0);
                tried.push_back(e);
                if (!best.valid || ms < best_ms) {
                    best = t;
                    best_ms = ms;
                }
            }
    }
    ctx->tuned[base] = best;
    if (!had_timing)
        mbs_set_timing(ctx, 0);
    if (rc)
        return rc;
    if (n_entries)
        *n_entries = (int32_t)tried.size();
    if (entries)
        for (size_t i = 0; i < tried.size() && (int32_t)i < max_entries; ++i)
            entries[i] = tried[i];
    return MBS_OK;
}

int mbs_apply_tune_entry(mbs_ctx *ctx, uint32_t flags, const mbs_tune_entry *e)
{
    if (!ctx || !e)
        return fail(ctx, MBS_ERR_INVALID, "mbs_apply_tune_entry: NULL argument");
    if (flags & ~(uint32_t)(MBS_PREC_MIXED | MBS_NOSHADOW))
        return fail(ctx, MBS_ERR_INVALID,
                    "mbs_apply_tune_entry takes MBS_PREC_MIXED and MBS_NOSHADOW");
    const int base = (flags & MBS_PREC_MIXED) ? 1 : 0;
    const bool waves_ok = e->waves == 0 || e->waves == MBS_MFMA_WAVES ||
                          ((e->waves == 4 || e->waves == 6) && base && e->mfma &&
                           e->work_mode == MBS_WORK_SEGMENTS);
    if (e->target_blocks < 1 || e->work_mode < 1 || e->work_mode > 3 || e->mfma < 0 ||
        e->mfma > 1 || (e->mfma && !base) || !waves_ok)
        return fail(ctx, MBS_ERR_INVALID, "mbs_apply_tune_entry: shape (%d, %d, %d, %d)",
                    e->target_blocks, e->work_mode, e->mfma, e->waves);
    mbs_ctx::Tuned t;
    t.valid = true;
    t.target_blocks = e->target_blocks;
    t.work_mode = e->work_mode;
    t.mfma = e->mfma;
    t.waves = (e->waves == 4 || e->waves == 6) ? e->waves : 0;
    ctx->tuned[base] = t;
    return MBS_OK;
}

}  // extern "C"

// ---- learned control variate: the MLP ------------------------------------

struct mbs_mlp {
    mbs_ctx *ctx = nullptr;
    int n_in = 0, width = 0, depth = 0, n_out = 0;
    std::vector<int64_t> w_off, b_off;  // parameter offsets per layer
    int64_t n_params = 0;
    DevBuf<float> params, grads, adam_m, adam_v;
    int64_t adam_t = 0;
    // activations of the last forward pass: Z_l (pre-activation), A_l (input of layer l)
    int cap_rows = 0;
    std::vector<DevBuf<float>> z, a;
    DevBuf<float> y, t, dy, da, dz, row_sq, x_all, t_all;
    DevBuf<int32_t> idx;
};

namespace {

int mlp_layer_in(const mbs_mlp *m, int l) { return l == 0 ? m->n_in : m->width; }
int mlp_layer_out(const mbs_mlp *m, int l) { return l == m->depth ? m->n_out : m->width; }

int mlp_ensure_rows(mbs_mlp *m, int rows)
{
    mbs_ctx *ctx = m->ctx;
    if (rows <= m->cap_rows)
        return MBS_OK;
    m->z.resize(m->depth);
    m->a.resize(m->depth + 1);
    for (int l = 0; l <= m->depth; ++l)
        HIP_TRY(ctx, m->a[l].ensure((size_t)rows * mlp_layer_in(m, l)));
    for (int l = 0; l < m->depth; ++l)
        HIP_TRY(ctx, m->z[l].ensure((size_t)rows * m->width));
    HIP_TRY(ctx, m->y.ensure((size_t)rows * m->n_out));
    HIP_TRY(ctx, m->t.ensure((size_t)rows * m->n_out));
    HIP_TRY(ctx, m->dy.ensure((size_t)rows * m->n_out));
    HIP_TRY(ctx, m->da.ensure((size_t)rows * m->width));
    HIP_TRY(ctx, m->dz.ensure((size_t)rows * m->width));
    HIP_TRY(ctx, m->row_sq.ensure(rows));
    m->cap_rows = rows;
    return MBS_OK;
}

template <int TA, int TB>
void mlp_gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B,
              int ldb, float *C, int ldc, const float *bias, int beta)
{
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    mlp_gemm_kernel<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, beta);
}

// Forward pass of `rows` inputs already in a[0]; output in y.
int mlp_forward_dev(mbs_mlp *m, int rows)
{
    mbs_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const float *P = m->params.p;
    for (int l = 0; l <= m->depth; ++l) {
        const int in = mlp_layer_in(m, l), out = mlp_layer_out(m, l);
        float *dst = l < m->depth ? m->z[l].p : m->y.p;
        mlp_gemm<0, 0>(s, rows, out, in, m->a[l].p, in, P + m->w_off[l], out, dst, out,
                       P + m->b_off[l], 0);
        if (l < m->depth) {
            const size_t n = (size_t)rows * out;
            mlp_gelu_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, m->z[l].p,
                                                                        m->a[l + 1].p);
        }
    }
    HIP_TRY(ctx, hipGetLastError());
    return MBS_OK;
}

// Loss (sum of squares) partials and parameter gradients of the batch whose
// forward pass is in place and whose targets are in t.
int mlp_backward_dev(mbs_mlp *m, int rows)
{
    mbs_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const float *P = m->params.p;
    float *G = m->grads.p;
    mlp_mse_grad_kernel<<<rows, 256, 0, s>>>(rows, m->n_out, m->y.p, m->t.p, m->dy.p,
                                             m->row_sq.p);
    const float *dcur = m->dy.p;  // gradient w.r.t. the output of layer l
    for (int l = m->depth; l >= 0; --l) {
        const int in = mlp_layer_in(m, l), out = mlp_layer_out(m, l);
        const float *dout = dcur;
        if (l < m->depth) {
            const size_t n = (size_t)rows * out;
            mlp_gelu_grad_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, m->z[l].p, dcur,
                                                                             m->dz.p);
            dout = m->dz.p;
        }
        // dW_l = A_l^T dOut, db_l = column sums of dOut
        mlp_gemm<1, 0>(s, in, out, rows, m->a[l].p, in, dout, out, G + m->w_off[l], out,
                       nullptr, 0);
        mlp_colsum_kernel<<<(out + 255) / 256, 256, 0, s>>>(rows, out, dout, G + m->b_off[l]);
        if (l > 0) {  // dA_l = dOut W_l^T
            mlp_gemm<0, 1>(s, rows, in, out, dout, out, P + m->w_off[l], out, m->da.p, in,
                           nullptr, 0);
            dcur = m->da.p;
        }
    }
    HIP_TRY(ctx, hipGetLastError());
    return MBS_OK;
}

uint64_t mlp_splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int mbs_mlp_create(mbs_ctx *ctx, int32_t n_in, int32_t width, int32_t depth, int32_t n_out,
                   uint64_t seed, mbs_mlp **out)
{
    if (!ctx || !out)
        return fail(ctx, MBS_ERR_INVALID, "mbs_mlp_create: NULL argument");
    *out = nullptr;
    if (n_in < 1 || n_in > 64 || width < 8 || width > 4096 || depth < 1 || depth > 16 ||
        n_out < 1 || n_out > 65536)
        return fail(ctx, MBS_ERR_INVALID, "MLP shape out of range");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    mbs_mlp *m = new mbs_mlp();
    m->ctx = ctx;
    m->n_in = n_in;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    int64_t off = 0;
    for (int l = 0; l <= depth; ++l) {
        m->w_off.push_back(off);
        off += (int64_t)mlp_layer_in(m, l) * mlp_layer_out(m, l);
        m->b_off.push_back(off);
        off += mlp_layer_out(m, l);
    }
    m->n_params = off;
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from splitmix64
    std::vector<float> h(off, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l <= depth; ++l) {
        const int in = mlp_layer_in(m, l), o = mlp_layer_out(m, l);
        const double bound = std::sqrt(6.0 / in) * (l == depth ? 0.5 : 1.0);
        for (int64_t i = 0; i < (int64_t)in * o; ++i) {
            const double u = (double)(mlp_splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            h[m->w_off[l] + i] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    hipError_t e = m->params.ensure(off);
    if (e == hipSuccess)
        e = m->grads.ensure(off);
    if (e == hipSuccess)
        e = m->adam_m.ensure(off);
    if (e == hipSuccess)
        e = m->adam_v.ensure(off);
    if (e == hipSuccess)
        e = hipMemcpy(m->params.p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(m->adam_m.p, 0, sizeof(float) * off);
    if (e == hipSuccess)
        e = hipMemset(m->adam_v.p, 0, sizeof(float) * off);
    if (e != hipSuccess) {
        delete m;
        return fail(ctx, MBS_ERR_DEVICE, "MLP allocation: %s", hipGetErrorString(e));
    }
    *out = m;
    return MBS_OK;
}

int mbs_mlp_destroy(mbs_mlp *m)
{
    if (!m)
        return MBS_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    for (auto &b : m->z)
        b.release();
    for (auto &b : m->a)
        b.release();
    DevBuf<float> *bufs[] = {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->y, &m->t,
                             &m->dy, &m->da, &m->dz, &m->row_sq, &m->x_all, &m->t_all};
    for (DevBuf<float> *b : bufs)
        b->release();
    m->idx.release();
    delete m;
    return MBS_OK;
}

int64_t mbs_mlp_param_count(const mbs_mlp *m) { return m ? m->n_params : -1; }

int mbs_mlp_set_params(mbs_mlp *m, const float *params)
{
    if (!m || !params)
        return fail(m ? m->ctx : nullptr, MBS_ERR_INVALID, "mbs_mlp_set_params: NULL argument");
    mbs_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(m->params.p, params, sizeof(float) * m->n_params,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_mlp_get_params(mbs_mlp *m, float *params)
{
    if (!m || !params)
        return fail(m ? m->ctx : nullptr, MBS_ERR_INVALID, "mbs_mlp_get_params: NULL argument");
    mbs_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(params, m->params.p, sizeof(float) * m->n_params,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_mlp_forward(mbs_mlp *m, const float *x, int32_t rows, float *y)
{
    if (!m || !x || !y || rows < 1)
        return fail(m ? m->ctx : nullptr, MBS_ERR_INVALID, "mbs_mlp_forward: bad argument");
    mbs_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = mlp_ensure_rows(m, rows);
    if (rc)
        return rc;
    HIP_TRY(ctx, hipMemcpyAsync(m->a[0].p, x, sizeof(float) * rows * m->n_in,
                                hipMemcpyHostToDevice, ctx->stream));
    rc = mlp_forward_dev(m, rows);
    if (rc)
        return rc;
    HIP_TRY(ctx, hipMemcpyAsync(y, m->y.p, sizeof(float) * rows * m->n_out,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBS_OK;
}

int mbs_mlp_gradient(mbs_mlp *m, const float *x, const float *t, int32_t rows, double *loss,
                     float *grad)
{
    if (!m || !x || !t || rows < 1)
        return fail(m ? m->ctx : nullptr, MBS_ERR_INVALID, "mbs_mlp_gradient: bad argument");
    mbs_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = mlp_ensure_rows(m, rows);
    if (rc)
        return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(m->a[0].p, x, sizeof(float) * rows * m->n_in,
                                hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(m->t.p, t, sizeof(float) * rows * m->n_out,
                                hipMemcpyHostToDevice, s));
    rc = mlp_forward_dev(m, rows);
    if (!rc)
        rc = mlp_backward_dev(m, rows);
    if (rc)
        return rc;
    std::vector<float> sq(rows);
    HIP_TRY(ctx, hipMemcpyAsync(sq.data(), m->row_sq.p, sizeof(float) * rows,
                                hipMemcpyDeviceToHost, s));
    if (grad)
        HIP_TRY(ctx, hipMemcpyAsync(grad, m->grads.p, sizeof(float) * m->n_params,
                                    hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (loss) {
        double acc = 0.0;
        for (float v : sq)
            acc += v;
        *loss = acc / ((double)rows * m->n_out);
    }
    return MBS_OK;
}

int mbs_mlp_train(mbs_mlp *m, const float *x, const float *t, int32_t n, int32_t batch,
                  int32_t epochs, float lr, uint64_t seed, double *final_loss)
{
    if (!m || !x || !t || n < 1 || batch < 1 || batch > n || epochs < 0 || !(lr > 0.0f))
        return fail(m ? m->ctx : nullptr, MBS_ERR_INVALID, "mbs_mlp_train: bad argument");
    mbs_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = mlp_ensure_rows(m, batch);
    if (rc)
        return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, m->x_all.ensure((size_t)n * m->n_in));
    HIP_TRY(ctx, m->t_all.ensure((size_t)n * m->n_out));
    HIP_TRY(ctx, m->idx.ensure(n));
    HIP_TRY(ctx, hipMemcpyAsync(m->x_all.p, x, sizeof(float) * n * m->n_in,
                                hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(m->t_all.p, t, sizeof(float) * n * m->n_out,
                                hipMemcpyHostToDevice, s));
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; ++i)
        perm[i] = i;
    uint64_t st = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0.0;
    std::vector<float> sq(batch);
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = n - 1; i > 0; --i)  // Fisher-Yates, deterministic in seed
            std::swap(perm[i], perm[(int)(mlp_splitmix(st) % (uint64_t)(i + 1))]);
        HIP_TRY(ctx, hipMemcpyAsync(m->idx.p, perm.data(), sizeof(int32_t) * n,
                                    hipMemcpyHostToDevice, s));
        double ep_sum = 0.0;
        int ep_rows = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            const size_t nx = (size_t)batch * m->n_in, nt = (size_t)batch * m->n_out;
            mlp_gather_kernel<<<(unsigned)((nx + 255) / 256), 256, 0, s>>>(
                batch, m->n_in, m->idx.p + b0, m->x_all.p, m->a[0].p);
            mlp_gather_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(
                batch, m->n_out, m->idx.p + b0, m->t_all.p, m->t.p);
            rc = mlp_forward_dev(m, batch);
            if (!rc)
                rc = mlp_backward_dev(m, batch);
            if (rc)
                return rc;
            ++m->adam_t;
            const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
            const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
            mlp_adam_kernel<<<(unsigned)((m->n_params + 255) / 256), 256, 0, s>>>(
                (size_t)m->n_params, m->params.p, m->grads.p, m->adam_m.p, m->adam_v.p, lr, b1,
                b2, eps, c1, c2);
            HIP_TRY(ctx, hipGetLastError());
            if (ep == epochs - 1) {
                HIP_TRY(ctx, hipMemcpyAsync(sq.data(), m->row_sq.p, sizeof(float) * batch,
                                            hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                for (float v : sq)
                    ep_sum += v;
                ep_rows += batch;
            }
        }
        if (ep == epochs - 1 && ep_rows)
            last = ep_sum / ((double)ep_rows * m->n_out);
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (final_loss)
        *final_loss = last;
    return MBS_OK;
}

}  // extern "C"
