/*
 * mbs_hip.h -- C ABI of the MI355X (gfx950) physical-optics diffraction engine.
 *
 * The engine evaluates, for every prepared beam of every orientation and every
 * scattering direction (theta, phi) of a grid, the polygon-aperture Fourier
 * integral, sums the 2x2 Jones matrices of all beams of one orientation
 * coherently, converts the sum to a 4x4 Mueller matrix and accumulates
 * weight * Mueller over orientations.  This is the work the reference does in
 * HandlerPO::HandleBeamsToLocal + HandlerPO::AddToMuellerLocal on the CPU
 * (src/handler/HandlerPO.cpp:2561-2796, :2888-2907) and in
 * HandlerPO::HandleOrientationsToLocalGpu on CUDA
 * (src/cuda/GpuDiffraction.cu:4159-4778).
 *
 * Only plain C types cross this boundary: pointers, sizes and POD structs.  All
 * entry points return an mbs_status; mbs_last_error() returns the message of
 * the last failure on that context (or of the last context-less failure when
 * ctx is NULL).  A context is bound to one HIP device and is not safe to use
 * from two host threads at once; create one context per host thread.
 */
#ifndef MBS_HIP_H
#define MBS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBS_ABI_VERSION 1

typedef enum mbs_status {
    MBS_OK = 0,
    MBS_ERR_INVALID = 1,     /* bad argument or inconsistent sizes          */
    MBS_ERR_BAD_BEAM = 2,    /* beam outside what the kernels accept        */
    MBS_ERR_DEVICE = 3,      /* HIP runtime error                           */
    MBS_ERR_NO_GRID = 4,     /* mbs_set_grid has not been called            */
    MBS_ERR_NO_BEAMS = 5,    /* mbs_upload_beams has not been called        */
    MBS_ERR_CAPACITY = 6     /* request exceeds a documented maximum        */
} mbs_status;

/* Columns of one row of mbs_beam_batch.beam_scalars: the scalar fields of the
 * reference's PreparedBeam (src/handler/HandlerPO.h:14-46), in this order. */
enum {
    MBS_B_HOR = 0,    /* aperture horizontal axis (3)        horAx..horAz   */
    MBS_B_VER = 3,    /* aperture vertical axis (3)          verAx..verAz   */
    MBS_B_DIR = 6,    /* beam propagation direction (3)      bdx..bdz       */
    MBS_B_CEN = 9,    /* beam centre (3)                     cenx..cenz     */
    MBS_B_AREA = 12,  /* beam area                           beam_area      */
    MBS_B_NT = 13,    /* normal x beamBasis (3)              pNTx..pNTz     */
    MBS_B_NP = 16,    /* normal x polarisation basis (3)     pNPx..pNPz     */
    MBS_B_NXDT = 19,  /* normal x (dir x beamBasis) (3)      pnxDTx..       */
    MBS_B_NXDP = 22,  /* normal x (dir x polarisation) (3)   pnxDPx..       */
    MBS_B_JP = 25,    /* J_phased, 8 reals: 00r 00i 01r 01i 10r 10i 11r 11i */
    MBS_B_NORMAL = 33,/* aperture normal (3), informational                 */
    MBS_BEAM_SCALARS = 36
};

/* Aperture-edge flags, one byte per vertex/edge (edge e joins vertex e and
 * vertex (e+1) mod nv), the edge_valid_x / edge_valid_y of BeamEdgeData
 * (src/handler/Handler.h:372-388). */
#define MBS_EDGE_VALID_X 1u
#define MBS_EDGE_VALID_Y 2u

/* Vertex limits of the kernels, the same as the reference GPU path
 * (GpuDiffraction.cu:4301-4303): 3 <= nv <= MBS_MAX_VERTICES. */
#define MBS_MIN_VERTICES 3
#define MBS_MAX_VERTICES 31

/* Prepared beams of a batch of orientations, as flat host arrays.  Replaces
 * std::vector<PreparedOrientation> (src/handler/HandlerPO.h:49-55) plus the
 * packing of pack_prepared_gpu_beam (src/cuda/GpuDiffraction.cu:273-358).
 * An orientation holds at most MBS_MAX_BEAMS_PER_ORIENT beams (the upload
 * returns MBS_ERR_CAPACITY beyond it). */
#define MBS_MAX_BEAMS_PER_ORIENT 8388607
typedef struct mbs_beam_batch {
    int32_t n_orient;
    int32_t n_beams;
    const int64_t *orient_offsets; /* [n_orient+1] beam ranges per orientation */
    const double *weights;         /* [n_orient] orientation weight (sinZenith) */
    const double *beam_scalars;    /* [n_beams * MBS_BEAM_SCALARS]             */
    const int32_t *beam_nv;        /* [n_beams] aperture vertex count          */
    const int32_t *beam_external;  /* [n_beams] 1 = external (shadow) beam     */
    const int64_t *vert_offsets;   /* [n_beams+1] vertex ranges per beam       */
    const double *vx;              /* [n_vert] aperture-plane x (edgeData.x)   */
    const double *vy;              /* [n_vert] aperture-plane y (edgeData.y)   */
    const double *slope_yx;        /* [n_vert] dy/dx of edge e (if valid_x)    */
    const double *slope_xy;        /* [n_vert] dx/dy of edge e (if valid_y)    */
    const uint8_t *edge_flags;     /* [n_vert] MBS_EDGE_VALID_X | _Y           */
} mbs_beam_batch;

/* Scattering grid: n_az phi bins (phi_p = p*az_step) by n_zen+1 theta rows,
 * the ScatteringRange of src/handler/Handler.h:15-189.  Cell (p, t) is stored
 * at index p*(n_zen+1)+t in every grid-shaped array. */
typedef struct mbs_grid {
    int32_t n_az;
    int32_t n_zen;
    const double *theta; /* [n_zen+1] radians                                */
    double az_step;      /* radians                                           */
    const double *vf;    /* [n_az*(n_zen+1)*3] polarisation basis vf per cell */
} mbs_grid;

/* Physical constants of the handler (src/handler/Handler.cpp:102-126). */
typedef struct mbs_consts {
    double wave_index;        /* k = 2*pi/lambda                              */
    double eps1;              /* edge-coefficient threshold (1e9*DBL_EPSILON) */
    double compl_wave[2];     /* -i*lambda/(4*pi^2)                           */
    double inv_compl_wave[2]; /* i/lambda                                     */
    int32_t legacy_sign;      /* m_legacySign                                 */
    int32_t reserved;
} mbs_consts;

/* Flags of mbs_diffract. */
#define MBS_PREC_FP64 0u      /* FP64 throughout                                */
#define MBS_PREC_MIXED 1u     /* FP64 phase arguments, their reduction, edge
                                 coefficients and polygon moments; FP32
                                 phasors, edge sums, rotation and Jones sums.
                                 Edges with |C| <= max(eps1, 3e-5) take the
                                 exact half-angle form in FP64, as the
                                 reference FP32 build does                    */
#define MBS_NOSHADOW 4u       /* also accumulate the no-shadow Mueller grid     */
#define MBS_PER_ORIENT 8u     /* also keep one Mueller grid per orientation     */
#define MBS_NO_MFMA 16u       /* mixed precision: rotation and coherent sum on
                                 the vector ALU instead of the MFMA contraction
                                 (same arithmetic, for comparison)              */
#define MBS_INCOHERENT 32u    /* FP64 sum of every beam's own Mueller matrix,
                                 sum_o w_o sum_b Mueller(d_b), instead of the
                                 Mueller matrix of the coherent sum; its M11 is
                                 what the reference's theta probe sums
                                 (theta_m11_kernel, GpuDiffraction.cu:1909-1957).
                                 Not with MBS_PREC_MIXED, MBS_NOSHADOW or
                                 MBS_PER_ORIENT (MBS_ERR_INVALID)               */
#define MBS_SPECTRUM 64u      /* mixed precision (MFMA contraction): the beams
                                 mbs_spectrum_source selected take F from
                                 their FFT spectra, the others the direct
                                 integral.  Not with MBS_PER_ORIENT or
                                 MBS_INCOHERENT; needs a current
                                 mbs_spectrum_source (MBS_ERR_INVALID)         */

typedef struct mbs_ctx mbs_ctx;

/* Library version and build information. */
int mbs_abi_version(void);
const char *mbs_build_info(void);

/* ---- GPU runtime queries: replace src/cuda/GpuSupport.h:30-34 ---------- */
typedef struct mbs_device_info {
    int32_t device_id;
    int32_t visible_devices;
    int64_t total_mem;
    int64_t free_mem;
    int32_t compute_units;
    int32_t clock_khz;
    int32_t runtime_version;
    int32_t driver_version;
    char name[128];
    char arch[64];
} mbs_device_info;

/* CheckGpuRuntime (GpuSupport.h:30): fills *info for the current device. */
int mbs_check_runtime(mbs_device_info *info);
/* QueryActiveGpuMemory (GpuSupport.h:31). */
int mbs_query_memory(int64_t *free_bytes, int64_t *total_bytes);
/* VisibleGpuDeviceCount (GpuSupport.h:33); 0 when no device is usable. */
int mbs_visible_device_count(void);

/* ---- contexts ------------------------------------------------------------ */
/* Creates a context on HIP device `device` with its own non-blocking stream. */
int mbs_ctx_create(int device, mbs_ctx **out);
int mbs_ctx_destroy(mbs_ctx *ctx);
const char *mbs_last_error(const mbs_ctx *ctx);

/* Uploads the direction grid (tables are derived on the device).  The grid
 * stays resident until replaced. */
int mbs_set_grid(mbs_ctx *ctx, const mbs_grid *grid);

/* Uploads one batch of orientations' prepared beams and repacks them on the
 * device.  Rejects beams with nv outside [MBS_MIN_VERTICES, MBS_MAX_VERTICES]
 * with MBS_ERR_BAD_BEAM, as the reference GPU path does. */
int mbs_upload_beams(mbs_ctx *ctx, const mbs_beam_batch *batch,
                     const mbs_consts *consts);

/* Page-locked staging of an upload (no reference counterpart; the reference
 * copies from pageable packing buffers, src/cuda/GpuDiffraction.cu:4572-4606).
 * mbs_stage_beams hands out host arrays in page-locked memory owned by the
 * context, sized for a batch of n_orient orientations, n_beams beams and
 * n_vert vertices, in the layout of mbs_beam_batch; the caller fills them
 * and calls mbs_upload_staged, which checks and uploads them like
 * mbs_upload_beams but returns as soon as the copies are queued: they run as
 * DMA at the link's rate while the host goes on, and the next mbs_diffract
 * on the context follows them on its stream.  Two staging areas are used in
 * turn, so the caller can fill the next batch while the previous one is
 * still in flight; mbs_stage_beams waits for the copies out of the area it
 * hands back.  The arrays stay valid until the second mbs_stage_beams after
 * this one. */
typedef struct mbs_beam_staging {
    int32_t n_orient;
    int32_t n_beams;
    int64_t n_vert;
    int64_t *orient_offsets; /* [n_orient+1] */
    double *weights;         /* [n_orient]   */
    double *beam_scalars;    /* [n_beams * MBS_BEAM_SCALARS] */
    int32_t *beam_nv;        /* [n_beams]    */
    int32_t *beam_external;  /* [n_beams]    */
    int64_t *vert_offsets;   /* [n_beams+1]  */
    double *vx, *vy;         /* [n_vert]     */
    double *slope_yx, *slope_xy;
    uint8_t *edge_flags;     /* [n_vert]     */
} mbs_beam_staging;
int mbs_stage_beams(mbs_ctx *ctx, int32_t n_orient, int32_t n_beams, int64_t n_vert,
                    mbs_beam_staging *out);
int mbs_upload_staged(mbs_ctx *ctx, const mbs_consts *consts);

/* Per-beam data from which the device re-derives a size-rescaled beam set
 * (the multi-size path: the reference packs every scale on the host,
 * src/cuda/GpuDiffraction.cu:273-358, applied when |scale - 1| > 1e-15,
 * :4324; its fused kernel rescales phases on the device, :1204-1238). */
typedef struct mbs_beam_scaling {
    const double *raw_jones;     /* [n_beams*8] the beam's Jones matrix J      */
    const double *path;          /* [n_beams] optical path (external beams) or
                                    projected length (internal) at scale 1    */
    const double *sign;          /* [n_beams] -1 for external beams and internal
                                    ones with an odd number of interactions  */
    const int64_t *abs_offsets;  /* [n_beams+1] ranges in abs_paths            */
    const double *abs_paths;     /* absorption path lengths at scale 1         */
    double c_abs;                /* absorption coefficient                     */
    double wave_index;           /* k of the phase exp(i k path scale)         */
} mbs_beam_scaling;

/* Uploads the scaling data of the batch of the last mbs_upload_beams
 * (n_beams entries); it stays until the next mbs_upload_beams.  No
 * reference counterpart: it replaces the host re-pack of every scale. */
int mbs_upload_beam_scaling(mbs_ctx *ctx, const mbs_beam_scaling *sc);

/* Re-derives the device beam records of the uploaded batch for size scale s,
 * taking the uploaded batch as scale 1: aperture vertices and beam centres
 * times s, areas times s^2, and J_phased = sign * a(s) * J * exp(i k path s)
 * with a(s) the mean over the beam's absorption paths p of exp(c_abs p s)
 * (1 for p <= DBL_EPSILON, and 1 without paths).  For |s - 1| <= 1e-15 the
 * batch's own records are restored.  Needs mbs_upload_beam_scaling. */
int mbs_rescale_beams(mbs_ctx *ctx, double scale);

/* Evaluates the uploaded batch on the current grid and leaves the result on
 * the device: the weighted Mueller sum over the batch's orientations (and
 * the no-shadow sum / per-orientation grids when flagged).  Asynchronous on
 * the context's stream. */
int mbs_diffract(mbs_ctx *ctx, uint32_t flags);

/* Copies the device result of the last mbs_diffract to the host and ADDS it
 * to m_full / m_noshadow ([n_az*(n_zen+1)*16], row-major 4x4 per cell), the
 * accumulate semantics of HandleOrientationsToLocalGpu.  m_noshadow may be
 * NULL.  Synchronises the context's stream. */
int mbs_accumulate_mueller(mbs_ctx *ctx, double *m_full, double *m_noshadow);

/* Copies the per-orientation grids of the last mbs_diffract (flag
 * MBS_PER_ORIENT) to out ([n_orient*n_az*(n_zen+1)*16]), each multiplied by
 * its orientation weight. */
int mbs_fetch_per_orient(mbs_ctx *ctx, double *out);

/* One-call convenience: upload, diffract, accumulate. */
int mbs_diffract_batch(mbs_ctx *ctx, const mbs_grid *grid,
                       const mbs_beam_batch *batch, const mbs_consts *consts,
                       uint32_t flags, double *m_full, double *m_noshadow);

/* Azimuth (laboratory-alpha) average of the last mbs_diffract result, on the
 * device: the rows HandlerPOTotal::WriteMatricesToFile writes to its .dat
 * (src/handler/HandlerPOTotal.cpp:23-122), with the pole forms of
 * src/handler/PoleMueller.h:25-55.  rows: [n_zen+1][18] = theta in degrees,
 * 2*pi*dcos, M11..M44 (row-major 4x4).  noshadow != 0 averages the no-shadow
 * grid (requires MBS_NOSHADOW on the last mbs_diffract).  Overwrites rows. */
int mbs_azimuth_average(mbs_ctx *ctx, int noshadow, double *rows);
/* The same for every orientation of the last mbs_diffract with
 * MBS_PER_ORIENT: rows [n_orient][n_zen+1][18] (each grid carries its
 * orientation's weight, as mbs_fetch_per_orient returns it). */
int mbs_azimuth_average_per_orient(mbs_ctx *ctx, double *rows);

/* Band-limited (Fourier) upsampling of the last mbs_diffract result on the
 * device: the interpolation step of the reference's angular FFT path
 * (fft_upsample_phi_arr2d / fft_upsample_theta_arr2d and the pad kernel,
 * src/cuda/GpuDiffraction.cu:3195-3420, :794-814).  The result grid is taken
 * as samples of a function periodic in azimuth (n_az samples per period)
 * and, when n_zen_out > n_zen, even about both ends of the zenith range
 * (2 n_zen samples per period; the zenith rows must be uniformly spaced);
 * each line is resampled to n_az_out azimuths and n_zen_out + 1 rows over
 * the same range: forward DFT, the spectrum zero-padded (an even count's
 * Nyquist term split evenly between +n/2 and -n/2, which for real data is
 * the real part of the one-sided placement), inverse DFT, scale 1/n.
 * Zenith first, then azimuth; FP64 throughout.  Sizes may only grow; every
 * line that is resampled is limited to 4096 samples (n_az_out, 2 n_zen_out),
 * else MBS_ERR_CAPACITY.
 *   noshadow    upsample the no-shadow result (needs MBS_NOSHADOW)
 *   grid_out    NULL, or [n_az_out*(n_zen_out+1)*16], written (accumulate =
 *               0) or added to (accumulate != 0), cell (q, t) at
 *               q*(n_zen_out+1)+t
 *   rows_out    NULL, or [n_zen_out+1][18]: the azimuth average of the
 *               upsampled grid as mbs_azimuth_average computes it; theta_out
 *               [n_zen_out+1] (radians) gives its rows' zenith angles */
int mbs_upsample_result(mbs_ctx *ctx, int noshadow, int32_t n_az_out, int32_t n_zen_out,
                        const double *theta_out, int accumulate, double *grid_out,
                        double *rows_out);

/* Orientation sets of the orientation average, generated on the device
 * (one row of MBS_ORIENT_STRIDE doubles per orientation: beta, gamma, weight,
 * qx, qy, qz, qw, R00..R22 of the particle rotation):
 *   MBS_ORIENT_SOBOL       Owen-scrambled 2-D Sobol (seed), beta = acos(1 -
 *                          (1 - cos beta_sym) u), gamma = gamma_sym v, weight
 *                          1/n: TracerPOTotal::TraceFromSobolSeed
 *                          (src/scattering/TracerPOTotal.cpp:5465-5492) with
 *                          Sobol2D (src/math/Sobol.cpp:1-106)
 *   MBS_ORIENT_HAMMERSLEY  u = (i+1/2)/n, v = radical inverse base 2 of i:
 *                          TraceFromSO3Quaternion (:5347-5426), unpaired
 *   MBS_ORIENT_SHOEMAKE    full-SO(3) quaternions from (u, RI2(i), RI3(i)):
 *                          TraceFromSO3FullQuaternion (:5428-5463); beta and
 *                          gamma are NaN, beta_sym/gamma_sym are ignored
 * R follows Particle::SetRotateMatrix(beta, gamma, 0) /
 * SetRotateMatrixFromQuaternion (src/particle/Particle.cpp:1593-1642).
 * out: [n][MBS_ORIENT_STRIDE]. */
#define MBS_ORIENT_SOBOL 0
#define MBS_ORIENT_HAMMERSLEY 1
#define MBS_ORIENT_SHOEMAKE 2
#define MBS_ORIENT_STRIDE 16
#define MBS_MAX_ORIENT_SET (1 << 26)
int mbs_orientation_set(mbs_ctx *ctx, int kind, int32_t n, uint32_t seed, double beta_sym,
                        double gamma_sym, double *out);

/* Aperture FFT (near-forward path), after mbs_upload_beams: for each listed
 * beam of the uploaded batch, the aperture polygon (vx, vy) is sampled on an
 * n x n grid of pixel centres spanning pad times its bounding box, Fourier
 * transformed in FP32 on the device, and the transform
 *   I(A, B) = integral over the aperture of exp(i k (A x + B y)) dx dy
 * (k = wave_index of the upload) is interpolated at each (A, B) point with a
 * phase-corrected Lanczos kernel of the given radius.  The method of the
 * reference's aperture probe (tools/fft_aperture_probe.cu:130-169, :549-551,
 * :50-102), for convex polygons.  (Multiply by -i/lambda for the aperture
 * factor of mbs_diffract's integral.)
 *   beams   [n_beams] indices into the uploaded batch, 1 <= n_beams <= 64
 *   n       grid side, a power of two in MBS_FFT_MIN_N..MBS_FFT_MAX_N
 *   ab      [n_points][2] (A, B);  out [n_beams][n_points][2] (re, im)
 * n_points = 0 computes the spectra only (kept on the device until the next
 * call).  Device workspace: n_beams * n * n * 8 bytes. */
#define MBS_FFT_MAX_BATCH 64
#define MBS_FFT_MIN_N 64
#define MBS_FFT_MAX_N 2048
int mbs_aperture_fft(mbs_ctx *ctx, const int64_t *beams, int32_t n_beams, int32_t n,
                     double pad, int32_t n_points, const double *ab, int32_t radius,
                     double *out);
/* Spectrum `index` of the last mbs_aperture_fft batch: the DFT of its raster,
 * [n rows (y)][n columns (x)] of (re, im) FP32, unnormalised. */
int mbs_aperture_fft_spectrum(mbs_ctx *ctx, int32_t index, float *out);
/* With timing enabled: device times of the last mbs_aperture_fft's raster+row
 * pass, column pass and sampling. */
int mbs_aperture_fft_timing(mbs_ctx *ctx, float *rows_ms, float *cols_ms, float *sample_ms);

/* ---- near-forward F source ----------------------------------------------
 * mbs_spectrum_source selects, for the uploaded batch on the current grid,
 * the beams whose scattering directions all lie close to their own
 * (max over the grid of max(|A|, |B|) at most a_max), whose polygon is
 * convex with at most 16 vertices, and whose spectrum fits an FFT of at most
 * max_n points per side; for each it rasterises the aperture (tent-weighted:
 * every pixel the polygon integral of a tent of half-width h around it,
 * exact), transforms it, and samples F at every cell of the grid by
 * non-uniform FFT interpolation (exponential-of-semicircle kernel, 6 taps
 * per axis, oversampling 2).  mbs_diffract with MBS_SPECTRUM then uses those
 * F values for the selected beams.  The pixel size is h = min(pi / (k a_b
 * oversample), extent / 32, width / 16), a_b the beam's own max(|A|, |B|)
 * (so |f h| <= 1/(2 oversample) on the grid), extent the larger side of its
 * bounding box and width its least width, with halved oversampling (down
 * to 2) when the FFT would exceed max_n.  The grid side is the power of two
 * >= 2 (extent + 4 h) / h, at least min_n.  This restates the reference's
 * aperture probe
 * (tools/fft_aperture_probe.cu: 0/1 raster :130-169, 2-D FFT :549-551,
 * Lanczos sampling :50-102) with an anti-aliased raster and a non-uniform
 * FFT interpolation, composed into the coherent sum.  A new upload, rescale
 * or grid drops the selection.  p = NULL drops it too.
 * Device memory: sum over selected beams of n^2 * 8 bytes (spectra) plus
 * selected beams * cells * 8 bytes (F). */
typedef struct mbs_spectrum_params {
    double a_max;       /* selection bound on max(|A|, |B|) over the grid     */
    double oversample;  /* pixels per half period of the highest frequency    */
    int32_t min_n;      /* smallest FFT side (>= MBS_FFT_MIN_N)               */
    int32_t max_n;      /* largest FFT side (<= MBS_FFT_MAX_N)                */
} mbs_spectrum_params;
typedef struct mbs_spectrum_info {
    int32_t n_selected;       /* beams on the spectrum path                    */
    int32_t n_by_size[6];     /* of which with FFT side 64, 128, ..., 2048     */
    int32_t reserved;
    int64_t spectrum_bytes;   /* device memory of the spectra                  */
    double selected_area;     /* sum of the selected beams' aperture areas     */
    double total_area;        /* sum over all beams                            */
    float scan_ms, fft_ms, sample_ms;  /* device times (timing enabled)        */
    float reserved2;
} mbs_spectrum_info;
int mbs_spectrum_source(mbs_ctx *ctx, const mbs_spectrum_params *p, mbs_spectrum_info *info);
/* The current selection: beams[n_selected] (indices into the uploaded batch,
 * in slot order) and, if f is not NULL, f[n_selected][cells][2] the F values
 * (re, im) the spectrum variant uses, cells in the device's theta-major order
 * (cell t * n_az + p), in the kernels' convention
 * s k^2 I(A, B) e^{i psi} (that is, F / complWave of mbs_diffract's aperture
 * factor). */
int mbs_spectrum_values(mbs_ctx *ctx, int32_t *beams, float *f);

/* Learned control variate: a multilayer perceptron trained on the device
 * (no counterpart in the reference).  Layers: n_in -> depth x width (GELU,
 * exact erf form) -> n_out (linear).  GEMMs run on BF16 MFMA with FP32
 * accumulation; weights are FP32 masters, rounded to BF16 as the GEMMs load
 * them.  Parameters are packed per layer as W [in][out] row-major then b
 * [out].  Created with He-uniform weights from `seed`, zero biases.  The
 * model lives on ctx's device and stream. */
typedef struct mbs_mlp mbs_mlp;
int mbs_mlp_create(mbs_ctx *ctx, int32_t n_in, int32_t width, int32_t depth, int32_t n_out,
                   uint64_t seed, mbs_mlp **out);
int mbs_mlp_destroy(mbs_mlp *m);
int64_t mbs_mlp_param_count(const mbs_mlp *m);
int mbs_mlp_set_params(mbs_mlp *m, const float *params);
int mbs_mlp_get_params(mbs_mlp *m, float *params);
/* y [rows][n_out] = MLP(x [rows][n_in]). */
int mbs_mlp_forward(mbs_mlp *m, const float *x, int32_t rows, float *y);
/* Mean-squared-error loss over rows x n_out and its gradient w.r.t. every
 * parameter (packed like the parameters). */
int mbs_mlp_gradient(mbs_mlp *m, const float *x, const float *t, int32_t rows, double *loss,
                     float *grad);
/* Adam (beta 0.9, 0.999, eps 1e-8) over `epochs` passes of shuffled
 * minibatches of `batch` rows (a trailing partial batch is dropped);
 * final_loss = mean loss of the last epoch's batches. */
int mbs_mlp_train(mbs_mlp *m, const float *x, const float *t, int32_t n, int32_t batch,
                  int32_t epochs, float lr, uint64_t seed, double *final_loss);

/* ---- the entry point that allocates nothing ------------------------------
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  mbs_diffract_noalloc evaluates the uploaded batch on
 * the current grid like mbs_diffract(flags) and WRITES (does not add) the
 * weighted Mueller sum to caller-owned device memory d_m_full
 * ([n_az*(n_zen+1)*16] FP64, row-major 4x4 per cell, cell (p, t) at
 * p*(n_zen+1)+t), and with MBS_NOSHADOW the no-shadow sum to d_m_noshadow
 * (NULL without it).  It calls no hipMalloc, hipFree or allocating library
 * routine: its work list and partial sums live in one fixed device array of
 * MBS_NOALLOC_SCRATCH_BYTES per device, which the library holds until the
 * process ends, and the launch is sized to fit it (on large grids fewer,
 * larger work items than mbs_diffract uses).  The allocations it relies on
 * are made beforehand by mbs_ctx_create, mbs_set_grid and mbs_upload_beams.
 * Flags: MBS_PREC_MIXED, MBS_NO_MFMA, MBS_NOSHADOW.  MBS_ERR_CAPACITY when
 * one work item's Mueller grid (n_az*(n_zen+1)*128 bytes, twice that with
 * MBS_NOSHADOW) exceeds the array.  Asynchronous on the context's stream.
 * Every call on one device shares that device's array and leaves its
 * contents behind: calls that overlap on one device (two contexts or host
 * threads) race, and the caller must serialise them. */
#define MBS_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int mbs_diffract_noalloc(mbs_ctx *ctx, uint32_t flags, double *d_m_full, double *d_m_noshadow);

/* Device pointer of the result grid of the last mbs_diffract (FP64,
 * [n_az*(n_zen+1)*16]); valid until the next mbs_upload_beams/mbs_set_grid. */
int mbs_device_result(mbs_ctx *ctx, const double **m_full,
                      const double **m_noshadow);

/* Stage timing, the counterpart of the reference's MBS_GPU_TIMING breakdown
 * (src/cuda/GpuDiffraction.cu:550-554, :4767-4775).  When enabled, HIP
 * events on the context's stream bracket the upload+pack of
 * mbs_upload_beams and the main kernel and the reduction of mbs_diffract;
 * mbs_get_timing waits for them and reports milliseconds of the last call of
 * each. */
typedef struct mbs_timing {
    float upload_ms;   /* H2D copies + device packing of the last upload   */
    float kernel_ms;   /* main diffraction kernel of the last mbs_diffract */
    float reduce_ms;   /* orientation/segment reduction of the same call   */
    int32_t n_items;   /* work items of the last mbs_diffract              */
    int32_t n_blocks;  /* workgroups launched by it                        */
} mbs_timing;
int mbs_set_timing(mbs_ctx *ctx, int enable);
int mbs_get_timing(mbs_ctx *ctx, mbs_timing *out);

/* Tuning: approximate number of workgroups one mbs_diffract launch aims
 * for (default 65536).  Work items are sized from it. */
int mbs_set_target_blocks(mbs_ctx *ctx, int blocks);

/* Work decomposition of mbs_diffract: 0 = automatic (beam segments of equal
 * size unless their partial Jones grids exceed 8 GiB), 1 = runs of whole
 * orientations, 2 = beam segments, 3 = beam segments planned on the device
 * and run by a persistent launch that takes (segment, direction tile) units
 * from a device work queue (one atomic counter per XCD over a contiguous
 * range of units, ranges of other XCDs taken when its own is empty); with
 * graph replay the plan is part of the graph, so a replay follows a new
 * upload with the same orientation count without any host work.
 * MBS_PER_ORIENT and MBS_INCOHERENT run mode 3 as mode 2.  Results agree
 * to rounding; modes 2 and 3 bit for bit. */
#define MBS_WORK_AUTO 0
#define MBS_WORK_ORIENTATIONS 1
#define MBS_WORK_SEGMENTS 2
#define MBS_WORK_QUEUE 3
int mbs_set_work_mode(mbs_ctx *ctx, int mode);

/* Autotuner of the launch shape (no counterpart in the reference, whose CUDA
 * path sizes its launch from fixed heuristics, GpuDiffraction.cu:2948-3029):
 * times mbs_diffract(flags) on the uploaded batch for every candidate
 * (target workgroups 4096..131072 x whole-orientation / beam-segment /
 * device-queue work items x, for mixed precision, MFMA / vector-ALU coherent
 * sum, and for the MFMA kernel with beam segments and no MBS_NOSHADOW its
 * register allocation held to 4, 5 or 6 waves per SIMD; after a warm-up
 * call, the least of three timed calls), keeps the fastest for later
 * mbs_diffract calls of the same precision, and reports what it timed.
 * mbs_set_target_blocks / mbs_set_work_mode discard it.
 * flags: MBS_PREC_MIXED and/or MBS_NOSHADOW only.  Leaves a result like
 * mbs_diffract(flags). */
typedef struct mbs_tune_entry {
    int32_t target_blocks;
    int32_t work_mode;  /* 1 whole orientations, 2 beam segments, 3 device queue */
    int32_t mfma;       /* 1 MFMA contraction, 0 vector ALU (mixed only) */
    float ms;           /* main kernel + reduction, device time (least of 3) */
    int32_t waves;      /* MFMA kernel: waves per SIMD its registers are held to
                           (4, 5 or 6; 4 and 6 with beam segments only); 0 for
                           the other kernels */
} mbs_tune_entry;
int mbs_autotune(mbs_ctx *ctx, uint32_t flags, mbs_tune_entry *entries, int32_t max_entries,
                 int32_t *n_entries);
/* Applies one launch shape (an entry mbs_autotune reported, e.g. from an
 * earlier process) for later mbs_diffract calls of the precision in flags,
 * as if the autotuner had chosen it; entry->ms is ignored.  MBS_ERR_INVALID
 * for a shape the tuner would not try. */
int mbs_apply_tune_entry(mbs_ctx *ctx, uint32_t flags, const mbs_tune_entry *entry);

/* Graph replay: with enable != 0, mbs_diffract captures its kernel sequence
 * (diffraction + reduction, no copies or memsets; with MBS_WORK_QUEUE also
 * the device's segment planning) into a HIP graph and replays it while
 * launch shape, buffers and sizes are unchanged; re-captured otherwise.  A
 * host-planned work list is copied to the device before each replay; the
 * device queue plans each replay on the device, so a new upload of the same
 * sizes replays without host work. */
int mbs_set_graph(mbs_ctx *ctx, int enable);

/* Graph use of the context so far: captures made and graph launches (a
 * launch that needed no capture is a replay). */
int mbs_graph_stats(mbs_ctx *ctx, int64_t *captures, int64_t *launches);

/* The HIP stream the context launches on (a hipStream_t, as void*), for
 * timing with events recorded on the same stream. */
void *mbs_ctx_stream(mbs_ctx *ctx);

/* Statistics of the uploaded batch: total beams, beam.direction evaluation
 * units (sum of beams x directions), histogram of vertex counts
 * (hist[nv], 32 entries), histogram of valid edges per beam. */
int mbs_batch_stats(mbs_ctx *ctx, int64_t *n_beams, double *units,
                    int64_t *nv_hist32, double *mean_edges);

#ifdef __cplusplus
}
#endif

#endif /* MBS_HIP_H */
