// po_kernels.hip -- gfx950 kernels of the PO diffraction engine.
//
// Work mapping.  One lane owns one scattering direction (grid cell, in the
// device's theta-major numbering, see abi_cell); a 256-lane workgroup owns 256
// consecutive cells of one "work item".  A work
// item is either a run of whole orientations (Mueller mode: the lane sums the
// Jones matrices of every beam of an orientation in registers, converts to
// Mueller, and accumulates weight*Mueller over the run) or one segment of one
// orientation's beams (Jones mode, used when there are too few orientations
// to fill the chip: the lane writes the partial Jones sum and a second kernel
// adds the segments and converts).  Every beam of an item is the same for all
// lanes, so beam records are read with wave-uniform loads (scalar cache), and
// all arithmetic per (beam, direction) stays in VGPRs.  Results are reduced
// in a fixed order by separate kernels, so output is bit-reproducible.
//
// Per (beam, direction) the lane evaluates, following the reference CPU loop
// src/handler/HandlerPO.cpp:2632-2776:
//   A = (b - d).hor, B = (b - d).ver                       (k-k0 projected)
//   phi_v = A*k*x_v + B*k*y_v  (+ psi = -k d.c for internal beams, folded into
//   the vertex phasors as the reference CUDA path does, GpuDiffraction.cu:
//   1110-1122), F = complWave * sum_e (e^{i phi_e+1} - e^{i phi_e})/C_e / B
//   (|B|>|A|) or ... / (-A), with the small-phase moment form when
//   max|phi_v| <= 1e-3 and the half-angle form when |C_e| <= eps1;
//   r_ij = u_i(d).T_j(beam) (the Jones rotation of HandlerPO_fast.h:580-683
//   rewritten exactly with per-direction vectors u_i), d = F R J_phased.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "po_math.h"
#include "po_layout.h"

// Occupancy of the vector-ALU kernels: at least 4 waves per SIMD (128
// VGPRs).  The FP64 kernel, whose batched vertex phasors would otherwise take
// ~140 VGPRs and 3 waves, spills only a few scalars per beam to keep it.
// Builds may pass -DMBS_MIN_WAVES=N for tuning.
#ifndef MBS_MIN_WAVES
#define MBS_MIN_WAVES 4
#endif
#define MBS_OCCUPANCY __attribute__((amdgpu_waves_per_eu(MBS_MIN_WAVES, 8)))
// The MFMA kernel is held to WAVES waves per SIMD (default 5: 96 VGPRs)
// when it keeps one accumulator set -- its 20.5 KB of LDS per workgroup lets
// seven workgroups share a CU; the no-shadow variant (two sets) gets 2.  The
// autotuner also tries 4 and 6 for the default beam-segment kernel.
#ifndef MBS_MFMA_WAVES
#define MBS_MFMA_WAVES 5
#endif
#define MBS_MFMA_OCCUPANCY __attribute__((amdgpu_waves_per_eu(NS ? 2 : WAVES, 8)))

namespace mbs {

// ---------------------------------------------------------------------------
// Direction cells are numbered theta-major on the device (cell = t n_az + p:
// the azimuths of one theta row are consecutive lanes), while the C ABI's
// grids are azimuth-major (p (n_zen + 1) + t).  Every azimuth cell of a pole
// row has the same direction, and a beam leaving along it meets singular edge
// coefficients there in every such cell; theta-major order keeps those cells
// in two waves instead of one wave per azimuth.  abi_cell maps a device cell
// to its place in the grids the kernels hand back.
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t abi_cell(int cell, int n_az, int cells)
{
    const int t = cell / n_az;
    const int p = cell - t * n_az;
    return (size_t)p * (cells / n_az) + t;
}

// ---------------------------------------------------------------------------
// Grid tables: per cell dir(3), u0(6), u1(6) as SoA [DT_COUNT][cells].
// u0 = (vt x d, vt - d (d.vt)), u1 = (vf x d, vf - d (d.vf)) with
// vt = (vf x d)/|vf x d| (kept unnormalised when |vf x d|^2 <= 1e-30, as
// rotate_jones_inline does).  Then cpT.vt = NT.(vt x d) + nxDT.(vt - d(d.vt))
// for cpT = d x NT + nxDT - d (d.nxDT), and likewise for cpP, vf.
// ---------------------------------------------------------------------------
__global__ void grid_tables_kernel(const double *__restrict__ theta,
                                   const double *__restrict__ vf, int n_az,
                                   int n_zen1, double az_step,
                                   double *__restrict__ tab, float *__restrict__ tab32)
{
    const int cells = n_az * n_zen1;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells)
        return;
    // device cell order is theta-major (see abi_cell); vf is azimuth-major
    const int t = cell / n_az;
    const int p = cell - t * n_az;
    double st, ct, sp, cp;
    sincos(theta[t], &st, &ct);
    sincos(p * az_step, &sp, &cp);
    const double d[3] = {st * cp, st * sp, -ct};
    const double *vfc = vf + 3 * ((size_t)p * n_zen1 + t);
    const double f[3] = {vfc[0], vfc[1], vfc[2]};
    double vt[3] = {f[1] * d[2] - f[2] * d[1], f[2] * d[0] - f[0] * d[2],
                    f[0] * d[1] - f[1] * d[0]};
    const double l2 = vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2];
    if (l2 > 1e-30) {
        const double inv = 1.0 / sqrt(l2);
        vt[0] *= inv;
        vt[1] *= inv;
        vt[2] *= inv;
    }
    const double dvt = d[0] * vt[0] + d[1] * vt[1] + d[2] * vt[2];
    const double dvf = d[0] * f[0] + d[1] * f[1] + d[2] * f[2];
    double u[DT_COUNT];
    u[DT_DIR + 0] = d[0];
    u[DT_DIR + 1] = d[1];
    u[DT_DIR + 2] = d[2];
    // u0: (vt x d, vt - d (d.vt))
    u[DT_U0 + 0] = vt[1] * d[2] - vt[2] * d[1];
    u[DT_U0 + 1] = vt[2] * d[0] - vt[0] * d[2];
    u[DT_U0 + 2] = vt[0] * d[1] - vt[1] * d[0];
    u[DT_U0 + 3] = vt[0] - d[0] * dvt;
    u[DT_U0 + 4] = vt[1] - d[1] * dvt;
    u[DT_U0 + 5] = vt[2] - d[2] * dvt;
    u[DT_U1 + 0] = f[1] * d[2] - f[2] * d[1];
    u[DT_U1 + 1] = f[2] * d[0] - f[0] * d[2];
    u[DT_U1 + 2] = f[0] * d[1] - f[1] * d[0];
    u[DT_U1 + 3] = f[0] - d[0] * dvf;
    u[DT_U1 + 4] = f[1] - d[1] * dvf;
    u[DT_U1 + 5] = f[2] - d[2] * dvf;
    for (int q = 0; q < DT_COUNT; ++q)
        tab[(size_t)q * cells + cell] = u[q];
    // FP32 u0, u1 for the mixed kernels (half the bytes per work unit), one
    // 48-byte record per cell
    for (int q = 0; q < DT32_COUNT; ++q)
        tab32[(size_t)cell * DT32_COUNT + q] = (float)u[DT_U0 + q];
}

// ---------------------------------------------------------------------------
// Beam packing: one thread per beam, from the C-ABI input arrays to the
// kernel record (po_layout.h).  Small-phase moments follow
// small_phase_polygon_factor (src/handler/HandlerPO_fast.h:154-216).
// ---------------------------------------------------------------------------
__global__ void pack_beams_kernel(int n_beams, const double *__restrict__ bs,
                                  const int32_t *__restrict__ nv_in,
                                  const int32_t *__restrict__ ext_in,
                                  const int64_t *__restrict__ voff,
                                  const double *__restrict__ vx,
                                  const double *__restrict__ vy,
                                  const double *__restrict__ syx,
                                  const double *__restrict__ sxy,
                                  const uint8_t *__restrict__ ef, double k,
                                  double al_re, double al_im, double cw_re,
                                  double cw_im, double sc,
                                  const double *__restrict__ raw_j,
                                  const double *__restrict__ path,
                                  const double *__restrict__ sign,
                                  const int64_t *__restrict__ abs_off,
                                  const double *__restrict__ abs_path, double c_abs,
                                  double k_j,
                                  double *__restrict__ hdr,
                                  float *__restrict__ hdr32,
                                  double *__restrict__ vtx,
                                  double *__restrict__ vdd,
                                  int32_t *__restrict__ vflag,
                                  float *__restrict__ wb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_beams)
        return;
    const double *s = bs + (size_t)b * 36;
    const int nv = nv_in[b];
    const int64_t v0 = voff[b];
    // size rescaling (raw_j set): vertices and centre times sc, area times
    // sc^2, J_phased rebuilt below; sc = 1 leaves every value as it is
    auto X = [&](int i) { return vx[v0 + i] * sc; };
    auto Y = [&](int i) { return vy[v0 + i] * sc; };
    double *h = hdr + (size_t)b * HDR_STRIDE;
    float *h32 = hdr32 + (size_t)b * HDR32_STRIDE;
    const double *hor = s + 0, *ver = s + 3, *bd = s + 6;
    const double cen[3] = {s[9] * sc, s[10] * sc, s[11] * sc};
    h[H_HOR + 0] = hor[0];
    h[H_HOR + 1] = hor[1];
    h[H_HOR + 2] = hor[2];
    h[H_A0] = bd[0] * hor[0] + bd[1] * hor[1] + bd[2] * hor[2];
    h[H_VER + 0] = ver[0];
    h[H_VER + 1] = ver[1];
    h[H_VER + 2] = ver[2];
    h[H_B0] = bd[0] * ver[0] + bd[1] * ver[1] + bd[2] * ver[2];
    const bool ext = ext_in[b] != 0;
    // psi = -k d.c for internal beams; external beams carry no direction phase
    h[H_KC + 0] = ext ? 0.0 : -k * cen[0];
    h[H_KC + 1] = ext ? 0.0 : -k * cen[1];
    h[H_KC + 2] = ext ? 0.0 : -k * cen[2];
    // complWave is folded into J_phased below; the small-phase area limit is
    // therefore stored divided by complWave.
    const double area = s[12] * (sc * sc);
    const double cw2 = cw_re * cw_re + cw_im * cw_im;
    const double lr = al_re * area, li = al_im * area;
    h[H_AL + 0] = (lr * cw_re + li * cw_im) / cw2;
    h[H_AL + 1] = (li * cw_re - lr * cw_im) / cw2;
    // polygon moments about the aperture origin
    double tw = 0, fx6 = 0, fy6 = 0, sx12 = 0, sy12 = 0, sxy24 = 0;
    double minX = 1e300, maxX = -1e300, minY = 1e300, maxY = -1e300;
    for (int i = 0; i < nv; ++i) {
        const int j = (i + 1 < nv) ? i + 1 : 0;
        const double xi = X(i), yi = Y(i);
        const double xj = X(j), yj = Y(j);
        const double cr = xi * yj - xj * yi;
        tw += cr;
        fx6 += (xi + xj) * cr;
        fy6 += (yi + yj) * cr;
        sx12 += (xi * xi + xi * xj + xj * xj) * cr;
        sy12 += (yi * yi + yi * yj + yj * yj) * cr;
        sxy24 += (2.0 * xi * yi + xi * yj + xj * yi + 2.0 * xj * yj) * cr;
        minX = fmin(minX, xi);
        maxX = fmax(maxX, xi);
        minY = fmin(minY, yi);
        maxY = fmax(maxY, yi);
    }
    const double areaScale = fmax((maxX - minX) * (maxY - minY), 2.2250738585072014e-308);
    double mX = 0, mY = 0, mXX = 0, mYY = 0, mXY = 0;
    if (!(fabs(tw) <= 0x1p-40 * areaScale)) {
        const double inv = 2.0 / tw;
        mX = (fx6 / 6.0) * inv;
        mY = (fy6 / 6.0) * inv;
        mXX = (sx12 / 12.0) * inv;
        mYY = (sy12 / 12.0) * inv;
        mXY = (sxy24 / 24.0) * inv;
    }
    h[H_MOM + 0] = k * mX;
    h[H_MOM + 1] = k * mY;
    h[H_MOM + 2] = k * k * mXX;
    h[H_MOM + 3] = 2.0 * k * k * mXY;
    h[H_MOM + 4] = k * k * mYY;
    // Orders 3-5.  Over the triangle (0, P, Q) of each edge, with x = s Px +
    // t Qx, y = s Py + t Qy and the simplex integral of s^i t^j equal to
    // i! j! / (i + j + 2)!, the integral of x^a y^b is det(P, Q) times
    //   sum_{i<=a, j<=b} C(a,i) C(b,j) Px^i Qx^(a-i) Py^j Qy^(b-j)
    //                    (i+j)! (a+b-i-j)! / (a+b+2)!.
    // Summed over the edges it is the polygon's moment; divided by the area
    // (tw / 2), the mean <x^a y^b>.
    {
        const double fact[8] = {1, 1, 2, 6, 24, 120, 720, 5040};
        const double binom[6][6] = {{1, 0, 0, 0, 0, 0}, {1, 1, 0, 0, 0, 0},
                                    {1, 2, 1, 0, 0, 0}, {1, 3, 3, 1, 0, 0},
                                    {1, 4, 6, 4, 1, 0}, {1, 5, 10, 10, 5, 1}};
        const bool ok = !(fabs(tw) <= 0x1p-40 * areaScale);
        const int base_of[6] = {0, 0, 0, H_MOM3, H_MOM4, H_MOM5};
        double kn = k * k;
        for (int n = 3; n <= 5; ++n) {
            kn *= k;
            for (int a = 0; a <= n; ++a) {
                const int b2 = n - a;
                double sum = 0.0;
                for (int e = 0; e < nv; ++e) {
                    const int f = (e + 1 < nv) ? e + 1 : 0;
                    const double px = X(e), py = Y(e), qx = X(f), qy = Y(f);
                    const double det = px * qy - qx * py;
                    double acc = 0.0;
                    for (int i = 0; i <= a; ++i)
                        for (int j = 0; j <= b2; ++j)
                            acc += binom[a][i] * binom[b2][j] * pow(px, i) * pow(qx, a - i) *
                                   pow(py, j) * pow(qy, b2 - j) * fact[i + j] *
                                   fact[n - i - j];
                    sum += det * acc;
                }
                const double mean = ok ? sum / fact[n + 2] * (2.0 / tw) : 0.0;
                h[base_of[n] + a] = kn * binom[n][a] * mean / fact[n];
            }
        }
    }
    // Small-phase bound.  The vertex phases k (A x + B y) spread over at
    // least k |(A, B)| w, w the least width of the vertex set, so some phase
    // exceeds 1e-3 rad once max(|A|, |B|) > 2e-3 / (k w); the kernels test
    // the vertices one by one only below that (2 % margin for rounding).  w
    // is the least, over all vertex pairs, of the other vertices' spread
    // across the line through the pair: every edge of the convex hull is such
    // a pair, and the hull's least width lies across one of its edges.
    double wmin = __builtin_inf();
    if (nv <= 32) {
        for (int i = 0; i < nv; ++i) {
            const double xi = X(i), yi = Y(i);
            for (int j = i + 1; j < nv; ++j) {
                const double ex = X(j) - xi, ey = Y(j) - yi;
                const double len = sqrt(ex * ex + ey * ey);
                if (!(len > 0.0))
                    continue;
                double lo = 0.0, hi = 0.0;
                for (int l = 0; l < nv; ++l) {
                    const double c = ex * (Y(l) - yi) - ey * (X(l) - xi);
                    lo = fmin(lo, c);
                    hi = fmax(hi, c);
                }
                wmin = fmin(wmin, (hi - lo) / len);
            }
        }
    }
    // a degenerate or very large polygon keeps the per-vertex test everywhere
    h[H_PHB] = (wmin > 0.0 && wmin < __builtin_inf()) ? 2.04e-3 / (k * wmin)
                                                      : __builtin_inf();
    // the mixed kernels compare in FP32: the bound is kept as an FP32 value
    // in the slot's low word, so no conversion is needed per beam
    h[H_PHBM] = __builtin_bit_cast(
        double, (uint64_t)__builtin_bit_cast(
                    uint32_t, (float)(h[H_PHB] * (kSmallPhaseMixed / kSmallPhase))));
    // rotation vectors, pre-scaled by 1/2 (exact)
    for (int q = 0; q < 3; ++q) {
        h[H_T0 + q] = 0.5 * s[13 + q];      // NT
        h[H_T0 + 3 + q] = 0.5 * s[19 + q];  // nxDT
        h[H_T1 + q] = 0.5 * s[16 + q];      // NP
        h[H_T1 + 3 + q] = 0.5 * s[22 + q];  // nxDP
    }
    // J_phased, or for a rescaled set sign * a(sc) * J * exp(i k_j path sc),
    // a(sc) the mean of exp(c_abs p sc) over the absorption paths p (1 for
    // p <= DBL_EPSILON): the host rule of the multi-size pack
    double jpv[8];
    if (raw_j) {
        double amp = sign[b];
        const int64_t a0 = abs_off[b], a1 = abs_off[b + 1];
        if (a1 > a0) {
            double sum = 0.0;
            for (int64_t a = a0; a < a1; ++a)
                sum += abs_path[a] > 2.220446049250313e-16 ? exp(c_abs * abs_path[a] * sc) : 1.0;
            amp *= sum / (double)(a1 - a0);
        }
        double sn, cs;
        sincos(k_j * (path[b] * sc), &sn, &cs);
        for (int q = 0; q < 4; ++q) {
            const double jr = raw_j[(size_t)b * 8 + 2 * q] * amp;
            const double ji = raw_j[(size_t)b * 8 + 2 * q + 1] * amp;
            jpv[2 * q] = jr * cs - ji * sn;
            jpv[2 * q + 1] = jr * sn + ji * cs;
        }
    } else {
        for (int q = 0; q < 8; ++q)
            jpv[q] = s[25 + q];
    }
    for (int q = 0; q < 4; ++q) {  // complWave * J_phased
        const double jr = jpv[2 * q], ji = jpv[2 * q + 1];
        h[H_JP + 2 * q] = cw_re * jr - cw_im * ji;
        h[H_JP + 2 * q + 1] = cw_re * ji + cw_im * jr;
    }
    // MFMA operand: W[m][c] = T0[m] Jp[0][c] + T1[m] Jp[1][c] (FP64, then FP32)
    float *w = wb + (size_t)b * WB_STRIDE;
    for (int n = 0; n < 32; ++n) {
        double re = 0.0, im = 0.0;
        if (n < WB_COLS) {
            const int m = n >> 2, c = (n >> 1) & 1;
            const double t0 = h[H_T0 + m], t1 = h[H_T1 + m];
            const double wr = t0 * h[H_JP + 2 * c] + t1 * h[H_JP + 4 + 2 * c];
            const double wi = t0 * h[H_JP + 2 * c + 1] + t1 * h[H_JP + 4 + 2 * c + 1];
            const bool imag_col = n & 1;
            re = imag_col ? wi : wr;   // row 0 multiplies Re F
            im = imag_col ? wr : -wi;  // row 1 multiplies Im F
        }
        w[n] = (float)re;
        w[32 + n] = (float)im;
    }
    // FP32 copy of the rotation data the vector-ALU mixed kernel uses
    for (int q = 0; q < 6; ++q) {
        h32[H32_T0 + q] = (float)h[H_T0 + q];
        h32[H32_T1 + q] = (float)h[H_T1 + q];
    }
    for (int q = 0; q < 8; ++q)
        h32[H32_JP + q] = (float)h[H_JP + q];
    for (int i = 0; i < nv; ++i) {
        double *v = vtx + (size_t)(v0 + i) * VTX_STRIDE;
        const uint8_t f = ef[v0 + i];
        v[V_KX] = k * X(i);
        v[V_KY] = k * Y(i);
        v[V_SYX] = (f & 1) ? syx[v0 + i] : kInvalidSlope;
        v[V_SXY] = (f & 2) ? sxy[v0 + i] : kInvalidSlope;
        const int j = (i + 1 < nv) ? i + 1 : 0;
        double *w = vdd + (size_t)(v0 + i) * VDD_STRIDE;
        w[D_DX] = X(j) - X(i);
        w[D_DY] = Y(j) - Y(i);
        vflag[v0 + i] = f;
    }
}

// ---------------------------------------------------------------------------
// Per-(beam, direction) evaluation, FP64.  Returns false when the aperture
// integral is NaN (the reference skips the beam, HandlerPO.cpp:2742); else
// adds d = F R Jp to j[8] (and to jn[8] when ACC_NS).
//
// The two edge-sum branches of the reference (|B| > |A|: sum over edges with
// valid dx, coefficient A + (dy/dx) B, divide by B; otherwise edges with
// valid dy, coefficient A (dx/dy) + B, divide by -A) are evaluated by one
// loop whose operands are selected per lane, so lanes that take different
// branches do not diverge.  complWave is folded into J_phased and the
// small-phase area limit at packing time.
// ---------------------------------------------------------------------------
// A wave-uniform FP64 value kept in scalar registers.  Both edge slopes are
// read as scalars and the per-lane choice is made between registers; without
// this the compiler may select the address per lane instead, turning two
// scalar loads into one vector load on every edge.
__device__ __forceinline__ double uniform_f64(double x)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Every vertex phase within 1e-3 rad: the cell takes the small-phase moment
// form and its edge sum is not used, so the near-singular edges need not be
// evaluated.  That matters: a beam leaving along the incident direction
// (through parallel facets) meets it with A = B = 0 on the theta = 0 row,
// whose azimuth cells lie in as many different waves.
__device__ __forceinline__ bool all_phases_small(const double *__restrict__ vtx, int nv, double A,
                                                 double B, double limit = kSmallPhase)
{
    bool small = true;
    for (int e = 0; e < nv; ++e) {
        const double *v = vtx + (size_t)e * VTX_STRIDE;
        small = small && __builtin_fabs(__builtin_fma(A, v[V_KX], B * v[V_KY])) <= limit;
    }
    return small;
}

template <int NV>
__device__ __forceinline__ bool all_phases_small_nv(const double *kx, const double *ky, double A,
                                                    double B, double limit = kSmallPhase)
{
    bool small = true;
#pragma unroll
    for (int e = 0; e < NV; ++e)
        small = small && __builtin_fabs(__builtin_fma(A, kx[e], B * ky[e])) <= limit;
    return small;
}

// Lanes for which edge e enters the sum, as a wave mask: an edge with a
// valid dx in the lanes of mB (|B| > |A|), with a valid dy in the others.
// The flags are wave-uniform, so this is scalar mask arithmetic and the
// per-lane choice costs no vector instruction.  Used by the FP64 sums; in
// the mixed ones the extra scalar work measured slower than the two vector
// instructions it saves.
__device__ __forceinline__ uint64_t edge_valid_mask(int flags, uint64_t mB)
{
    const uint64_t mx = (flags & EDGE_VX) ? ~0ull : 0ull;
    const uint64_t my = (flags & EDGE_VY) ? ~0ull : 0ull;
    return (mB & mx) | (~mB & my);
}

// One edge of the aperture sum (FP64 phasors): the phasor of vertex e+1
// (vertex 0 again for the closing edge), the edge coefficient
// C = A + (dy/dx) B (|B| > |A|) or A (dx/dy) + B (otherwise) and its
// contribution (e^{i phi_{e+1}} - e^{i phi_e}) / C, or the exact half-angle
// form when |C| <= eps1.
__device__ __forceinline__ void edge_f64(int e, int n, const double *__restrict__ vtx,
                                         const double *__restrict__ vdd,
                                         const int32_t *__restrict__ vflag, double A, double B,
                                         double psi, bool useB, uint64_t mB,
                                         const double2 *__restrict__ tab, double k,
                                         double eps1, double s0, double c0, double &sp,
                                         double &cp, double &sr, double &si, bool small)
{
    const double *v = vtx + (size_t)e * VTX_STRIDE;
    // both slopes as scalars and both coefficients, then a per-lane select
    // (a select between the two loads becomes a divergent branch around each
    // scalar load)
    const double syx = uniform_f64(v[V_SYX]), sxy = uniform_f64(v[V_SXY]);
    const int f = __builtin_amdgcn_readfirstlane(vflag[e]);
    double sn = s0, cn = c0;
    if (e + 1 < n) {
        const double xn = __builtin_fma(B, v[VTX_STRIDE + V_KY], psi);
        sincos_table(__builtin_fma(A, v[VTX_STRIDE + V_KX], xn), tab, sn, cn);
    }
    const double cc = useB ? __builtin_fma(syx, B, A) : __builtin_fma(sxy, A, B);
    const uint64_t mvalid = edge_valid_mask(f, mB);
    const uint64_t mbig = __builtin_amdgcn_ballot_w64(__builtin_fabs(cc) > eps1);
    const double r = rcp_f64_fast(cc);
    const double inv = __builtin_amdgcn_inverse_ballot_w64(mvalid & mbig) ? r : 0.0;
    sr = __builtin_fma(cn - cp, inv, sr);
    si = __builtin_fma(sn - sp, inv, si);
    if (__builtin_expect(__builtin_amdgcn_inverse_ballot_w64(mvalid & ~mbig) && !small, 0)) {
        const double *w = vdd + (size_t)e * VDD_STRIDE;
        const double dd = useB ? uniform_f64(w[D_DX]) : uniform_f64(w[D_DY]);
        const double2 q = stable_edge_quotient(cp, sp, cn, sn, dd, cc, k);
        sr += q.x;
        si += q.y;
    }
    sp = sn;
    cp = cn;
}

// The edge sum of aperture_f64: both coefficients from scalar slopes, then
// a per-lane select; the lanes an edge enters as scalar masks.
template <int NV>
__device__ __forceinline__ void edge_sum_f64_nv(
    const double *syx, const double *sxy, const int *fl, const double *sv, const double *cv,
    const double *__restrict__ vdd, double A, double B, bool useB, uint64_t mB, double k,
    double eps1, bool small, double &sr, double &si)
{
#pragma unroll
    for (int e = 0; e < NV; ++e) {
        const int f = e + 1 < NV ? e + 1 : 0;
        const double cx = __builtin_fma(syx[e], B, A);
        const double cy = __builtin_fma(sxy[e], A, B);
        const double cc = useB ? cx : cy;
        const uint64_t mvalid = edge_valid_mask(fl[e], mB);
        const uint64_t mbig = __builtin_amdgcn_ballot_w64(__builtin_fabs(cc) > eps1);
        const double r = rcp_f64_fast(cc);
        const double inv = __builtin_amdgcn_inverse_ballot_w64(mvalid & mbig) ? r : 0.0;
        sr = __builtin_fma(cv[f] - cv[e], inv, sr);
        si = __builtin_fma(sv[f] - sv[e], inv, si);
        const bool rare = __builtin_amdgcn_inverse_ballot_w64(mvalid & ~mbig) && !small;
        if (__builtin_expect(rare, 0)) {
            const double *w = vdd + (size_t)e * VDD_STRIDE;
            const double dx = uniform_f64(w[D_DX]), dy = uniform_f64(w[D_DY]);
            const double2 q =
                stable_edge_quotient(cv[e], sv[e], cv[f], sv[f], useB ? dx : dy, cc, k);
            sr += q.x;
            si += q.y;
        }
    }
}

// Aperture integral F(beam, direction) / complWave in FP64 (A4).  NV > 0
// is the vertex count known at compile time (the loops are unrolled, the
// scalar loads issued up front and every vertex's table read issued before
// the first is used); NV = 0 loops over nv.  Returns false when
// the result is NaN (degenerate beam, skipped).
template <int NV>
__device__ __forceinline__ bool aperture_f64(const double *__restrict__ h,
                                             const double *__restrict__ vtx,
                                             const double *__restrict__ vdd,
                                             const int32_t *__restrict__ vflag, int nv,
                                             const double *d, const double2 *__restrict__ tab,
                                             double k, double eps1, double &fr, double &fi)
{
    // A, B: projections of k - k0 = b - d on the aperture axes
    const double A = h[H_A0] - (d[0] * h[H_HOR] + d[1] * h[H_HOR + 1] + d[2] * h[H_HOR + 2]);
    const double B = h[H_B0] - (d[0] * h[H_VER] + d[1] * h[H_VER + 1] + d[2] * h[H_VER + 2]);
    const double psi = d[0] * h[H_KC] + d[1] * h[H_KC + 1] + d[2] * h[H_KC + 2];
    const bool useB = __builtin_fabs(B) > __builtin_fabs(A);
    const uint64_t mB = __builtin_amdgcn_ballot_w64(useB);  // lanes summing the dx-valid edges
    // psi is folded into the vertex phasors (the integral is linear in each
    // phasor), as the reference CUDA path does (GpuDiffraction.cu:1110-1122)
    double sr = 0.0, si = 0.0;
    bool small;
    if constexpr (NV > 0) {
        // all vertex phasors first (scalar loads, then every table read), so
        // the LDS latency of one vertex overlaps the others; then the edges
        double kx[NV], ky[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            kx[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_KX]);
            ky[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_KY]);
        }
        const double phb = uniform_f64(h[H_PHB]);
        double dv[NV];
        double2 ev[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e)
            sincos_reduce(__builtin_fma(A, kx[e], __builtin_fma(B, ky[e], psi)), tab, dv[e],
                          ev[e]);
        double sv[NV], cv[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e)
            sincos_finish(dv[e], ev[e], sv[e], cv[e]);
        // the vertex phases one by one only where the beam's bound leaves
        // the small-phase form possible, after the table reads are consumed
        small = false;
        if (__builtin_expect(__builtin_fabs(useB ? B : -A) <= phb, 0))
            small = all_phases_small_nv<NV>(kx, ky, A, B);
        double syx[NV], sxy[NV];
        int fl[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            syx[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_SYX]);
            sxy[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_SXY]);
            fl[e] = __builtin_amdgcn_readfirstlane(vflag[e]);
        }
        edge_sum_f64_nv<NV>(syx, sxy, fl, sv, cv, vdd, A, B, useB, mB, k, eps1, small, sr, si);
    } else {
        small = false;
        if (__builtin_expect(__builtin_fabs(useB ? B : -A) <= h[H_PHB], 0))
            small = all_phases_small(vtx, nv, A, B);
        double s0, c0;
        sincos_table(__builtin_fma(A, vtx[V_KX], __builtin_fma(B, vtx[V_KY], psi)), tab, s0, c0);
        double sp = s0, cp = c0;
        for (int e = 0; e < nv; ++e)
            edge_f64(e, nv, vtx, vdd, vflag, A, B, psi, useB, mB, tab, k, eps1, s0,
                     c0, sp, cp, sr, si, small);
    }
    const double nrm = rcp_f64(useB ? B : -A);
    fr = sr * nrm;
    fi = si * nrm;
    if (__builtin_expect(small, 0)) {
        // small-phase moment form (HandlerPO_fast.h:161-216); the area limit
        // is stored divided by complWave
        const double p1 = A * h[H_MOM] + B * h[H_MOM + 1];
        const double p2 = A * A * h[H_MOM + 2] + A * B * h[H_MOM + 3] + B * B * h[H_MOM + 4];
        const double gr = 1.0 - 0.5 * p2, gi = p1;
        const double ar = h[H_AL] * gr - h[H_AL + 1] * gi;
        const double ai = h[H_AL] * gi + h[H_AL + 1] * gr;
        double sd, cd;
        sincos_table(psi, tab, sd, cd);
        fr = ar * cd - ai * sd;
        fi = ar * sd + ai * cd;
    }
    return !__builtin_isnan(fr);
}

// Per-beam aperture integral dispatched on the vertex count (3 and 4 cover
// most beams of a hexagonal particle).
__device__ __forceinline__ bool aperture_f64_any(const double *__restrict__ h,
                                                 const double *__restrict__ vtx,
                                                 const double *__restrict__ vdd,
                                                 const int32_t *__restrict__ vflag, int nv,
                                                 const double *d,
                                                 const double2 *__restrict__ tab, double k,
                                                 double eps1, double &fr, double &fi)
{
    if (nv == 4)
        return aperture_f64<4>(h, vtx, vdd, vflag, 4, d, tab, k, eps1, fr, fi);
    if (nv == 3)
        return aperture_f64<3>(h, vtx, vdd, vflag, 3, d, tab, k, eps1, fr, fi);
    return aperture_f64<0>(h, vtx, vdd, vflag, nv, d, tab, k, eps1, fr, fi);
}

template <int ACC_NS>
__device__ __forceinline__ void beam_dir_f64(
    const double *__restrict__ h, const double *__restrict__ vtx,
    const double *__restrict__ vdd, const int32_t *__restrict__ vflag, int nv,
    const double *d, const double *__restrict__ su, const double2 *__restrict__ tab,
    double k, double eps1, double *j, double *jn)
{
    double fr, fi;
    if (!aperture_f64_any(h, vtx, vdd, vflag, nv, d, tab, k, eps1, fr, fi))
        return;
    // rotation vectors of this lane's direction, from LDS ([12][256])
    const int l = threadIdx.x;
    double u0[6], u1[6];
    for (int q = 0; q < 6; ++q) {
        u0[q] = su[q * 256 + l];
        u1[q] = su[(6 + q) * 256 + l];
    }
    const double *T0 = h + H_T0, *T1 = h + H_T1;
    const double r00 = u0[0] * T0[0] + u0[1] * T0[1] + u0[2] * T0[2] + u0[3] * T0[3] + u0[4] * T0[4] + u0[5] * T0[5];
    const double r01 = u0[0] * T1[0] + u0[1] * T1[1] + u0[2] * T1[2] + u0[3] * T1[3] + u0[4] * T1[4] + u0[5] * T1[5];
    const double r10 = u1[0] * T0[0] + u1[1] * T0[1] + u1[2] * T0[2] + u1[3] * T0[3] + u1[4] * T0[4] + u1[5] * T0[5];
    const double r11 = u1[0] * T1[0] + u1[1] * T1[1] + u1[2] * T1[2] + u1[3] * T1[3] + u1[4] * T1[4] + u1[5] * T1[5];
    const double *jp = h + H_JP;  // complWave * J_phased
    // q = R Jp (real 2x2 times complex 2x2), then d = F q
    double q[8];
    q[0] = r00 * jp[0] + r01 * jp[4];
    q[1] = r00 * jp[1] + r01 * jp[5];
    q[2] = r00 * jp[2] + r01 * jp[6];
    q[3] = r00 * jp[3] + r01 * jp[7];
    q[4] = r10 * jp[0] + r11 * jp[4];
    q[5] = r10 * jp[1] + r11 * jp[5];
    q[6] = r10 * jp[2] + r11 * jp[6];
    q[7] = r10 * jp[3] + r11 * jp[7];
    if constexpr (ACC_NS) {
        for (int t = 0; t < 4; ++t) {
            const double dr = fr * q[2 * t] - fi * q[2 * t + 1];
            const double di = fr * q[2 * t + 1] + fi * q[2 * t];
            j[2 * t] += dr;
            j[2 * t + 1] += di;
            jn[2 * t] += dr;
            jn[2 * t + 1] += di;
        }
    } else {
        for (int t = 0; t < 4; ++t) {
            j[2 * t] = __builtin_fma(fr, q[2 * t], __builtin_fma(-fi, q[2 * t + 1], j[2 * t]));
            j[2 * t + 1] = __builtin_fma(fr, q[2 * t + 1], __builtin_fma(fi, q[2 * t], j[2 * t + 1]));
        }
    }
}

// One edge of the aperture sum, mixed precision: phase arguments, their
// reduction and the edge coefficient in FP64, phasors and sums in FP32.
__device__ __forceinline__ void edge_mixed(int e, int n, const double *__restrict__ vtx,
                                           const double *__restrict__ vdd,
                                           const int32_t *__restrict__ vflag, double A,
                                           double B, double psi, bool useB, int vmask,
                                           const float4 *__restrict__ tab, double k,
                                           double eps1, float s0, float c0, float &sp,
                                           float &cp, float &sr, float &si, bool small)
{
    const double *v = vtx + (size_t)e * VTX_STRIDE;
    const double syx = uniform_f64(v[V_SYX]), sxy = uniform_f64(v[V_SXY]);
    const int f = __builtin_amdgcn_readfirstlane(vflag[e]);
    float sn = s0, cn = c0;
    if (e + 1 < n) {
        const double xn = __builtin_fma(B, v[VTX_STRIDE + V_KY], psi);
        sincos_table_f32(__builtin_fma(A, v[VTX_STRIDE + V_KX], xn), tab, sn, cn);
    }
    // the edge coefficient is cancellation-prone: FP64
    const double cc = useB ? __builtin_fma(syx, B, A) : __builtin_fma(sxy, A, B);
    const bool big = __builtin_fabs(cc) > eps1;
    // hardware reciprocal (1 ulp): the FP32 phasor differences it scales
    // carry more rounding than that; an invalid edge's stored slope makes
    // it exactly zero
    const float r = __builtin_amdgcn_rcpf((float)cc);
    const float inv = big ? r : 0.0f;
    sr = __builtin_fmaf(cn - cp, inv, sr);
    si = __builtin_fmaf(sn - sp, inv, si);
    if (__builtin_expect(!big && !small && (f & vmask) != 0, 0)) {
        const double *w = vdd + (size_t)e * VDD_STRIDE;
        const double dd = useB ? uniform_f64(w[D_DX]) : uniform_f64(w[D_DY]);
        const double2 q =
            stable_edge_quotient((double)cp, (double)sp, (double)cn, (double)sn, dd, cc, k);
        sr += (float)q.x;
        si += (float)q.y;
    }
    sp = sn;
    cp = cn;
}

// The end of the mixed aperture integral: the 1/B or -1/A normalisation,
// the small-phase moment form when every vertex phase is within 0.25 rad,
// and a NaN result (degenerate beam) returned as 0, i.e. the beam skipped.
// The normaliser is the hardware reciprocal (1 ulp), as for the edge
// coefficients: the FP32 edge sum it scales carries more rounding than that.
// Word 3 of a raw buffer resource on gfx9-family devices (32-bit data
// format; the element and index fields unused).
constexpr int kBufferRsrcWord3 = 0x00020000;

// The FP32 small-phase bound stored in the low word of header slot H_PHBM.
__device__ __forceinline__ float phase_bound_f32(double slot)
{
    return __builtin_bit_cast(float, low_bits(slot));
}

__device__ __forceinline__ void aperture_mixed_finish(const double *__restrict__ h, double A,
                                                      double B, double psi, float nin,
                                                      bool small, float sr, float si,
                                                      const float4 *__restrict__ tab,
                                                      float &out_r, float &out_i)
{
    const float nrm = __builtin_amdgcn_rcpf(nin);  // nin: B or -A
    float fr = sr * nrm, fi = si * nrm;
    if (__builtin_expect(small, 0)) {
        // <e^{i phi}> = sum_n i^n <phi^n> / n! to fifth order
        const double p1 = A * h[H_MOM] + B * h[H_MOM + 1];
        const double p2 = A * A * h[H_MOM + 2] + A * B * h[H_MOM + 3] + B * B * h[H_MOM + 4];
        const double A2 = A * A, B2 = B * B, AB = A * B;
        const double *c3 = h + H_MOM3, *c4 = h + H_MOM4, *c5 = h + H_MOM5;
        const double q3 = B * (B2 * c3[0] + AB * c3[1]) + A * (AB * c3[2] + A2 * c3[3]);
        const double q4 = B2 * (B2 * c4[0] + AB * c4[1]) + A2 * (B2 * c4[2] + AB * c4[3] + A2 * c4[4]);
        const double q5 = B * B2 * (B2 * c5[0] + AB * c5[1] + A2 * c5[2]) +
                          A * A2 * (B2 * c5[3] + AB * c5[4] + A2 * c5[5]);
        const double gr = 1.0 - 0.5 * p2 + q4, gi = p1 - q3 + q5;
        const double ar = h[H_AL] * gr - h[H_AL + 1] * gi;
        const double ai = h[H_AL] * gi + h[H_AL + 1] * gr;
        float sd, cd;
        sincos_table_f32(psi, tab, sd, cd);
        fr = (float)ar * cd - (float)ai * sd;
        fi = (float)ar * sd + (float)ai * cd;
    }
    // a NaN F (degenerate polygon) counts as zero; the selects only in
    // waves that have one
    out_r = fr;
    out_i = fi;
    const bool bad = __builtin_isnan(fr);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(bad) != 0, 0)) {
        out_r = bad ? 0.0f : fr;
        out_i = bad ? 0.0f : fi;
    }
}

// The edge sum of aperture_mixed_nv.  Both coefficients come from scalar
// slopes, then a per-lane select of the result (selecting the slope first
// costs two moves and a select per half of the FP64 operand).  An edge that
// is invalid for the lane's branch carries kInvalidSlope, so its reciprocal
// is zero and the sum needs no validity test.  A separate pass for waves
// whose lanes all take one branch of the reference's sum (one coefficient
// per edge; about half the waves of an orientation average) gained 0.4 %
// for three copies of the loop, and none in the FP64 sum.  Near-singular edges
// (|C| <= eps1, their term below zero) are evaluated in one pass after the
// sum, taken only when some lane has one; the validity test lives there.
template <int NV>
__device__ __forceinline__ void edge_sum_mixed_nv(
    const double *syx, const double *sxy, const int *fl, const float *sv, const float *cv,
    const double *__restrict__ vdd, double A, double B, bool useB, int vmask, double k,
    double eps1, bool small, float &sr, float &si)
{
    bool rare[NV];
    bool any_rare = false;
    // every term taken as it comes: a lane with a near-singular edge (its
    // reciprocal huge or infinite) sums again below, so the common sum
    // carries no select
#pragma unroll
    for (int e = 0; e < NV; ++e) {
        const int f = e + 1 < NV ? e + 1 : 0;
        const double cx = __builtin_fma(syx[e], B, A);
        const double cy = __builtin_fma(sxy[e], A, B);
        const double cc = useB ? cx : cy;
        const bool big = __builtin_fabs(cc) > eps1;
        // hardware reciprocal (1 ulp): the FP32 phasor differences it scales
        // carry more rounding than that
        const float r = __builtin_amdgcn_rcpf((float)cc);
        sr = __builtin_fmaf(cv[f] - cv[e], r, sr);
        si = __builtin_fmaf(sv[f] - sv[e], r, si);
        // near-singular lanes as scalar mask arithmetic (a second vector
        // compare for !big otherwise)
        rare[e] = __builtin_amdgcn_inverse_ballot_w64(~__builtin_amdgcn_ballot_w64(big));
        any_rare = any_rare || rare[e];
    }
    // (lanes on the moment form discard the sum)
    if (__builtin_expect(any_rare && !small, 0)) {
        // the sum without the near-singular terms (the same operations, so
        // the same bits as a masked first pass) ...
        sr = 0.0f;
        si = 0.0f;
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int f = e + 1 < NV ? e + 1 : 0;
            const double cc =
                useB ? __builtin_fma(syx[e], B, A) : __builtin_fma(sxy[e], A, B);
            const float inv = rare[e] ? 0.0f : __builtin_amdgcn_rcpf((float)cc);
            sr = __builtin_fmaf(cv[f] - cv[e], inv, sr);
            si = __builtin_fmaf(sv[f] - sv[e], inv, si);
        }
        // ... and those terms in the stable form
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            if (rare[e] && (fl[e] & vmask) != 0) {
                const int f = e + 1 < NV ? e + 1 : 0;
                const double cc =
                    useB ? __builtin_fma(syx[e], B, A) : __builtin_fma(sxy[e], A, B);
                const double *w = vdd + (size_t)e * VDD_STRIDE;
                const double dx = uniform_f64(w[D_DX]), dy = uniform_f64(w[D_DY]);
                const double2 q = stable_edge_quotient((double)cv[e], (double)sv[e],
                                                       (double)cv[f], (double)sv[f],
                                                       useB ? dx : dy, cc, k);
                sr += (float)q.x;
                si += (float)q.y;
            }
        }
    }
}

// Aperture integral F(beam, direction) / complWave, mixed precision: A, B,
// the phases and their reduction, the edge coefficients and the moments in
// FP64; phasors and edge sums in FP32.
//
// NV > 0: the vertex count is known at compile time.  The phases of all
// vertices are reduced and their table reads issued together, so the LDS
// latency of one vertex overlaps the arithmetic of the others; the edges
// follow.
template <int NV>
__device__ __forceinline__ void aperture_mixed_nv(
    const double *__restrict__ h, const double *__restrict__ vtx,
    const double *__restrict__ vdd, const int32_t *__restrict__ vflag, int nv,
    const double *d, const float4 *__restrict__ tab, double k, double eps1, float &out_r,
    float &out_i)
{
    // a product sum, then one subtraction: an FMA chain ending on A0 would
    // pair two scalar operands in one instruction and cost two moves
    const double A = h[H_A0] - (d[0] * h[H_HOR] + d[1] * h[H_HOR + 1] + d[2] * h[H_HOR + 2]);
    const double B = h[H_B0] - (d[0] * h[H_VER] + d[1] * h[H_VER + 1] + d[2] * h[H_VER + 2]);
    const double psi = d[0] * h[H_KC] + d[1] * h[H_KC + 1] + d[2] * h[H_KC + 2];
    const bool useB = __builtin_fabs(B) > __builtin_fabs(A);
    const int vmask = useB ? EDGE_VX : EDGE_VY;
    const float nin = useB ? (float)B : -(float)A;
    float sr = 0.0f, si = 0.0f;
    bool small;
    if constexpr (NV > 0) {
        // scalar loads first: LDS reads and scalar loads share one wait
        // counter, so a scalar load between two table reads would drain both
        double kx[NV], ky[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            kx[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_KX]);
            ky[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_KY]);
        }
        // the moment form's bound (loaded with the header's first fields,
        // so no scalar load waits among the table reads)
        const float bnd = phase_bound_f32(uniform_f64(h[H_PHBM]));
        float dv[NV];
        float4 ev[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e)
            sincos_reduce_f32(__builtin_fma(A, kx[e], __builtin_fma(B, ky[e], psi)), tab, dv[e],
                              ev[e]);
        float sv[NV], cv[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e)
            sincos_finish_f32(dv[e], ev[e], sv[e], cv[e]);
        // the vertex phases one by one only in waves where the beam's bound
        // leaves the moment form possible for some lane
        small = false;
        const bool maybe = __builtin_fabsf(nin) <= bnd;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(maybe) != 0, 0))
            small = maybe && all_phases_small_nv<NV>(kx, ky, A, B, kSmallPhaseMixed);
        double syx[NV], sxy[NV];
        int fl[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            syx[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_SYX]);
            sxy[e] = uniform_f64(vtx[(size_t)e * VTX_STRIDE + V_SXY]);
            fl[e] = __builtin_amdgcn_readfirstlane(vflag[e]);
        }
        edge_sum_mixed_nv<NV>(syx, sxy, fl, sv, cv, vdd, A, B, useB, vmask, k, eps1, small, sr,
                              si);
    } else {
        small = false;
        if (__builtin_expect(__builtin_fabsf(nin) <= phase_bound_f32(h[H_PHBM]), 0))
            small = all_phases_small(vtx, nv, A, B, kSmallPhaseMixed);
        float s0, c0;
        sincos_table_f32(__builtin_fma(A, vtx[V_KX], __builtin_fma(B, vtx[V_KY], psi)), tab, s0,
                         c0);
        float sp = s0, cp = c0;
        for (int e = 0; e < nv; ++e)
            edge_mixed(e, nv, vtx, vdd, vflag, A, B, psi, useB, vmask, tab, k, eps1,
                       s0, c0, sp, cp, sr, si, small);
    }
    aperture_mixed_finish(h, A, B, psi, nin, small, sr, si, tab, out_r, out_i);
}

__device__ __forceinline__ void aperture_mixed(
    const double *__restrict__ h, const double *__restrict__ vtx,
    const double *__restrict__ vdd, const int32_t *__restrict__ vflag, int nv,
    const double *d, const float4 *__restrict__ tab, double k, double eps1, float &out_r,
    float &out_i)
{
    if (nv == 4)
        aperture_mixed_nv<4>(h, vtx, vdd, vflag, 4, d, tab, k, eps1, out_r, out_i);
    else if (nv == 3)
        aperture_mixed_nv<3>(h, vtx, vdd, vflag, 3, d, tab, k, eps1, out_r, out_i);
    else
        aperture_mixed_nv<0>(h, vtx, vdd, vflag, nv, d, tab, k, eps1, out_r, out_i);
}

// Mixed precision with the rotation and coherent sum on the vector ALU.
template <int ACC_NS>
__device__ __forceinline__ void beam_dir_mixed(
    const double *__restrict__ h, const float *__restrict__ h32,
    const double *__restrict__ vtx, const double *__restrict__ vdd,
    const int32_t *__restrict__ vflag, int nv, const double *d,
    const float *__restrict__ su, const float4 *__restrict__ tab, double k, double eps1,
    float *j, float *jn)
{
    float fr, fi;
    aperture_mixed(h, vtx, vdd, vflag, nv, d, tab, k, eps1, fr, fi);
    if (fr == 0.0f && fi == 0.0f)
        return;
    const int l = threadIdx.x;
    float u0[6], u1[6];
    for (int q = 0; q < 6; ++q) {
        u0[q] = su[q * 256 + l];
        u1[q] = su[(6 + q) * 256 + l];
    }
    const float *T0 = h32 + H32_T0, *T1 = h32 + H32_T1;
    const float r00 = u0[0] * T0[0] + u0[1] * T0[1] + u0[2] * T0[2] + u0[3] * T0[3] + u0[4] * T0[4] + u0[5] * T0[5];
    const float r01 = u0[0] * T1[0] + u0[1] * T1[1] + u0[2] * T1[2] + u0[3] * T1[3] + u0[4] * T1[4] + u0[5] * T1[5];
    const float r10 = u1[0] * T0[0] + u1[1] * T0[1] + u1[2] * T0[2] + u1[3] * T0[3] + u1[4] * T0[4] + u1[5] * T0[5];
    const float r11 = u1[0] * T1[0] + u1[1] * T1[1] + u1[2] * T1[2] + u1[3] * T1[3] + u1[4] * T1[4] + u1[5] * T1[5];
    const float *jp = h32 + H32_JP;
    float q[8];
    q[0] = r00 * jp[0] + r01 * jp[4];
    q[1] = r00 * jp[1] + r01 * jp[5];
    q[2] = r00 * jp[2] + r01 * jp[6];
    q[3] = r00 * jp[3] + r01 * jp[7];
    q[4] = r10 * jp[0] + r11 * jp[4];
    q[5] = r10 * jp[1] + r11 * jp[5];
    q[6] = r10 * jp[2] + r11 * jp[6];
    q[7] = r10 * jp[3] + r11 * jp[7];
    if constexpr (ACC_NS) {
        for (int t = 0; t < 4; ++t) {
            const float dr = fr * q[2 * t] - fi * q[2 * t + 1];
            const float di = fr * q[2 * t + 1] + fi * q[2 * t];
            j[2 * t] += dr;
            j[2 * t + 1] += di;
            jn[2 * t] += dr;
            jn[2 * t + 1] += di;
        }
    } else {
        for (int t = 0; t < 4; ++t) {
            j[2 * t] = __builtin_fmaf(fr, q[2 * t], __builtin_fmaf(-fi, q[2 * t + 1], j[2 * t]));
            j[2 * t + 1] = __builtin_fmaf(fr, q[2 * t + 1], __builtin_fmaf(fi, q[2 * t], j[2 * t + 1]));
        }
    }
}

// XCD-aware block -> (item, tile): consecutive workgroup ids land on the 8
// XCDs round-robin; remapping gives each XCD a contiguous range of items so
// the beam records of an item are shared through that XCD's L2.
__device__ __forceinline__ bool map_block(int n_items, int n_tiles, int &item,
                                          int &tile)
{
    const int total = n_items * n_tiles;
    const int nblk = gridDim.x;
    const int per_xcd = (nblk + 7) / 8;
    const int bid = blockIdx.x;
    const int logical = (bid % 8) * per_xcd + bid / 8;
    if (logical >= total)
        return false;
    item = logical / n_tiles;
    tile = logical - item * n_tiles;
    return true;
}

// ---------------------------------------------------------------------------
// Device work queue (work mode MBS_WORK_QUEUE).  plan_segments_kernel cuts
// every orientation into beam segments on the device (the host plan's rule:
// segments of max(32, ceil(n_beams / want)) beams, at least one per
// orientation) and splits the (item, tile) units into 8 contiguous ranges,
// one per XCD, so the units of one item stay in one XCD's L2.  A persistent
// workgroup takes units from its XCD's range with an atomic counter and,
// when that range is empty, from the next XCDs' in turn; it exits when all
// eight are empty.  Each unit still writes its own partial sums, so results
// do not depend on which workgroup ran it.
// ---------------------------------------------------------------------------
struct QueueCursor {
    int range = -1;   // XCD range being drained (thread 0 only)
    int tried = 0;    // ranges found empty so far
};

__device__ __forceinline__ bool next_unit(int *__restrict__ queue, int n_tiles, QueueCursor &qc,
                                          int &s_unit, int &item, int &tile)
{
    if (threadIdx.x == 0) {
        if (qc.range < 0)
            qc.range = blockIdx.x & 7;
        int u = -1;
        while (qc.tried < 8) {
            const int t = atomicAdd(queue + Q_NEXT + qc.range, 1);
            if (t < queue[Q_END + qc.range]) {
                u = t;
                break;
            }
            qc.range = (qc.range + 1) & 7;
            ++qc.tried;
        }
        s_unit = u;
    }
    __syncthreads();
    const int u = s_unit;
    __syncthreads();
    if (u < 0)
        return false;
    item = u / n_tiles;
    tile = u - item * n_tiles;
    return true;
}

// One workgroup: the segment plan of the resident batch, and the queue
// header.  ooff [n_orient + 1] device orientation offsets; want: the target
// item count; items has room for want + n_orient items (the plan never
// needs more: sum over o of ceil(n_o / seg) <= n_beams / seg + n_orient).
__global__ void __launch_bounds__(1024) plan_segments_kernel(
    const int64_t *__restrict__ ooff, int n_orient, int want, int n_tiles,
    WorkItem *__restrict__ items, int *__restrict__ queue)
{
    __shared__ int s_scan[1024];
    __shared__ int s_base;
    const int64_t nb = ooff[n_orient];
    const int64_t seg =
        max((int64_t)32, (nb + (int64_t)want - 1) / max((int64_t)want, (int64_t)1));
    if (threadIdx.x == 0)
        s_base = 0;
    __syncthreads();
    for (int o0 = 0; o0 < n_orient; o0 += blockDim.x) {
        const int o = o0 + threadIdx.x;
        int64_t b0 = 0, n = 0;
        int kseg = 0;
        if (o < n_orient) {
            b0 = ooff[o];
            n = ooff[o + 1] - b0;
            kseg = (int)max((int64_t)1, (n + seg - 1) / seg);
        }
        // inclusive scan of the segment counts over this chunk
        s_scan[threadIdx.x] = kseg;
        __syncthreads();
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            const int v = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0;
            __syncthreads();
            s_scan[threadIdx.x] += v;
            __syncthreads();
        }
        const int first = s_base + s_scan[threadIdx.x] - kseg;
        for (int i = 0; i < kseg; ++i)
            items[first + i] = WorkItem{o, o + 1, b0 + n * i / kseg, b0 + n * (i + 1) / kseg};
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            s_base += s_scan[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int total = s_base;
        const int64_t units = (int64_t)total * n_tiles;
        queue[Q_ITEMS] = total;
        for (int x = 0; x < 8; ++x) {
            queue[Q_NEXT + x] = (int)(units * x / 8);
            queue[Q_END + x] = (int)(units * (x + 1) / 8);
        }
    }
}

struct KernelArgs {
    const WorkItem *items;
    int n_items, n_tiles, cells, n_az;
    const double *tab;       // grid tables [DT_COUNT][cells]
    const double *hdr;       // [n_beams][HDR_STRIDE]
    const float *hdr32;      // [n_beams][HDR32_STRIDE]
    const float *wb;         // [n_beams][WB_STRIDE]
    const double *vtx;       // [n_vert][VTX_STRIDE]
    const double *vdd;       // [n_vert][VDD_STRIDE]
    const int32_t *vflag;    // [n_vert]
    const int32_t *nv;       // [n_beams]
    const int64_t *voff;     // [n_beams+1]
    const int32_t *ext;      // [n_beams]
    const int64_t *ooff;     // [n_orient+1]
    const double *weights;   // [n_orient]
    double k, eps1;
    double *out;             // Mueller mode: [n_items][cells][16]; Jones: [n_items][cells][8]
    double *out_ns;          // same shape, no-shadow (NS only)
    double *per_orient;      // [n_orient][cells][16] (PER only, Mueller mode)
    int *queue;              // work-queue header (queue mode), else null
    const float *tab32;      // FP32 u0, u1 [cells][DT32_COUNT] (mixed kernels)
    const int32_t *sidx;     // [n_beams] spectrum slot of each beam, -1: direct
    const float2 *fbuf;      // [spectrum beams][cells] F from the spectra
    int waves;               // MFMA kernel occupancy variant (0: the default)
};

// Mueller of one orientation's Jones sum, weighted, stored to (first) or
// added to (later orientations of the item) a lane-owned global cell: the
// running orientation sum lives in memory, not in registers -- it is touched
// once per orientation, against thousands of beams of work in between.
template <typename T>
__device__ __forceinline__ void flush_mueller(const T *j, double w, double *dst, bool first,
                                              double *per)
{
    double mm[16];
    for (int q = 0; q < 16; ++q)
        mm[q] = 0.0;
    add_mueller(mm, j, w);
    if (per)
        for (int q = 0; q < 16; ++q)
            per[q] = mm[q];
    if (first) {
        for (int q = 0; q < 16; ++q)
            dst[q] = mm[q];
    } else {
        for (int q = 0; q < 16; ++q)
            dst[q] += mm[q];
    }
}

// All pointers are separate __restrict__ kernel arguments: the per-orientation
// stores into `out` must be provably disjoint from the beam records for the
// compiler to keep reading those through the scalar cache.
template <int PREC, int NS, int JONES, int PER, int QUEUE>
__global__ __launch_bounds__(256) MBS_OCCUPANCY void diffract_kernel(
    const WorkItem *__restrict__ items, int n_items, int n_tiles, int cells, int n_az,
    const double *__restrict__ tab, const double *__restrict__ hdr,
    const float *__restrict__ hdr32, const double *__restrict__ vtx,
    const double *__restrict__ vdd, const int32_t *__restrict__ vflag,
    const int32_t *__restrict__ nv,
    const int64_t *__restrict__ voff, const int32_t *__restrict__ ext,
    const int64_t *__restrict__ ooff, const double *__restrict__ weights, double k,
    double eps1, double *__restrict__ out, double *__restrict__ out_ns,
    double *__restrict__ per_orient, int *__restrict__ queue, const float *__restrict__ tab32)
{
    using U = typename std::conditional<PREC == 1, float, double>::type;
    using U2 = typename std::conditional<PREC == 1, float4, double2>::type;
    // LDS: this lane's rotation vectors u0, u1 ([12][256], read once per beam
    // after the vertex loop instead of occupying 12-24 VGPRs through it) and
    // the phasor table e^{i 2 pi j/TN} (512 FP32 entries, 256 FP64 ones)
    constexpr int TN = PREC == 1 ? kTabN32 : kTabN;
    __shared__ U s_u[12 * 256];
    __shared__ U2 s_tab[TN];
    __shared__ int s_unit;
    int item = 0, tile = 0;
    if constexpr (!QUEUE) {
        if (!map_block(n_items, n_tiles, item, tile))
            return;
    }
    for (int i = threadIdx.x; i < TN; i += blockDim.x) {
        if constexpr (PREC == 1) {
            s_tab[i] = tab32_entry(i);
        } else {
            double ts, tc;
            sincospi((double)i / (TN / 2), &ts, &tc);
            s_tab[i] = make_double2(tc, ts);
        }
    }
    QueueCursor qc;
    for (;;) {
    if constexpr (QUEUE) {
        if (!next_unit(queue, n_tiles, qc, s_unit, item, tile))
            return;
    }
    const int cell = tile * 256 + threadIdx.x;
    const bool live = cell < cells;
    const int c = live ? cell : cells - 1;  // dead lanes compute, never store
    double d[3];
    for (int q = 0; q < 3; ++q)
        d[q] = tab[(size_t)(DT_DIR + q) * cells + c];
    for (int q = 0; q < 12; ++q) {
        if constexpr (PREC == 1)
            s_u[q * 256 + threadIdx.x] = tab32[(size_t)c * DT32_COUNT + q];
        else
            s_u[q * 256 + threadIdx.x] = tab[(size_t)(DT_U0 + q) * cells + c];
    }
    __syncthreads();
    const WorkItem it = items[item];
    for (int o = it.o_begin; o < it.o_end; ++o) {
        int64_t b0 = ooff[o], b1 = ooff[o + 1];
        if constexpr (JONES) {
            b0 = it.b_begin;
            b1 = it.b_end;
        }
        U j[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        U jn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        // the next beam's vertex offset and count are loaded one beam
        // ahead, so a beam's header and vertex loads issue together
        int64_t v0_next = b0 < b1 ? voff[b0] : 0;
        int nv_next = b0 < b1 ? nv[b0] : 0;
        const int nbeam = (int)(b1 - b0);  // 32-bit loop test (scalar unit)
        for (int ib = 0; ib < nbeam; ++ib) {
            const int64_t b = b0 + ib;
            const double *h = hdr + (size_t)b * HDR_STRIDE;
            const int64_t v0 = v0_next;
            const int nvb = nv_next;
            if (ib + 1 < nbeam) {
                v0_next = voff[b + 1];
                nv_next = nv[b + 1];
            }
            const double *bv = vtx + (size_t)v0 * VTX_STRIDE;
            const double *bd = vdd + (size_t)v0 * VDD_STRIDE;
            if constexpr (PREC == 0) {
                if (NS && ext[b] == 0)
                    beam_dir_f64<1>(h, bv, bd, vflag + v0, nvb, d, s_u, s_tab, k, eps1, j, jn);
                else
                    beam_dir_f64<0>(h, bv, bd, vflag + v0, nvb, d, s_u, s_tab, k, eps1, j,
                                    nullptr);
            } else {
                const float *h32 = hdr32 + (size_t)b * HDR32_STRIDE;
                if (NS && ext[b] == 0)
                    beam_dir_mixed<1>(h, h32, bv, bd, vflag + v0, nvb, d, s_u, s_tab, k, eps1,
                                      j, jn);
                else
                    beam_dir_mixed<0>(h, h32, bv, bd, vflag + v0, nvb, d, s_u, s_tab, k, eps1,
                                      j, nullptr);
            }
        }
        if (!live)
            continue;
        if constexpr (JONES) {
            // Jones partials in the kernel's own precision (FP32 for the
            // mixed kernel: half the bytes, the same values)
            U *dst = reinterpret_cast<U *>(out) + ((size_t)item * cells + cell) * 8;
            for (int q = 0; q < 8; ++q)
                dst[q] = j[q];
            if constexpr (NS) {
                U *dn = reinterpret_cast<U *>(out_ns) + ((size_t)item * cells + cell) * 8;
                for (int q = 0; q < 8; ++q)
                    dn[q] = jn[q];
            }
        } else {
            const double w = weights[o];
            const bool first = (o == it.o_begin);
            double *per = PER ? per_orient + ((size_t)o * cells + abi_cell(cell, n_az, cells)) * 16
                              : nullptr;
            flush_mueller(j, w, out + ((size_t)item * cells + cell) * 16, first, per);
            if constexpr (NS)
                flush_mueller(jn, w, out_ns + ((size_t)item * cells + cell) * 16, first,
                              (double *)nullptr);
        }
    }
    if constexpr (!QUEUE)
        return;
    __syncthreads();  // the next unit overwrites s_u
    }
}

// Incoherent sum (MBS_INCOHERENT), FP64: per orientation the Mueller
// matrices of the single beams are added -- sum_b Mueller(d_b) rather than
// Mueller(sum_b d_b) -- so beams do not interfere.  Its M11 element, the sum
// of |d_b|^2 / 2, is what the reference's theta probe accumulates
// (theta_m11_kernel, src/cuda/GpuDiffraction.cu:1909-1957).  Orientation
// items as in diffract_kernel's orientation mode: each item writes the
// weighted sum of its orientations, which reduce_mueller_kernel adds.
__global__ __launch_bounds__(256) MBS_OCCUPANCY void diffract_incoherent_kernel(
    const WorkItem *__restrict__ items, int n_items, int n_tiles, int cells,
    const double *__restrict__ tab, const double *__restrict__ hdr,
    const double *__restrict__ vtx, const double *__restrict__ vdd,
    const int32_t *__restrict__ vflag, const int32_t *__restrict__ nv,
    const int64_t *__restrict__ voff, const int64_t *__restrict__ ooff,
    const double *__restrict__ weights, double k, double eps1, double *__restrict__ out)
{
    int item, tile;
    if (!map_block(n_items, n_tiles, item, tile))
        return;
    const int cell = tile * 256 + threadIdx.x;
    const bool live = cell < cells;
    const int c = live ? cell : cells - 1;  // dead lanes compute, never store
    double d[3];
    for (int q = 0; q < 3; ++q)
        d[q] = tab[(size_t)(DT_DIR + q) * cells + c];
    __shared__ double s_u[12 * 256];
    __shared__ double2 s_tab[kTabN];
    for (int q = 0; q < 12; ++q)
        s_u[q * 256 + threadIdx.x] = tab[(size_t)(DT_U0 + q) * cells + c];
    {
        double ts, tc;
        sincospi((double)threadIdx.x / (kTabN / 2), &ts, &tc);
        s_tab[threadIdx.x] = make_double2(tc, ts);
    }
    __syncthreads();
    const WorkItem it = items[item];
    for (int o = it.o_begin; o < it.o_end; ++o) {
        double m[16];
        for (int q = 0; q < 16; ++q)
            m[q] = 0.0;
        for (int64_t b = ooff[o]; b < ooff[o + 1]; ++b) {
            const int64_t v0 = voff[b];
            double dj[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            beam_dir_f64<0>(hdr + (size_t)b * HDR_STRIDE, vtx + (size_t)v0 * VTX_STRIDE,
                            vdd + (size_t)v0 * VDD_STRIDE, vflag + v0, nv[b], d, s_u, s_tab, k,
                            eps1, dj, nullptr);
            add_mueller(m, dj, 1.0);
        }
        if (!live)
            continue;
        const double w = weights[o];
        double *dst = out + ((size_t)item * cells + cell) * 16;
        const bool first = (o == it.o_begin);
        for (int q = 0; q < 16; ++q)
            dst[q] = (first ? 0.0 : dst[q]) + w * m[q];
    }
}

// ---- coherent sum as a matrix contraction on MFMA (mixed precision) -------
//
// The rotation is linear in the beam's polarisation data: with the
// direction-only vectors u0, u1 (6 components each) and the per-beam complex
// W[m][c] = T0[m] Jp[0][c] + T1[m] Jp[1][c],
//   J[i][c](dir) = sum_b F_b(dir) (R_b Jp_b)[i][c]
//                = sum_m u_i[m](dir) G[m][c](dir),  G = sum_b F_b(dir) W_b[m][c].
// G is a complex [64 directions x beams] . [beams x 12] product per wave,
// done as a real one with K = 2 per beam (Re F, Im F) and N = 24 columns on
// v_mfma_f32_32x32x2f32 (two 32-direction blocks per wave).  The vector ALU
// keeps only the aperture integral; the rotation (4 six-term dot products),
// R.Jp and the accumulation (56 operations per beam and direction) move to
// the matrix cores.  The 24 columns pad to 32; the unpadded form on
// v_mfma_f32_4x4x1f32 (16 blocks of 4 directions, 12 instructions and 48
// bytes of operands per lane and beam) measured 30.0 ms against 28.4: its
// operand registers cost the occupancy more than the matrix-core cycles it
// saves.
typedef float f32x16 __attribute__((ext_vector_type(16)));
enum { G_LD = 25 };  // LDS row stride (floats) of the per-wave G transpose

// G (this wave's accumulators) -> this lane's Jones sum.  Accumulator v of
// lane l holds G[dir][n] with n = l % 32, dir = 8 (v/4) + 4 (l/32) + v % 4
// (+32 for the second block).  The two 32-direction blocks pass through one
// 32-row LDS tile in turn (half the LDS of staging both at once, which lets
// five workgroups share a CU).  Called by every thread of the block.
__device__ __forceinline__ void mfma_epilogue(const f32x16 &g0, const f32x16 &g1,
                                              float *__restrict__ sg,
                                              const float *__restrict__ u, int cells,
                                              float *j)
{
    const int lane = threadIdx.x & 63;
    const int n = lane & 31;
    for (int blk = 0; blk < 2; ++blk) {
        const f32x16 &g = blk ? g1 : g0;
        if (n < WB_COLS) {
            for (int v = 0; v < 16; ++v) {
                const int i = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                sg[i * G_LD + n] = g[v];
            }
        }
        __syncthreads();
        if ((lane >> 5) == blk) {
            const float *gr = sg + (lane & 31) * G_LD;
            for (int t = 0; t < 4; ++t) {  // t = 2 i + c
                const int i = t >> 1, c = t & 1;
                float re = 0.0f, im = 0.0f;
                for (int m = 0; m < 6; ++m) {
                    const float um = u[6 * i + m];
                    re = __builtin_fmaf(um, gr[4 * m + 2 * c], re);
                    im = __builtin_fmaf(um, gr[4 * m + 2 * c + 1], im);
                }
                j[2 * t] = re;
                j[2 * t + 1] = im;
            }
        }
        __syncthreads();
    }
}

// FSRC: beams with sidx[b] >= 0 take F from fbuf (the near-forward spectrum
// source) instead of the aperture integral.  WAVES: the occupancy the
// register allocation is held to (see MBS_MFMA_OCCUPANCY).
template <int NS, int JONES, int PER, int QUEUE, int FSRC = 0, int WAVES = MBS_MFMA_WAVES>
__global__ __launch_bounds__(256) MBS_MFMA_OCCUPANCY void diffract_mfma_kernel(
    const WorkItem *__restrict__ items, int n_items, int n_tiles, int cells, int n_az,
    const double *__restrict__ tab, const double *__restrict__ hdr,
    const float *__restrict__ wb, const double *__restrict__ vtx,
    const double *__restrict__ vdd, const int32_t *__restrict__ vflag,
    const int32_t *__restrict__ nv, const int64_t *__restrict__ voff,
    const int32_t *__restrict__ ext, const int64_t *__restrict__ ooff,
    const double *__restrict__ weights, double k, double eps1, double *__restrict__ out,
    double *__restrict__ out_ns, double *__restrict__ per_orient, int *__restrict__ queue,
    const float *__restrict__ tab32, const int32_t *__restrict__ sidx = nullptr,
    const float2 *__restrict__ fbuf = nullptr)
{
    __shared__ float4 s_tab[kTabN32];
    __shared__ float s_g[4 * 32 * G_LD];
    __shared__ int s_unit;
    int item = 0, tile = 0;
    if constexpr (!QUEUE) {
        if (!map_block(n_items, n_tiles, item, tile))
            return;
    }
    for (int i = threadIdx.x; i < kTabN32; i += blockDim.x)
        s_tab[i] = tab32_entry(i);
    QueueCursor qc;
    for (;;) {
    if constexpr (QUEUE) {
        if (!next_unit(queue, n_tiles, qc, s_unit, item, tile))
            return;
    }
    const int cell = tile * 256 + threadIdx.x;
    const bool live = cell < cells;
    const int c = live ? cell : cells - 1;  // dead lanes compute, never store
    double d[3];
    for (int q = 0; q < 3; ++q)
        d[q] = tab[(size_t)(DT_DIR + q) * cells + c];
    // this lane's u0, u1, read by the epilogue straight from its 48-byte
    // FP32 record (once per work item; staging them in LDS cost 12 KB per
    // workgroup, and one base address keeps the loads from holding twelve)
    const float *u = tab32 + (size_t)c * DT32_COUNT;
    __syncthreads();  // the phasor table
    float *sg = s_g + (threadIdx.x >> 6) * 32 * G_LD;
    const int lane = threadIdx.x & 63;
    const WorkItem it = items[item];
    for (int o = it.o_begin; o < it.o_end; ++o) {
        int64_t b0 = ooff[o], b1 = ooff[o + 1];
        if constexpr (JONES) {
            b0 = it.b_begin;
            b1 = it.b_end;
        }
        f32x16 g0 = {}, g1 = {}, n0 = {}, n1 = {};
        // the next beam's vertex offset and count are loaded one beam
        // ahead, so a beam's header and vertex loads issue together
        int64_t v0_next = b0 < b1 ? voff[b0] : 0;
        int nv_next = b0 < b1 ? nv[b0] : 0;
        // a 32-bit trip count: the scalar unit has no 64-bit ordered compare,
        // so a 64-bit loop test would go to the vector unit
        const int nbeam = (int)(b1 - b0);
        // B operand rows through a buffer resource based at the range's first
        // beam: the beam's offset is a scalar register and the lane's a loop-
        // invariant vector register, so no per-lane 64-bit address is formed
        // or kept (a range holds far fewer than 2^24 beams; reads past it
        // return zero)
        const __amdgpu_buffer_rsrc_t wres = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(wb + (size_t)b0 * WB_STRIDE), 0,
            nbeam * (int)(WB_STRIDE * sizeof(float)), kBufferRsrcWord3);
        const int lane_off = lane * (int)sizeof(float);
        for (int ib = 0; ib < nbeam; ++ib) {
            const int64_t b = b0 + ib;
            const float bop = __builtin_bit_cast(
                float, __builtin_amdgcn_raw_buffer_load_b32(
                           wres, lane_off, ib * (int)(WB_STRIDE * sizeof(float)), 0));
            const int64_t v0 = v0_next;
            const int nvb = nv_next;
            if (ib + 1 < nbeam) {
                v0_next = voff[b + 1];
                nv_next = nv[b + 1];
            }
            float fr, fi;
            const int slot = FSRC ? __builtin_amdgcn_readfirstlane(sidx[b]) : -1;
            if (FSRC && slot >= 0) {
                const float2 f = fbuf[(size_t)slot * cells + c];
                fr = f.x;
                fi = f.y;
            } else {
                aperture_mixed(hdr + (size_t)b * HDR_STRIDE, vtx + (size_t)v0 * VTX_STRIDE,
                               vdd + (size_t)v0 * VDD_STRIDE, vflag + v0, nvb, d, s_tab, k,
                               eps1, fr, fi);
            }
            // A operands: lanes 0-31 (Re F), lanes 32-63 (Im F) of the
            // wave's directions 0-31, then the same for directions 32-63
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(fr),
                                                             __float_as_uint(fi), false, false);
            const float a0 = __uint_as_float(sw[0]), a1 = __uint_as_float(sw[1]);
            g0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bop, g0, 0, 0, 0);
            g1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bop, g1, 0, 0, 0);
            if (NS && ext[b] == 0) {
                n0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bop, n0, 0, 0, 0);
                n1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bop, n1, 0, 0, 0);
            }
        }
        float j[8], jn[8];
        mfma_epilogue(g0, g1, sg, u, cells, j);
        if constexpr (NS)
            mfma_epilogue(n0, n1, sg, u, cells, jn);
        if (!live)
            continue;
        if constexpr (JONES) {
            // FP32 Jones partials (the sums are FP32)
            float *dst = reinterpret_cast<float *>(out) + ((size_t)item * cells + cell) * 8;
            for (int q = 0; q < 8; ++q)
                dst[q] = j[q];
            if constexpr (NS) {
                float *dn = reinterpret_cast<float *>(out_ns) + ((size_t)item * cells + cell) * 8;
                for (int q = 0; q < 8; ++q)
                    dn[q] = jn[q];
            }
        } else {
            const double w = weights[o];
            const bool first = (o == it.o_begin);
            double *per = PER ? per_orient + ((size_t)o * cells + abi_cell(cell, n_az, cells)) * 16
                              : nullptr;
            flush_mueller(j, w, out + ((size_t)item * cells + cell) * 16, first, per);
            if constexpr (NS)
                flush_mueller(jn, w, out_ns + ((size_t)item * cells + cell) * 16, first,
                              (double *)nullptr);
        }
    }
    if constexpr (!QUEUE)
        return;
    __syncthreads();  // the next unit overwrites s_g
    }
}

// Zero fill (used instead of a memset so that no memset node is ever recorded
// into a captured graph).
__global__ void zero_kernel(double *__restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0.0;
}

// Sum of per-item Mueller partials in item order: res[c] = sum_i part[i][c].
__global__ void reduce_mueller_kernel(const double *__restrict__ part, int n_items,
                                      int cells, int n_az, double *__restrict__ res)
{
    // (Mueller-mode items are always planned on the host)
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)cells * 16;
    if (idx >= n)
        return;
    double s = 0.0;
    for (int i = 0; i < n_items; ++i)
        s += part[(size_t)i * n + idx];
    res[abi_cell((int)(idx / 16), n_az, cells) * 16 + idx % 16] = s;
}

// Jones segments -> Mueller: items are ordered by orientation; for each
// orientation the segment Jones sums are added in segment order, converted
// and accumulated with the orientation weight (TracerPOTotal.cpp:5985-5994).
template <typename T>
__global__ void jones_to_mueller_kernel(const T *__restrict__ part,
                                        const WorkItem *__restrict__ items,
                                        int n_items, int cells, int n_az,
                                        const double *__restrict__ weights,
                                        double *__restrict__ res,
                                        double *__restrict__ per_orient,
                                        const int *__restrict__ queue)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells)
        return;
    if (queue)  // planned on the device
        n_items = queue[Q_ITEMS];
    const size_t out_cell = abi_cell(cell, n_az, cells);
    double m[16];
    for (int q = 0; q < 16; ++q)
        m[q] = 0.0;
    int i = 0;
    while (i < n_items) {
        const int o = items[i].o_begin;
        double j[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        while (i < n_items && items[i].o_begin == o) {
            const T *src = part + ((size_t)i * cells + cell) * 8;
            for (int q = 0; q < 8; ++q)
                j[q] += src[q];
            ++i;
        }
        add_mueller(m, j, weights[o]);
        if (per_orient) {
            double po[16];
            for (int q = 0; q < 16; ++q)
                po[q] = 0.0;
            add_mueller(po, j, weights[o]);
            double *dst = per_orient + ((size_t)o * cells + out_cell) * 16;
            for (int q = 0; q < 16; ++q)
                dst[q] = po[q];
        }
    }
    double *dst = res + out_cell * 16;
    for (int q = 0; q < 16; ++q)
        dst[q] = m[q];
}

// Azimuth (laboratory-alpha) average of an orientation-summed Mueller grid,
// one row per zenith: HandlerPOTotal::WriteMatricesToFile
// (src/handler/HandlerPOTotal.cpp:23-122).  Row t = [theta in degrees,
// 2 pi dcos, M11..M44] with Mbar = (1/nAz) sum_p M(p,t) L(-phi_p), where
// L(a) rotates the Stokes Q/U block by 2a (at the poles the plain mean is
// taken); then the pole forms of PoleMueller (src/handler/PoleMueller.h:25-55)
// and the solid-angle weight of ScatteringRange::Compute2PiDcos
// (src/handler/Handler.h:70-125) in the stable sin*sin form.  One wave per
// row; the 64 lanes stride over azimuths and combine by a fixed butterfly, so
// the result is deterministic.
// blockIdx.y selects one of several grids stored back to back (the
// per-orientation grids of an MBS_PER_ORIENT run).
__global__ void __launch_bounds__(64) azimuth_average_kernel(
    const double *__restrict__ m, const double *__restrict__ theta, int n_az, int n_zen,
    double az_step, double *__restrict__ rows)
{
    const int t = blockIdx.x;
    const int nZ = n_zen + 1;
    m += (size_t)blockIdx.y * n_az * nZ * 16;
    rows += (size_t)blockIdx.y * nZ * 18;
    const double th = theta[t];
    const double tol = 256.0 * 2.220446049250313e-16;
    const bool fwd = __builtin_fabs(th) <= tol;
    const bool bwd = __builtin_fabs(th - kPi) <= tol;
    double s[16];
    for (int q = 0; q < 16; ++q)
        s[q] = 0.0;
    for (int p = threadIdx.x; p < n_az; p += 64) {
        const double *c = m + ((size_t)p * nZ + t) * 16;
        double sn, cs;
        sincos(2.0 * (-(double)p * az_step), &sn, &cs);
        for (int r = 0; r < 4; ++r) {
            const double *row = c + 4 * r;
            if (fwd || bwd) {
                for (int q = 0; q < 4; ++q)
                    s[4 * r + q] += row[q];
            } else {
                s[4 * r + 0] += row[0];
                s[4 * r + 1] += row[1] * cs - row[2] * sn;
                s[4 * r + 2] += row[1] * sn + row[2] * cs;
                s[4 * r + 3] += row[3];
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 16; ++q)
            s[q] += __shfl_xor(s[q], off, 64);
    if (threadIdx.x != 0)
        return;
    for (int q = 0; q < 16; ++q)
        s[q] /= (double)n_az;
    if (fwd || bwd) {
        const double m00 = s[0], m33 = s[15], m03 = s[3];
        const double m11 = fwd ? 0.5 * (s[5] + s[10]) : 0.5 * (s[5] - s[10]);
        for (int q = 0; q < 16; ++q)
            s[q] = 0.0;
        s[0] = m00;
        s[5] = m11;
        s[10] = fwd ? m11 : -m11;
        s[15] = m33;
        s[3] = m03;
        s[12] = m03;
    }
    double lo, hi;
    if (t == 0) {
        lo = theta[0];
        hi = 0.5 * (theta[0] + theta[n_zen > 0 ? 1 : 0]);
    } else if (t == n_zen) {
        lo = 0.5 * (theta[n_zen - 1] + theta[n_zen]);
        hi = theta[n_zen];
    } else {
        lo = 0.5 * (theta[t - 1] + theta[t]);
        hi = 0.5 * (theta[t] + theta[t + 1]);
    }
    double *o = rows + (size_t)t * 18;
    o[0] = (th * 180.0) / kPi;  // RadToDeg, src/common/common.cpp:140-143
    o[1] = 2.0 * sin(0.5 * (lo + hi)) * sin(0.5 * (hi - lo)) * (2.0 * kPi);
    for (int q = 0; q < 16; ++q)
        o[2 + q] = s[q];
}

}  // namespace mbs
