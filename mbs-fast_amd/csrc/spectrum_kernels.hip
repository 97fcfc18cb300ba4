// spectrum_kernels.hip -- the aperture integral of selected beams from their
// FFT spectra (the near-forward F source).
//
// For a beam whose scattering directions all lie close to its own direction
// (|A|, |B| small over the grid), F(A, B) is a low-frequency sample of the
// aperture's Fourier transform.  Each such beam gets one spectrum:
//
//   1. raster: the polygon on an N x N grid of pixel centres (spacing h), each
//      value the polygon integral of the separable tent t(u) t(v),
//      t(u) = max(0, 1 - |u|/h), centred on the pixel, over h^2.  That is the
//      polygon convolved with a tent of width 2h and sampled; its transform
//      is the polygon's times sinc^2(f_x h) sinc^2(f_y h), and the
//      aliased copies are suppressed by the same factor at their frequency
//      (|f h - m| -> (f h)^2 / m^2), where a 0/1 raster leaves them whole.
//      Pixels whose tent support lies inside every edge line are 1, outside
//      one edge line 0; the others (boundary pixels) are integrated exactly
//      by an edge form of Green's theorem (spec_tent_exact), one per lane
//      from a per-row queue;
//   2. deapodisation: each raster value divided by the transform of the
//      interpolation kernel at its position (below);
//   3. the 2-D FFT (the Stockham stages of the aperture FFT,
//      fft_kernels.hip), N = 2 x the padded extent (oversampling 2), pruned
//      to the band of bins the grid samples: with every |A|, |B| <= a_max
//      the taps stay within M = ceil(k a_max N h / 2 pi + W/2) + 1 bins of
//      zero, so the row pass stores only those P = 2M + 1 columns of each
//      row's transform and the column pass transforms only those columns
//      and stores P rows (P / N of the column work and traffic, (P / N)^2 of
//      the footprint).  Rows whose tent support misses the polygon are zero:
//      neither transformed nor stored, and read as zero by the column pass.
//      Each stored bin (mx, my) is multiplied by the phases
//      e^{i pi (N - 1) m / N} that centre the pixel grid, once,
//      instead of at each of the W^2 taps of every sample;
//   4. sampling at the grid's cells (spectrum_sample_kernel): the spectrum
//      interpolated with the "exponential of semicircle" kernel
//      phi(z) = exp(beta (sqrt(1 - z^2) - 1)) over W bins per axis
//      (Barnett, Magland, af Klinteberg, "A parallel non-uniform fast Fourier
//      transform library based on an exponential of semicircle kernel",
//      SIAM J. Sci. Comput. 41 (2019)), the tent's sinc^2 divided out, and
//      the phase of the raster's centre and of the beam's centre applied.
//
// The result is F in the diffraction kernels' convention:
//   F = s k^2 I(A, B) e^{i psi},   I(A, B) = integral over the polygon of
//   e^{i k (A x + B y)} dx dy,
// with s the sign of the polygon's signed area (the edge sum of the direct
// kernels, HandlerPO_fast.h:318-372, carries the orientation of the vertex
// list) and psi = -k d.centre (zero for external beams).  It is written to a
// buffer [spectrum beam][cell] (float2) that the spectrum variant of the
// MFMA kernel reads in place of its own aperture integral for those beams.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_kernels.hip"
#include "po_layout.h"
#include "po_math.h"

namespace mbs {

// Interpolation kernel: W taps per axis, beta = 2.30 W (oversampling 2).
constexpr int kSpecW = 6;
constexpr double kSpecBeta = 2.30 * kSpecW;
// Polygons up to this many vertices take the spectrum path.
constexpr int kSpecMaxNv = 16;
// Pixel size at most the polygon's extent over kSpecPixelsAcross and its
// least width over kSpecPixelsWidth.  A polygon not much larger than a pixel
// in some direction has a broad spectrum in that direction, and its aliases
// are then damped only by the tent's (f h)^2: the alias error is about
// area (h / width) (f h)^2 / 2.
constexpr double kSpecPixelsAcross = 32.0;
constexpr double kSpecPixelsWidth = 16.0;

// Per-beam record: raster centre, pixel size, orientation sign, vertex
// count, vertices, edge lines a x + b y + c (> 0 inside) and the tent test
// radius (|a| + |b|) h of each edge.
enum {
    SP_XC = 0,
    SP_YC = 1,
    SP_H = 2,
    SP_SG = 3,
    SP_NV = 4,
    SP_VX = 5,
    SP_VY = SP_VX + kSpecMaxNv,
    SP_EDGE = SP_VY + kSpecMaxNv,  // (a, b, c, r) per edge
    SP_M = SP_EDGE + 4 * kSpecMaxNv,  // band half-width (bins)
    SP_P = SP_M + 1,                   // band width 2M + 1, or N unpruned
    SP_YLO = SP_P + 1,                 // the polygon's y range
    SP_YHI = SP_YLO + 1,
    SP_STRIDE = SP_YHI + 1
};

// Band column of signed bin m (|m| <= M, or any m in [-N/2, N/2) when
// P = N): non-negative bins first, then the negative ones.
__device__ __forceinline__ int spec_band_col(int m, int p) { return m >= 0 ? m : m + p; }

// What the selection needs of every beam, one workgroup per beam:
// max over the grid's cells of max(|A|, |B|), the polygon's larger bounding
// box side and its least width (the least over edges of the largest vertex
// distance from the edge's line; both times the packed size scale), whether
// it is convex, and the sign of its signed area.  out: [n_beams][5].
__global__ void __launch_bounds__(256) spectrum_scan_kernel(
    int n_beams, const double *__restrict__ hdr, const double *__restrict__ tab, int cells,
    const int32_t *__restrict__ nv_in, const int64_t *__restrict__ voff,
    const double *__restrict__ vx, const double *__restrict__ vy, double scale,
    double *__restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= n_beams)
        return;
    const double *h = hdr + (size_t)b * HDR_STRIDE;
    double amax = 0.0;
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const double d0 = tab[(size_t)(DT_DIR + 0) * cells + c];
        const double d1 = tab[(size_t)(DT_DIR + 1) * cells + c];
        const double d2 = tab[(size_t)(DT_DIR + 2) * cells + c];
        const double A = h[H_A0] - (d0 * h[H_HOR] + d1 * h[H_HOR + 1] + d2 * h[H_HOR + 2]);
        const double B = h[H_B0] - (d0 * h[H_VER] + d1 * h[H_VER + 1] + d2 * h[H_VER + 2]);
        amax = fmax(amax, fmax(fabs(A), fabs(B)));
    }
    __shared__ double red[256];
    red[threadIdx.x] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x != 0)
        return;
    const int nv = nv_in[b];
    const int64_t v0 = voff[b];
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300, area2 = 0.0;
    int pos = 0, neg = 0;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0, g = f + 1 < nv ? f + 1 : 0;
        const double xe = vx[v0 + e], ye = vy[v0 + e];
        x0 = fmin(x0, xe);
        x1 = fmax(x1, xe);
        y0 = fmin(y0, ye);
        y1 = fmax(y1, ye);
        area2 += xe * vy[v0 + f] - vx[v0 + f] * ye;
        const double cr = (vx[v0 + f] - xe) * (vy[v0 + g] - vy[v0 + f]) -
                          (vy[v0 + f] - ye) * (vx[v0 + g] - vx[v0 + f]);
        pos += cr > 0.0;
        neg += cr < 0.0;
    }
    double width = 1e300;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double ex = vx[v0 + f] - vx[v0 + e], ey = vy[v0 + f] - vy[v0 + e];
        const double len = sqrt(ex * ex + ey * ey);
        if (len <= 0.0)
            continue;
        double far = 0.0;
        for (int q = 0; q < nv; ++q)
            far = fmax(far, fabs(ex * (vy[v0 + q] - vy[v0 + e]) - ey * (vx[v0 + q] - vx[v0 + e])));
        width = fmin(width, far / len);
    }
    double *o = out + (size_t)b * 5;
    o[0] = red[0];
    o[1] = scale * fmax(x1 - x0, y1 - y0);
    o[2] = (pos == 0 || neg == 0) ? 1.0 : 0.0;
    o[3] = area2 >= 0.0 ? 1.0 : -1.0;
    o[4] = scale * width;
}

// Records of the spectrum beams.  beam[i], hpx[i] (pixel size) and n (the
// group's grid size) come from the host's selection.
__global__ void spectrum_setup_kernel(int n_sb, const int32_t *__restrict__ beam,
                                      const double *__restrict__ hpx,
                                      const int32_t *__restrict__ logn,
                                      const int32_t *__restrict__ band,
                                      const int32_t *__restrict__ nv_in,
                                      const int64_t *__restrict__ voff,
                                      const double *__restrict__ vx,
                                      const double *__restrict__ vy, double scale,
                                      double *__restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sb)
        return;
    const int b = beam[i];
    const int nv = nv_in[b];
    const int64_t v0 = voff[b];
    // vertices at the packed size (the pack scales them the same way)
    auto X = [&](int e) { return scale * vx[v0 + e]; };
    auto Y = [&](int e) { return scale * vy[v0 + e]; };
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300, area2 = 0.0;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        x0 = fmin(x0, X(e));
        x1 = fmax(x1, X(e));
        y0 = fmin(y0, Y(e));
        y1 = fmax(y1, Y(e));
        area2 += X(e) * Y(f) - X(f) * Y(e);
    }
    double *r = rec + (size_t)i * SP_STRIDE;
    const double h = hpx[i];
    const double sg = area2 >= 0.0 ? 1.0 : -1.0;
    r[SP_XC] = 0.5 * (x0 + x1);
    r[SP_YC] = 0.5 * (y0 + y1);
    r[SP_H] = h;
    r[SP_SG] = sg;
    r[SP_NV] = nv;
    const int n = 1 << logn[i];
    const int m = band[i];
    r[SP_M] = 2 * m + 1 < n ? m : n / 2;
    r[SP_P] = 2 * m + 1 < n ? 2 * m + 1 : n;
    r[SP_YLO] = y0;
    r[SP_YHI] = y1;
    // vertices counter-clockwise (reversed when the list runs clockwise)
    for (int e = 0; e < nv; ++e) {
        const int s = sg > 0.0 ? e : nv - 1 - e;
        r[SP_VX + e] = X(s);
        r[SP_VY + e] = Y(s);
    }
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double xe = r[SP_VX + e], ye = r[SP_VY + e];
        const double a = -(r[SP_VY + f] - ye), bb = r[SP_VX + f] - xe;
        double *l = r + SP_EDGE + 4 * e;
        l[0] = a;
        l[1] = bb;
        l[2] = -__builtin_fma(a, xe, bb * ye);
        l[3] = (fabs(a) + fabs(bb)) * h;
    }
}

// Transform of the interpolation kernel, phi_hat(s) = integral over
// |u| <= W/2 (bins) of phi(2u/W) cos(2 pi u s) du, at the position
// s = (j - N/2 + 1/2) / N of raster column j; deap[j] = 1 / phi_hat.  Midpoint
// rule on 1024 nodes over [0, W/2] (the integrand is even): one wave per
// column, 16 nodes per lane, then a fixed-order butterfly sum.
__global__ void __launch_bounds__(256) spectrum_deap_kernel(int n, double *__restrict__ deap)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n)
        return;  // (whole waves)
    const double s = ((double)j - 0.5 * n + 0.5) / n;
    const int q = 1024;
    const double du = 0.5 * kSpecW / q;
    double acc = 0.0;
    for (int i = lane; i < q; i += 64) {
        const double u = (i + 0.5) * du;
        const double z = 2.0 * u / kSpecW;
        acc += exp(kSpecBeta * (sqrt(1.0 - z * z) - 1.0)) * cospi(2.0 * u * s);
    }
    for (int w = 32; w > 0; w >>= 1)
        acc += __shfl_xor(acc, w, 64);
    if (lane == 0)
        deap[j] = 1.0 / (2.0 * du * acc);
}

// Tent-weighted integral of the polygon over the support of the pixel at
// (px, py), over h^2 (1 for a pixel whose support is inside).  By Green's
// theorem, with the cumulative tent T(u) = integral of t up to u,
//   integral over P of t(x - px) t(y - py) = sum over the (counter-
//   clockwise) edges of integral T(x(y) - px) t(y - py) dy,
// and along an edge the integrand is a cubic in y between the kinks of t
// (y = py; the band |y - py| <= h bounds the integral) and of T (where
// x(y) - px = -h, 0, h), so Simpson's rule on the pieces is exact.  Edges
// that miss the band contribute nothing; registers only.
__device__ __forceinline__ double spec_tent_T(double u, double h)
{
    if (u <= -h)
        return 0.0;
    if (u <= 0.0)
        return (u + h) * (u + h) / (2.0 * h);
    if (u <= h)
        return h - (h - u) * (h - u) / (2.0 * h);
    return h;
}

__device__ double spec_tent_exact(const double *__restrict__ r, double px, double py, double h)
{
    const int nv = (int)r[SP_NV];
    const double ylo = py - h, yhi = py + h;
    double tot = 0.0;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double x0 = r[SP_VX + e], y0 = r[SP_VY + e];
        const double x1 = r[SP_VX + f], y1 = r[SP_VY + f];
        if (y0 == y1)
            continue;
        const double ya = fmax(fmin(y0, y1), ylo), yb = fmin(fmax(y0, y1), yhi);
        if (!(ya < yb))
            continue;
        const double m = (x1 - x0) / (y1 - y0);  // dx/dy
        // the kinks, clamped into [ya, yb] (outside ones give empty pieces)
        double c[4];
        c[0] = py;
        if (x1 != x0) {
            const double im = (y1 - y0) / (x1 - x0);
            c[1] = __builtin_fma(px - h - x0, im, y0);
            c[2] = __builtin_fma(px - x0, im, y0);
            c[3] = __builtin_fma(px + h - x0, im, y0);
        } else {
            c[1] = c[2] = c[3] = ya;
        }
        for (int q = 0; q < 4; ++q)
            c[q] = fmin(fmax(c[q], ya), yb);
        // sorting network on four values
        auto cx = [](double &a, double &b) {
            const double lo = fmin(a, b), hi = fmax(a, b);
            a = lo;
            b = hi;
        };
        cx(c[0], c[1]);
        cx(c[2], c[3]);
        cx(c[0], c[2]);
        cx(c[1], c[3]);
        cx(c[1], c[2]);
        auto fy = [&](double y) {
            return spec_tent_T(__builtin_fma(y - y0, m, x0) - px, h) *
                   fmax(0.0, 1.0 - fabs(y - py) / h);
        };
        const double pts[6] = {ya, c[0], c[1], c[2], c[3], yb};
        double fa = fy(pts[0]), s = 0.0;
        for (int q = 0; q < 5; ++q) {
            const double a = pts[q], b = pts[q + 1];
            const double fb = fy(b);
            s += (b - a) * (fa + 4.0 * fy(0.5 * (a + b)) + fb);
            fa = fb;
        }
        tot += (y1 > y0 ? s : -s) * (1.0 / 6.0);
    }
    return tot / (h * h);
}

// Class of the pixel at (px, py): 0 where the tent's support lies outside
// an edge line, 1 where it lies inside every edge line, 2 otherwise (the
// value is then the exact tent integral).
__device__ __forceinline__ int spec_pixel_class(const double *__restrict__ r, double px,
                                                double py)
{
    const int nv = (int)r[SP_NV];
    bool inside = true;
    for (int e = 0; e < nv; ++e) {
        const double *l = r + SP_EDGE + 4 * e;
        const double v = __builtin_fma(l[0], px, __builtin_fma(l[1], py, l[2]));
        if (v <= -l[3])
            return 0;  // the whole support is outside this edge's line
        inside = inside && v > l[3];
    }
    return inside ? 1 : 2;
}

// e^{i pi (n - 1) m / n} for signed bins m in [-n/2, n/2] (cph[m + n/2]),
// the phase that centres bin m of an n-point transform of pixel-centred
// samples, with the angle reduced exactly in integers: (n - 1) m mod 2n.
__global__ void spectrum_phase_kernel(int n, float2 *__restrict__ cph)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n)
        return;
    const long long m = q - n / 2;
    const long long a = (((long long)(n - 1) * m) % (2LL * n) + 2LL * n) % (2LL * n);
    double sn, cs;
    sincospi((double)a / n, &sn, &cs);
    cph[q] = make_float2((float)cs, (float)sn);
}

// Row whose tent support [py - h, py + h] misses the polygon's y range: all
// zero.  Such rows are neither transformed nor stored; the column pass reads
// them as zero.
__device__ __forceinline__ bool spec_zero_row(const double *__restrict__ r, int row, int n)
{
    const double py = fft_pixel(row, n, r[SP_H], r[SP_YC]);
    return py + r[SP_H] <= r[SP_YLO] || py - r[SP_H] >= r[SP_YHI];
}

// Raster + deapodisation + row FFTs of the spectrum beams of one group (all
// of grid size 2^LOGN), as fft_rows_kernel with the tent raster; each row
// stores its band of P columns ([N rows][P], bin mx at spec_band_col),
// times the centring phase of mx.  The raster goes to LDS first: pixels
// inside or outside are written at once, the boundary pixels (a few per
// row, many in a row along a near-horizontal edge) are queued and their
// exact tent integrals taken one per lane (evaluated where they fall, most
// lanes of a wave would idle through each one); the first FFT stage then
// reads LDS.  LDS: FFT_ROWS x fft_ld(N) float2 and FFT_ROWS x N queue slots.
template <int LOGN>
__global__ void __launch_bounds__(FFT_ROW_THREADS) spectrum_rows_kernel(
    int n_sb, const double *__restrict__ rec, const double *__restrict__ deap,
    const float2 *__restrict__ tw, const float2 *__restrict__ cph,
    const int64_t *__restrict__ off, float2 *__restrict__ spec)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ float2 lds[];
    const int units_per_beam = N / FFT_ROWS;
    const int unit = fft_logical_block();
    if (unit >= n_sb * units_per_beam)
        return;
    const int beam = unit / units_per_beam;
    const int row0 = (unit - beam * units_per_beam) * FFT_ROWS;
    const double *r = rec + (size_t)beam * SP_STRIDE;
    const double xc = r[SP_XC], yc = r[SP_YC], h = r[SP_H];
    const int M = (int)r[SP_M], P = (int)r[SP_P];
    float2 *g = spec + off[beam] + (size_t)row0 * P;
    bool all_zero = true;
    for (int rr = 0; rr < FFT_ROWS; ++rr)
        all_zero = all_zero && spec_zero_row(r, row0 + rr, N);
    if (all_zero)
        return;
    constexpr int LD = fft_ld(N);
    int *queue = reinterpret_cast<int *>(lds + FFT_ROWS * LD);
    __shared__ int s_count;
    if (threadIdx.x == 0)
        s_count = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < FFT_ROWS * N; q += blockDim.x) {
        const int seq = q / N, i = q - seq * N;
        const int cls =
            spec_pixel_class(r, fft_pixel(i, N, h, xc), fft_pixel(row0 + seq, N, h, yc));
        float v = 0.0f;
        if (cls == 1)
            v = (float)(1.0 * deap[row0 + seq] * deap[i]);
        else if (cls == 2)
            queue[atomicAdd(&s_count, 1)] = q;
        lds[seq * LD + fft_pad(i)] = make_float2(v, 0.0f);
    }
    __syncthreads();
    const int count = s_count;
    for (int t = threadIdx.x; t < count; t += blockDim.x) {
        const int q = queue[t];
        const int seq = q / N, i = q - seq * N;
        const double v =
            spec_tent_exact(r, fft_pixel(i, N, h, xc), fft_pixel(row0 + seq, N, h, yc), h);
        lds[seq * LD + fft_pad(i)] = make_float2((float)(v * deap[row0 + seq] * deap[i]), 0.0f);
    }
    __syncthreads();
    auto raster = [&](int seq, int i) { return lds[seq * LD + fft_pad(i)]; };
    auto out = [&](int seq, int i, float2 x) {
        const int m = i <= M ? i : (i >= N - M ? i - N : N);
        if (m != N)
            g[(size_t)seq * P + spec_band_col(m, P)] = cmulf(x, cph[m + N / 2]);
    };
    fft_stages<LOGN, FFT_ROWS, FFT_ROW_THREADS, false, 0>(lds, LD, tw, raster, out);
}

// Column FFTs of the band: work unit = (beam, FFT_COLS adjacent band
// columns); reads all N rows of its columns, stores rows my in the band
// (the first P rows, bin my at spec_band_col) times the centring phase of
// my.  In place: a unit owns its columns, and the first stage reads them
// whole before the last one writes.
template <int LOGN>
__global__ void __launch_bounds__(FFT_COL_THREADS) spectrum_cols_kernel(
    int n_sb, int units_per_beam, const double *__restrict__ rec,
    const float2 *__restrict__ tw, const float2 *__restrict__ cph,
    const int64_t *__restrict__ off, float2 *__restrict__ spec)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ float2 lds[];
    const int unit = fft_logical_block();
    if (unit >= n_sb * units_per_beam)
        return;
    const int beam = unit / units_per_beam;
    const int col0 = (unit - beam * units_per_beam) * FFT_COLS;
    const double *r = rec + (size_t)beam * SP_STRIDE;
    const int M = (int)r[SP_M], P = (int)r[SP_P];
    if (col0 >= P)
        return;
    float2 *g = spec + off[beam] + col0;
    const int ncol = P - col0;  // columns of this unit inside the band
    auto in = [&](int seq, int i) {
        return (seq < ncol && !spec_zero_row(r, i, N)) ? g[(size_t)i * P + seq]
                                                       : make_float2(0.0f, 0.0f);
    };
    auto out = [&](int seq, int i, float2 x) {
        const int m = i <= M ? i : (i >= N - M ? i - N : N);
        if (seq < ncol && m != N)
            g[(size_t)spec_band_col(m, P) * P + seq] = cmulf(x, cph[m + N / 2]);
    };
    fft_stages<LOGN, FFT_COLS, FFT_COL_THREADS, true, 0>(lds, fft_ld(N), tw, in, out);
}

// F of spectrum beam i at every cell: thread per (i, cell).  off[i] is the
// offset of beam i's band [P][P] (centring phases applied), logn[i] its grid
// size.  fbuf: [n_sb][cells].
__global__ void __launch_bounds__(256) spectrum_sample_kernel(
    int n_sb, int cells, const int32_t *__restrict__ beam, const int64_t *__restrict__ off,
    const int32_t *__restrict__ logn, const double *__restrict__ rec,
    const float2 *__restrict__ spectra, const double *__restrict__ hdr,
    const double *__restrict__ tab, double k, float2 *__restrict__ fbuf)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n_sb * cells)
        return;
    const int i = (int)(idx / cells);
    const int c = (int)(idx - (size_t)i * cells);
    const double *h = hdr + (size_t)beam[i] * HDR_STRIDE;
    const double *r = rec + (size_t)i * SP_STRIDE;
    const int n = 1 << logn[i];
    const int P = (int)r[SP_P];
    const double d0 = tab[(size_t)(DT_DIR + 0) * cells + c];
    const double d1 = tab[(size_t)(DT_DIR + 1) * cells + c];
    const double d2 = tab[(size_t)(DT_DIR + 2) * cells + c];
    const double A = h[H_A0] - (d0 * h[H_HOR] + d1 * h[H_HOR + 1] + d2 * h[H_HOR + 2]);
    const double B = h[H_B0] - (d0 * h[H_VER] + d1 * h[H_VER + 1] + d2 * h[H_VER + 2]);
    const double psi = d0 * h[H_KC] + d1 * h[H_KC + 1] + d2 * h[H_KC + 2];
    const double px_ = r[SP_H];
    const double inv2pi = 0.15915494309189533577;
    const double fx = -k * A * inv2pi, fy = -k * B * inv2pi;
    // positions in bins; the W taps nearest each (inside the band by the
    // choice of M)
    const double ux = fx * n * px_, uy = fy * n * px_;
    const int jx0 = (int)ceil(ux - 0.5 * kSpecW), jy0 = (int)ceil(uy - 0.5 * kSpecW);
    // kernel weights in FP32 (relative error ~1e-7, below the spectrum's)
    float wx[kSpecW], wy[kSpecW];
    for (int t = 0; t < kSpecW; ++t) {
        const float zx = (float)(2.0 * (ux - (jx0 + t)) / kSpecW);
        const float zy = (float)(2.0 * (uy - (jy0 + t)) / kSpecW);
        const float beta = (float)kSpecBeta;
        wx[t] = fabsf(zx) < 1.0f ? expf(beta * (sqrtf(1.0f - zx * zx) - 1.0f)) : 0.0f;
        wy[t] = fabsf(zy) < 1.0f ? expf(beta * (sqrtf(1.0f - zy * zy) - 1.0f)) : 0.0f;
    }
    const float2 *g = spectra + off[i];
    float sr = 0.0f, si = 0.0f;
    // (a tap outside the band cannot occur; it would read as zero)
    const int M = (int)r[SP_M];
    for (int ty = 0; ty < kSpecW; ++ty) {
        const int my = jy0 + ty;
        if (my < -M || my > M)
            continue;
        const float2 *row = g + (size_t)spec_band_col(my, P) * P;
        float rr = 0.0f, ri = 0.0f;
        for (int tx = 0; tx < kSpecW; ++tx) {
            const int mx = jx0 + tx;
            const float2 z = (mx < -M || mx > M) ? make_float2(0.0f, 0.0f)
                                                 : row[spec_band_col(mx, P)];
            rr = __builtin_fmaf(z.x, wx[tx], rr);
            ri = __builtin_fmaf(z.y, wx[tx], ri);
        }
        sr = __builtin_fmaf(rr, wy[ty], sr);
        si = __builtin_fmaf(ri, wy[ty], si);
    }
    // the tent's sinc^2 per axis, the pixel area, s k^2 and the phase of
    // the raster centre and of the beam centre
    const double sx = fx * px_, sy = fy * px_;
    const double snx = fabs(sx) > 1e-8 ? sinpi(sx) / (kPi * sx) : 1.0;
    const double sny = fabs(sy) > 1e-8 ? sinpi(sy) / (kPi * sy) : 1.0;
    const double scale = r[SP_SG] * k * k * px_ * px_ / (snx * snx * sny * sny);
    const double ph = psi - 2.0 * kPi * (fx * r[SP_XC] + fy * r[SP_YC]);
    double s, co;
    sincos(ph, &s, &co);
    const double vr = scale * ((double)sr * co - (double)si * s);
    const double vi = scale * ((double)sr * s + (double)si * co);
    fbuf[idx] = make_float2((float)vr, (float)vi);
}

}  // namespace mbs
