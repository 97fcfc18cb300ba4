// fft_kernels.hip -- aperture FFT of beams (near-forward path): the aperture
// polygon of each beam is sampled on an n x n grid of pixel centres, the grid
// is Fourier transformed in FP32 and the transform is read off at arbitrary
// (A, B) by phase-corrected Lanczos interpolation.  This is the method of the
// reference's aperture probe (tools/fft_aperture_probe.cu:130-169 raster,
// :549-551 2-D complex FFT, :50-102 Lanczos sampling), generalised from
// rectangles to the beams' convex polygons.
//
// Transform convention: I(A, B) = integral over the aperture of
// exp(i k (A x + B y)) dx dy = G(fx, fy) with G(f) = integral of
// exp(-2 pi i f.r) and f = -k (A, B) / (2 pi).  Pixel centres are
// r = c + ((2 j + 1 - n)/2) d per axis, so for bin m the continuous transform
// is approximated by d^2 X[m] exp(+i pi m (n - 1)/n) exp(-2 pi i f.c), where X
// is the DFT (exp(-2 pi i j m / n)) of the 0/1 raster.
//
// Layout: one grid per beam, row-major [n rows (y)][n columns (x)] of float2,
// the whole batch contiguous.  Two passes over HBM:
//   fft_rows_kernel   rasterises ROWS rows into LDS (no global read), does
//                     their n-point FFTs in LDS, writes them (n^2 x 8 B)
//   fft_cols_kernel   reads COLS adjacent columns (64-byte row segments),
//                     does their FFTs in LDS, writes them back (2 n^2 x 8 B)
// The 1-D FFT is radix-2 decimation in time, in place in LDS after a
// bit-reversed store, with two stages fused per pass (4 points per thread
// and pass, radix-4 style), twiddles from an LDS table made in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

// Workgroup sizes: the row kernel (2 workgroups of 74 KB LDS per CU) runs
// 512 threads, the column kernel (1 workgroup of 139 KB) 1024, so that each
// SIMD keeps 4 waves in flight either way.
enum { FFT_ROW_THREADS = 512, FFT_COL_THREADS = 1024 };
// LDS sequences are FFT_PAD elements longer than n, so that the 8 columns of
// the column kernel start 16 banks apart.
enum { FFT_PAD = 8 };
enum { FB_XC = 0, FB_YC = 1, FB_DX = 2, FB_DY = 3, FB_NE = 4, FB_EDGE = 5, FB_STRIDE = 5 + 3 * 32 };

// Per-beam raster set-up: bounding box, pixel size and the edge lines
// a x + b y + c (> 0 inside) of the convex aperture polygon.  One thread per
// beam.  rec: [FB_STRIDE] doubles.
__global__ void fft_setup_kernel(int n_beams, const int64_t *__restrict__ beams,
                                 const int32_t *__restrict__ nv_in,
                                 const int64_t *__restrict__ voff,
                                 const double *__restrict__ vx, const double *__restrict__ vy,
                                 int n, double pad, double *__restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_beams)
        return;
    const int64_t b = beams[i];
    const int nv = nv_in[b];
    const int64_t v0 = voff[b];
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300, area2 = 0.0;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double xe = vx[v0 + e], ye = vy[v0 + e];
        x0 = fmin(x0, xe);
        x1 = fmax(x1, xe);
        y0 = fmin(y0, ye);
        y1 = fmax(y1, ye);
        area2 += xe * vy[v0 + f] - vx[v0 + f] * ye;
    }
    double *r = rec + (size_t)i * FB_STRIDE;
    r[FB_XC] = 0.5 * (x0 + x1);
    r[FB_YC] = 0.5 * (y0 + y1);
    r[FB_DX] = pad * (x1 - x0) / n;
    r[FB_DY] = pad * (y1 - y0) / n;
    r[FB_NE] = nv;
    const double sg = area2 >= 0.0 ? 1.0 : -1.0;  // inside on the left of each edge
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double xe = vx[v0 + e], ye = vy[v0 + e];
        const double a = -(vy[v0 + f] - ye) * sg, bb = (vx[v0 + f] - xe) * sg;
        r[FB_EDGE + 3 * e + 0] = a;
        r[FB_EDGE + 3 * e + 1] = bb;
        r[FB_EDGE + 3 * e + 2] = -__builtin_fma(a, xe, bb * ye);
    }
}

// Pixel centre coordinate along one axis.
__device__ __forceinline__ double fft_pixel(int j, int n, double d, double c)
{
    return __builtin_fma((double)(2 * j + 1 - n), 0.5 * d, c);
}

__device__ __forceinline__ bool fft_inside(const double *__restrict__ r, int ne, double px,
                                           double py)
{
    bool in = true;
    for (int e = 0; e < ne; ++e) {
        const double *l = r + FB_EDGE + 3 * e;
        in = in && (__builtin_fma(l[0], px, __builtin_fma(l[1], py, l[2])) > 0.0);
    }
    return in;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// twiddles tw[m] = exp(-2 pi i m / n), m < n/2, computed once per call in
// FP64 and rounded; the FFT kernels copy them into LDS.
__global__ void fft_twiddle_kernel(int n, float2 *__restrict__ tw)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n / 2) {
        double s, c;
        sincospi(-2.0 * (double)m / (double)n, &s, &c);
        tw[m] = make_float2((float)c, (float)s);
    }
}

__device__ __forceinline__ void fft_twiddles(float2 *tw, const float2 *__restrict__ gtw, int n)
{
    for (int m = threadIdx.x; m < n / 2; m += blockDim.x)
        tw[m] = gtw[m];
}

// In-place FFTs of `count` length-n sequences stored in LDS at a stride of
// n + FFT_PAD (already in bit-reversed order).  All threads of the block
// take part.
__device__ __forceinline__ void fft_lds(float2 *data, int count, int n, int logn,
                                        const float2 *__restrict__ tw)
{
    const int quads = count * (n >> 2);
    int s = 0;
    for (; s + 1 < logn; s += 2) {
        const int h = 1 << s;
        const int sh1 = logn - 1 - s;  // k * n/(2h) = k << sh1
        for (int q = threadIdx.x; q < quads; q += blockDim.x) {
            const int seq = q >> (logn - 2);
            const int g = q & ((n >> 2) - 1);
            const int k = g & (h - 1);
            const int i = ((g >> s) << (s + 2)) | k;
            float2 *d = data + seq * (n + FFT_PAD);
            const float2 w1 = tw[k << sh1];
            const float2 w2 = tw[k << (sh1 - 1)];
            const float2 x0 = d[i], x1 = d[i + h], x2 = d[i + 2 * h], x3 = d[i + 3 * h];
            const float2 t1 = cmulf(w1, x1), t3 = cmulf(w1, x3);
            const float2 y0 = make_float2(x0.x + t1.x, x0.y + t1.y);
            const float2 y1 = make_float2(x0.x - t1.x, x0.y - t1.y);
            const float2 y2 = make_float2(x2.x + t3.x, x2.y + t3.y);
            const float2 y3 = make_float2(x2.x - t3.x, x2.y - t3.y);
            const float2 u2 = cmulf(w2, y2);
            const float2 w2y3 = cmulf(w2, y3);
            const float2 u3 = make_float2(w2y3.y, -w2y3.x);  // (-i) w2 y3
            d[i] = make_float2(y0.x + u2.x, y0.y + u2.y);
            d[i + 2 * h] = make_float2(y0.x - u2.x, y0.y - u2.y);
            d[i + h] = make_float2(y1.x + u3.x, y1.y + u3.y);
            d[i + 3 * h] = make_float2(y1.x - u3.x, y1.y - u3.y);
        }
        __syncthreads();
    }
    if (s < logn) {  // last single stage: h = n/2, twiddle tw[i]
        const int h = 1 << s;
        const int pairs = count * h;
        for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
            const int seq = p >> s;
            const int i = p & (h - 1);
            float2 *d = data + seq * (n + FFT_PAD);
            const float2 t = cmulf(tw[i], d[i + h]);
            const float2 a = d[i];
            d[i] = make_float2(a.x + t.x, a.y + t.y);
            d[i + h] = make_float2(a.x - t.x, a.y - t.y);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int fft_bitrev(int x, int logn)
{
    return (int)(__builtin_bitreverse32((uint32_t)x) >> (32 - logn));
}

// Grid index of workgroup work: XCD-aware, so that consecutive work units
// (which share cache lines) run on the same XCD.
__device__ __forceinline__ int fft_logical_block()
{
    const int nblk = gridDim.x;
    const int per_xcd = (nblk + 7) / 8;
    return (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
}

// Raster + row FFTs.  Work unit = (beam, ROWS rows); dynamic LDS:
// ROWS * (n + FFT_PAD) float2 + n/2 float2 twiddles.  Thread j rasterises
// pixel bitrev(j) into LDS slot j (linear, conflict-free stores).
template <int ROWS>
__global__ void __launch_bounds__(FFT_ROW_THREADS) fft_rows_kernel(
    int n_beams, int n, int logn, const double *__restrict__ rec,
    const float2 *__restrict__ gtw, float2 *__restrict__ grid)
{
    extern __shared__ float2 lds[];
    const int units_per_beam = n / ROWS;
    const int unit = fft_logical_block();
    if (unit >= n_beams * units_per_beam)
        return;
    const int beam = unit / units_per_beam;
    const int row0 = (unit - beam * units_per_beam) * ROWS;
    const int ld = n + FFT_PAD;
    float2 *data = lds;
    float2 *tw = lds + ROWS * ld;
    fft_twiddles(tw, gtw, n);
    const double *r = rec + (size_t)beam * FB_STRIDE;
    const int ne = (int)r[FB_NE];
    const double xc = r[FB_XC], yc = r[FB_YC], dx = r[FB_DX], dy = r[FB_DY];
    for (int rr = 0; rr < ROWS; ++rr) {
        const double py = fft_pixel(row0 + rr, n, dy, yc);
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const bool in = fft_inside(r, ne, fft_pixel(fft_bitrev(j, logn), n, dx, xc), py);
            data[rr * ld + j] = make_float2(in ? 1.0f : 0.0f, 0.0f);
        }
    }
    __syncthreads();
    fft_lds(data, ROWS, n, logn, tw);
    float2 *g = grid + ((size_t)beam * n + row0) * n;
    for (int rr = 0; rr < ROWS; ++rr)
        for (int x = threadIdx.x; x < n; x += blockDim.x)
            g[(size_t)rr * n + x] = data[rr * ld + x];
}

// Column FFTs.  Work unit = (beam, COLS adjacent columns); dynamic LDS:
// COLS * (n + FFT_PAD) float2 + n/2 float2 twiddles.  Lanes of a wave read
// COLS columns of 64/COLS rows (COLS*8-byte row segments); row bitrev(j)
// goes to slot j.
template <int COLS>
__global__ void __launch_bounds__(FFT_COL_THREADS) fft_cols_kernel(
    int n_beams, int n, int logn, const float2 *__restrict__ gtw, float2 *__restrict__ grid)
{
    extern __shared__ float2 lds[];
    const int units_per_beam = n / COLS;
    const int unit = fft_logical_block();
    if (unit >= n_beams * units_per_beam)
        return;
    const int beam = unit / units_per_beam;
    const int col0 = (unit - beam * units_per_beam) * COLS;
    const int ld = n + FFT_PAD;
    float2 *data = lds;
    float2 *tw = lds + COLS * ld;
    fft_twiddles(tw, gtw, n);
    float2 *g = grid + (size_t)beam * n * n + col0;
    const int c = threadIdx.x & (COLS - 1);
    const int r_step = blockDim.x / COLS;
    for (int j = threadIdx.x / COLS; j < n; j += r_step)
        data[c * ld + j] = g[(size_t)fft_bitrev(j, logn) * n + c];
    __syncthreads();
    fft_lds(data, COLS, n, logn, tw);
    for (int row = threadIdx.x / COLS; row < n; row += r_step)
        g[(size_t)row * n + c] = data[c * ld + row];
}

__device__ __forceinline__ double lanczos(double x, int radius)
{
    const double ax = fabs(x);
    if (ax >= radius)
        return 0.0;
    if (ax < 1e-12)
        return 1.0;
    const double px = 3.141592653589793116 * x;
    return sin(px) / px * (sin(px / radius) / (px / radius));
}

// Sample I(A, B) of each beam at each point, by Lanczos interpolation of the
// phase-corrected spectrum.  One thread per (beam, point).  out: [beam][pt][2].
__global__ void fft_sample_kernel(int n_beams, int n_points, int n, const double *__restrict__ rec,
                                  const float2 *__restrict__ grid, const double *__restrict__ ab,
                                  double k, int radius, double *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_beams * n_points)
        return;
    const int beam = idx / n_points;
    const int p = idx - beam * n_points;
    const double *r = rec + (size_t)beam * FB_STRIDE;
    const double dx = r[FB_DX], dy = r[FB_DY];
    const double inv2pi = 0.15915494309189533577;
    const double fx = -k * ab[2 * p] * inv2pi, fy = -k * ab[2 * p + 1] * inv2pi;
    const double ux = fx * n * dx, uy = fy * n * dy;
    const int cx = (int)floor(ux), cy = (int)floor(uy);
    const float2 *g = grid + (size_t)beam * n * n;
    const double corr = 3.141592653589793116 * (double)(n - 1) / (double)n;
    double sr = 0.0, si = 0.0, wsum = 0.0;
    for (int jy = cy - radius + 1; jy <= cy + radius; ++jy) {
        const double wy = lanczos(uy - jy, radius);
        if (wy == 0.0)
            continue;
        const int iy = ((jy % n) + n) % n;
        for (int jx = cx - radius + 1; jx <= cx + radius; ++jx) {
            const double w = wy * lanczos(ux - jx, radius);
            if (w == 0.0)
                continue;
            const int ix = ((jx % n) + n) % n;
            const float2 z = g[(size_t)iy * n + ix];
            double s, c;
            sincos(corr * (double)(jx + jy), &s, &c);
            sr += w * ((double)z.x * c - (double)z.y * s);
            si += w * ((double)z.x * s + (double)z.y * c);
            wsum += w;
        }
    }
    if (fabs(wsum) > 1e-30) {
        sr /= wsum;
        si /= wsum;
    }
    // pixel area and the grid-centre shift exp(-2 pi i f.c)
    double s, c;
    sincos(-2.0 * 3.141592653589793116 * (fx * r[FB_XC] + fy * r[FB_YC]), &s, &c);
    const double a = dx * dy;
    out[2 * (size_t)idx] = a * (sr * c - si * s);
    out[2 * (size_t)idx + 1] = a * (sr * s + si * c);
}

}  // namespace mbs
