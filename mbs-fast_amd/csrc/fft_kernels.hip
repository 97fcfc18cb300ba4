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
//   fft_rows_kernel   rasterises each row's window of L pixels (no global
//                     read), transforms it as n / L transforms of length L,
//                     writes the row (n^2 x 8 B at most: rows of pixel
//                     centres outside the polygon's y range are zero and
//                     skipped -- with padding p, all but about 1/p of them)
//   fft_cols_kernel   reads WCOLS adjacent columns (64-byte row segments) of
//                     the rows written, as n / L transforms of length L over
//                     the window of L rows that holds them, and writes the
//                     whole grid (n^2 x 8 B)
// The 1-D FFT is the Stockham autosort formulation (Govindaraju et al.,
// "High performance discrete Fourier transforms on graphics processors",
// SC08): radix-16 stages and one radix-8, -4 or -2 stage, each thread taking
// whole butterflies into registers, with a barrier between the read and the
// write phase so the stages run in place in one LDS buffer.  The first stage
// reads the raster (row pass) or HBM (column pass) directly and the last
// stage writes HBM directly, so each pass touches HBM once per element.
// Twiddles come from a table of exp(-2 pi i m / n) made once per call in
// FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

// Workgroups: 128 threads per row or column.  The row kernel holds FFT_ROWS
// rows: one row (17 KB of LDS at n = 2048, nine workgroups per CU) ran the
// row pass 16 % faster than four (68 KB, two per CU), whose barriers stall
// more of the CU at once.  FFT_COLS columns of full length per workgroup
// (the spectrum source's column pass): 4 (two 68 KB workgroups of 512
// threads per CU) ran a full-length column pass of 64 padded 2048^2 grids
// in 1.13 ms, 8 (one 135 KB workgroup) in 1.31 ms and 2 in 1.54 ms.
#ifndef MBS_FFT_COLS
#define MBS_FFT_COLS 4
#endif
#ifndef MBS_FFT_ROWS
#define MBS_FFT_ROWS 1
#endif
enum {
    FFT_ROWS = MBS_FFT_ROWS,
    FFT_ROW_THREADS = MBS_FFT_ROWS * 128,
    FFT_COLS = MBS_FFT_COLS,
    FFT_COL_THREADS = MBS_FFT_COLS * 128
};
enum {
    FB_XC = 0,
    FB_YC = 1,
    FB_DX = 2,
    FB_DY = 3,
    FB_NE = 4,
    FB_EDGE = 5,                // (a, b, c) per edge, up to 32 edges
    FB_YLO = FB_EDGE + 3 * 32,  // the polygon's y range
    FB_YHI = FB_YLO + 1,
    FB_STRIDE = FB_YHI + 1
};

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
    r[FB_YLO] = y0;
    r[FB_YHI] = y1;
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

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// twiddles tw[m] = exp(-2 pi i m / n), m < n, computed once per call in
// FP64 and rounded.
__global__ void fft_twiddle_kernel(int n, float2 *__restrict__ tw)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        double s, c;
        sincospi(-2.0 * (double)m / (double)n, &s, &c);
        tw[m] = make_float2((float)c, (float)s);
    }
}

// Grid index of workgroup work: XCD-aware, so that consecutive work units
// (which share cache lines) run on the same XCD.
__device__ __forceinline__ int fft_logical_block()
{
    const int nblk = gridDim.x;
    const int per_xcd = (nblk + 7) / 8;
    return (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
}

// LDS index of element i of a sequence: one pad slot per 32 elements
// (breaks the power-of-two strides of the early Stockham stages).
__device__ __forceinline__ int fft_pad(int i) { return i + (i >> 5); }

// In-place forward DFT of R points (exp(-2 pi i j k / R)), R = 2, 4, 8, 16.
template <int R>
__device__ __forceinline__ void fft_dft(float2 *v)
{
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const float2 a0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        const float2 a1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        const float2 a2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        const float2 a3 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        v[0] = make_float2(a0.x + a2.x, a0.y + a2.y);
        v[2] = make_float2(a0.x - a2.x, a0.y - a2.y);
        v[1] = make_float2(a1.x + a3.y, a1.y - a3.x);  // a1 - i a3
        v[3] = make_float2(a1.x - a3.y, a1.y + a3.x);  // a1 + i a3
    } else if constexpr (R == 8) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        fft_dft<4>(e);
        fft_dft<4>(o);
        const float h = 0.70710678118654752f;
        const float2 o1 = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));   // w^1
        const float2 o2 = make_float2(o[2].y, -o[2].x);                                // w^2 = -i
        const float2 o3 = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));  // w^3
        v[0] = make_float2(e[0].x + o[0].x, e[0].y + o[0].y);
        v[4] = make_float2(e[0].x - o[0].x, e[0].y - o[0].y);
        v[1] = make_float2(e[1].x + o1.x, e[1].y + o1.y);
        v[5] = make_float2(e[1].x - o1.x, e[1].y - o1.y);
        v[2] = make_float2(e[2].x + o2.x, e[2].y + o2.y);
        v[6] = make_float2(e[2].x - o2.x, e[2].y - o2.y);
        v[3] = make_float2(e[3].x + o3.x, e[3].y + o3.y);
        v[7] = make_float2(e[3].x - o3.x, e[3].y - o3.y);
    } else {
        // radix 2 over two 8-point DFTs: X[k] = E[k] + w^k O[k],
        // X[k+8] = E[k] - w^k O[k], w = exp(-2 pi i / 16)
        float2 e[8], o[8];
        for (int q = 0; q < 8; ++q) {
            e[q] = v[2 * q];
            o[q] = v[2 * q + 1];
        }
        fft_dft<8>(e);
        fft_dft<8>(o);
        const float c1 = 0.92387953251128676f;  // cos(pi/8)
        const float s1 = 0.38268343236508977f;  // sin(pi/8)
        const float h = 0.70710678118654752f;
        float2 w[8];
        w[0] = o[0];
        w[1] = make_float2(o[1].x * c1 + o[1].y * s1, o[1].y * c1 - o[1].x * s1);
        w[2] = make_float2(h * (o[2].x + o[2].y), h * (o[2].y - o[2].x));
        w[3] = make_float2(o[3].x * s1 + o[3].y * c1, o[3].y * s1 - o[3].x * c1);
        w[4] = make_float2(o[4].y, -o[4].x);
        w[5] = make_float2(o[5].y * c1 - o[5].x * s1, -o[5].y * s1 - o[5].x * c1);
        w[6] = make_float2(h * (o[6].y - o[6].x), -h * (o[6].x + o[6].y));
        w[7] = make_float2(o[7].y * s1 - o[7].x * c1, -o[7].y * c1 - o[7].x * s1);
        for (int q = 0; q < 8; ++q) {
            v[q] = make_float2(e[q].x + w[q].x, e[q].y + w[q].y);
            v[q + 8] = make_float2(e[q].x - w[q].x, e[q].y - w[q].y);
        }
    }
}

// One Stockham stage of radix R over `cnt` length-n sequences; ns = product
// of the radices of the earlier stages.  Butterfly idx covers sequence seq,
// position j (seq fastest when COLFAST, else j fastest); it reads elements
// j + r n/R (twiddled by exp(-2 pi i k r / (ns R)), k = j mod ns) and writes
// (j / ns) ns R + k + r ns.  load(seq, i) / store(seq, i, v) give the
// element access; all reads of the stage precede all its writes.  tw holds
// exp(-2 pi i m / tw_n), tw_n a multiple of n.
template <int R, int PB, bool COLFAST, typename Load, typename Store>
__device__ __forceinline__ void fft_stage(int n, int lg_ns, int cnt,
                                          const float2 *__restrict__ tw, int tw_n, Load load,
                                          Store store)
{
    const int nb = n / R;
    const int total = cnt * nb;
    const int ns = 1 << lg_ns;
    const int tstep = tw_n / (ns * R);
    float2 v[PB][R];
    for (int b = 0; b < PB; ++b) {
        const int idx = threadIdx.x + b * blockDim.x;
        if (idx >= total)
            break;
        const int seq = COLFAST ? idx % cnt : idx / nb;
        const int j = COLFAST ? idx / cnt : idx - seq * nb;
        const int k = j & (ns - 1);
        for (int r = 0; r < R; ++r) {
            float2 x = load(seq, j + r * nb);
            if (r > 0 && k > 0)
                x = cmulf(x, tw[k * r * tstep]);
            v[b][r] = x;
        }
    }
    __syncthreads();
    for (int b = 0; b < PB; ++b) {
        const int idx = threadIdx.x + b * blockDim.x;
        if (idx >= total)
            break;
        const int seq = COLFAST ? idx % cnt : idx / nb;
        const int j = COLFAST ? idx / cnt : idx - seq * nb;
        const int k = j & (ns - 1);
        fft_dft<R>(v[b]);
        const int base = ((j >> lg_ns) << lg_ns) * R + k;
        for (int r = 0; r < R; ++r)
            store(seq, base + r * ns, v[b][r]);
    }
    __syncthreads();
}

// Stages S..end of a length 2^LOGN transform: radix 2^LR while that many
// bits remain, then one stage of the rest.  Stage 0 reads through `first`,
// the last stage writes through `last`, stages in between stay in LDS.
// Radix 16 (three stages at n = 2048: 16, 16, 8) makes one LDS round trip and
// two barriers fewer than radix 8 (8, 8, 8, 4).  tw: exp(-2 pi i m / 2^LOGTW).
#ifndef MBS_FFT_LOG_RADIX
#define MBS_FFT_LOG_RADIX 4
#endif
// LDS index of element i of sequence seq, CNT sequences interleaved: the
// sequences' elements i sit together in one block of CNT x 8 bytes, and the
// low bits of the block index are XORed with the bits of i above the first
// stage's radix 2^LR, so that every read and write pattern of a Stockham
// stage (column-fastest lanes, 2^LR-radix butterflies) hits 32 lanes' worth
// of distinct banks: consecutive elements, and the first stage's
// outputs j 2^LR + r of consecutive j.
template <int CNT, int LR>
__device__ __forceinline__ int fft_swz(int seq, int i)
{
    constexpr int M = (32 / CNT > 1 ? 32 / CNT : 1) - 1;
    static_assert(M < (1 << LR), "the XOR source bits lie above the bits they change");
    return (i ^ ((i >> LR) & M)) * CNT + seq;
}

template <int LOGN, int CNT, int T, bool COLFAST, int S, int LR = MBS_FFT_LOG_RADIX,
          int LOGTW = LOGN, bool SWZ = false, typename First, typename Last>
__device__ __forceinline__ void fft_stages(float2 *buf, int ld, const float2 *__restrict__ tw,
                                           First first, Last last)
{
    constexpr int N = 1 << LOGN;
    constexpr int A = LOGN / LR, REM = LOGN % LR, NST = A + (REM ? 1 : 0);
    if constexpr (S < NST) {
        constexpr int R = S < A ? (1 << LR) : (1 << REM);
        constexpr int PB = (CNT * (N / R) + T - 1) / T;
        auto at = [&](int seq, int i) {
            if constexpr (SWZ)
                return fft_swz<CNT, LR>(seq, i);
            else
                return seq * ld + fft_pad(i);
        };
        auto lds_load = [&](int seq, int i) { return buf[at(seq, i)]; };
        auto lds_store = [&](int seq, int i, float2 x) { buf[at(seq, i)] = x; };
        constexpr int TWN = 1 << LOGTW;
        if constexpr (S == 0)
            fft_stage<R, PB, COLFAST>(N, LR * S, CNT, tw, TWN, first, lds_store);
        else if constexpr (S == NST - 1)
            fft_stage<R, PB, COLFAST>(N, LR * S, CNT, tw, TWN, lds_load, last);
        else
            fft_stage<R, PB, COLFAST>(N, LR * S, CNT, tw, TWN, lds_load, lds_store);
        fft_stages<LOGN, CNT, T, COLFAST, S + 1, LR, LOGTW, SWZ>(buf, ld, tw, first, last);
    }
}

// LDS elements per sequence: padded length, plus 8 so that consecutive
// sequences (the columns of the column pass) start 16 banks apart.
__host__ __device__ constexpr int fft_ld(int n) { return n + (n >> 5) + 8; }

// Row of pixel centres outside the polygon's y range: all zero.  Such rows
// are neither transformed nor stored; the column pass reads them as zero.
__device__ __forceinline__ bool fft_zero_row(const double *__restrict__ r, int row, int n)
{
    const double py = fft_pixel(row, n, r[FB_DY], r[FB_YC]);
    return py < r[FB_YLO] || py > r[FB_YHI];
}

// Column FFTs over a row window.  Rows of pixel centres outside the
// polygon's bounding box are zero, and with padding p those lie outside a
// window of L = 2^LOGL <= N rows, [j0, j0 + L), L the power of two at or
// above n / p (fft_window).  With D = N / L, w = exp(-2 pi i / N) and bins
// m = D q + s (s < D), w^(row m) = w^(row s) w_L^((row mod L) q), w_L = w^D,
// so
//   X[D q + s] = sum_p z_s[p] w_L^(p q),  z_s[row mod L] = x[row] w^(row s):
// each column's transform is D transforms of length L, one per residue s, of
// the window placed circularly (row mod L) and twiddled by w^(row s).  The
// LDS holds L elements per column instead of N; the window stays in
// registers across the residues, twiddled from one residue to the next by
// one complex product with w^row, and each residue's bins are stored while
// the next residue is transformed.  Nothing between the stores and the end
// reads global memory (the L-point twiddles w_L^i sit in LDS), so no wait on
// the load counter, which also counts stores, holds the stores up; the
// window's loads are issued unconditionally (zero rows are selected away
// afterwards) so that they are all in flight at once.
// Work unit = (beam, FFT_WCOLS adjacent columns); lanes take columns
// fastest, so a wave reads and writes FFT_WCOLS x 8-byte row segments.  LDS:
// FFT_WCOLS x L float2 in the fft_swz layout, then the L twiddles.
// Column pass of 64 grids of 2048^2 at padding 4 (L = 512, four residues),
// by columns x threads per workgroup: 8 x 512 0.67 ms, 16 x 1024 0.70 ms,
// 8 x 256 and 16 x 512 0.78 ms, 32 x 1024 1.00 ms, 4 x 256 1.04 ms; radix-16
// first stages (8 x 512) 1.02 ms.  The full-length column pass it replaced
// (4 columns of 2048 rows per workgroup) took 1.05 ms.
#ifndef MBS_FFT_WCOLS
#define MBS_FFT_WCOLS 8
#endif
#ifndef MBS_FFT_WCOL_THREADS
#define MBS_FFT_WCOL_THREADS 512
#endif
#ifndef MBS_FFT_WCOL_LOG_RADIX
#define MBS_FFT_WCOL_LOG_RADIX 3
#endif
enum { FFT_WCOLS = MBS_FFT_WCOLS, FFT_WCOL_THREADS = MBS_FFT_WCOL_THREADS };

// Shortest window: 64 rows (LOGL >= 6 keeps at least two Stockham stages),
// and n / 8 (at most eight residues).
enum { FFT_MIN_LOGL = 6 };
__host__ __device__ constexpr int fft_min_logl(int logn)
{
    return logn - 3 > FFT_MIN_LOGL ? logn - 3 : FFT_MIN_LOGL;
}

// The row window of an n-point grid at padding pad: pixel centres at
// j - (n - 1)/2 pixels from the box centre, the box n / (2 pad) pixels
// either side of it, so rows lo..hi below can carry the polygon (with 1e-3
// pixel to spare for rounding); L the power of two at or above hi - lo + 1
// (64 and n / 8 at least -- at most eight residues -- and n at most) and j0
// the window start, clamped into the grid.  The columns that can carry the
// polygon are the same window: the box is n / pad pixels wide either way.
__host__ inline void fft_window(int n, double pad, int *j0, int *logl)
{
    const double c = 0.5 * (n - 1), h = n / (2.0 * pad);
    const int lo = (int)ceil(c - h - 1e-3), hi = (int)floor(c + h + 1e-3);
    int lg = FFT_MIN_LOGL;
    while ((1 << lg) < n && ((1 << lg) < hi - lo + 1 || (8 << lg) < n))
        ++lg;  // ends at fft_min_logl(log2 n) or above
    const int l = 1 << lg;
    *logl = lg;
    *j0 = lo < 0 ? 0 : (lo > n - l ? n - l : lo);
}

template <int LOGN, int LOGL>
__global__ void __launch_bounds__(FFT_WCOL_THREADS) fft_cols_kernel(
    int n_beams, int j0, const double *__restrict__ rec, const float2 *__restrict__ tw,
    float2 *__restrict__ grid)
{
    constexpr int N = 1 << LOGN, L = 1 << LOGL, D = N / L;
    constexpr int LR = LOGL < MBS_FFT_WCOL_LOG_RADIX ? LOGL : MBS_FFT_WCOL_LOG_RADIX;
    constexpr int R = 1 << LR, NB = L / R;  // first-stage radix, butterflies per column
    constexpr int C = FFT_WCOLS, T = FFT_WCOL_THREADS;
    constexpr int PB = (C * NB + T - 1) / T;
    extern __shared__ float2 lds[];
    float2 *twl = lds + C * L;  // w_L^i = w^(i D), i < L
    const int units_per_beam = N / C;
    const int unit = fft_logical_block();
    if (unit >= n_beams * units_per_beam)
        return;
    const int beam = unit / units_per_beam;
    const int col0 = (unit - beam * units_per_beam) * C;
    float2 *g = grid + (size_t)beam * N * N + col0;
    const double *r = rec + (size_t)beam * FB_STRIDE;
    for (int i = threadIdx.x; i < L; i += T)
        twl[i] = tw[i * D];
    // the window in first-stage order: element p = j + e NB of butterfly b's
    // column is row j0 + ((p - j0) mod L); y = x w^(row s), wr = w^row
    float2 y[PB][R], wr[PB][R];
    for (int b = 0; b < PB; ++b) {
        const int idx = (threadIdx.x + b * T) % (C * NB);  // lanes past the end repeat
        const int seq = idx % C, j = idx / C;
        for (int e = 0; e < R; ++e) {
            const int row = j0 + ((j + e * NB - j0) & (L - 1));
            y[b][e] = g[(size_t)row * N + seq];
            wr[b][e] = tw[row];
        }
        for (int e = 0; e < R; ++e) {
            const int row = j0 + ((j + e * NB - j0) & (L - 1));
            if (fft_zero_row(r, row, N))
                y[b][e] = make_float2(0.0f, 0.0f);
        }
    }
    // every load is back here, before the first store: the compiler's own
    // wait for a loaded register used after a store would also wait for the
    // store (s_waitcnt vmcnt(0); expcnt and lgkmcnt left alone)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    auto none = [&](int, int) { return make_float2(0.0f, 0.0f); };
    for (int s = 0; s < D; ++s) {
        // first stage (ns = 1): R-point DFT of the twiddled window, stored at
        // j R + e
        for (int b = 0; b < PB; ++b) {
            const int idx = threadIdx.x + b * T;
            if (idx >= C * NB)
                break;
            const int seq = idx % C, j = idx / C;
            float2 v[R];
            for (int e = 0; e < R; ++e) {
                v[e] = y[b][e];
                if (s + 1 < D)
                    y[b][e] = cmulf(y[b][e], wr[b][e]);
            }
            fft_dft<R>(v);
            for (int e = 0; e < R; ++e)
                lds[fft_swz<C, LR>(seq, j * R + e)] = v[e];
        }
        __syncthreads();
        auto out = [&](int seq, int q, float2 v) { g[(size_t)(D * q + s) * N + seq] = v; };
        fft_stages<LOGL, C, T, true, 1, LR, LOGL, true>(lds, 0, twl, none, out);
    }
}

// Raster + row FFTs over the column window [j0, j0 + L) (fft_window): as in
// the column pass, each row's transform is D = N / L transforms of length L
// of the rasterised window placed circularly (col mod L) and twiddled by
// w^(col s).  Thread j keeps bins q = j + e L / R of every residue in
// registers and stores them at the end: bins m = D q .. D q + D - 1 are
// consecutive, so each wave's stores are contiguous (storing each residue's
// bins as it finishes, a stride-D scatter, took the row pass of 64 padded
// 2048^2 grids from 0.21 to 0.44 ms), and no store is outstanding.  Every global read happens before the first stage: the
// window's twiddles w^col into registers, the L-point stage twiddles into
// LDS (a dependent global read per stage and residue would put the load
// latency on the critical path twelve times per row).  Work unit = (beam,
// window row); rows outside the polygon's y range neither transform nor
// store (the column pass reads them as zero).  One or two waves per row;
// LDS fft_ld(L) + L float2 plus the row's edge offsets (b y + c per edge).
#ifndef MBS_FFT_WROW_LOG_RADIX
#define MBS_FFT_WROW_LOG_RADIX 3
#endif
// Threads per row: one first-stage butterfly each, 64 to 128.
__host__ __device__ constexpr int fft_wrow_threads(int logl)
{
    return logl - MBS_FFT_WROW_LOG_RADIX <= 6 ? 64 : 128;
}
// Registers held to four waves per SIMD up to four residues: the row pass
// of 64 padded 2048^2 grids took 0.20 ms, against 0.22 ms at the 132 VGPRs
// the compiler takes unbounded (three waves) and 0.24 ms at five (spills).
// Eight residues would spill at four waves and are left unbounded.
#ifndef MBS_FFT_WROW_WAVES
#define MBS_FFT_WROW_WAVES 4
#endif
template <int LOGN, int LOGL>
__global__ void __launch_bounds__(fft_wrow_threads(LOGL),
                                  LOGN - LOGL <= 2 ? MBS_FFT_WROW_WAVES : 1) fft_rows_kernel(
    int n_beams, int j0, const double *__restrict__ rec, const float2 *__restrict__ tw,
    float2 *__restrict__ grid)
{
    constexpr int N = 1 << LOGN, L = 1 << LOGL, D = N / L;
    constexpr int LR = LOGL < MBS_FFT_WROW_LOG_RADIX ? LOGL : MBS_FFT_WROW_LOG_RADIX;
    constexpr int R = 1 << LR, NB = L / R;  // first-stage radix, butterflies
    constexpr int T = fft_wrow_threads(LOGL);
    constexpr int PB = (NB + T - 1) / T;
    constexpr int LD = fft_ld(L);
    static_assert(D <= 8, "at most eight residues (fft_window)");
    extern __shared__ float2 lds[];
    const int unit = fft_logical_block();
    if (unit >= n_beams * L)
        return;
    const int beam = unit / L;
    const int row = j0 + (unit - beam * L);
    const double *r = rec + (size_t)beam * FB_STRIDE;
    if (fft_zero_row(r, row, N))
        return;
    const int ne = (int)r[FB_NE];
    const double xc = r[FB_XC], dx = r[FB_DX];
    float2 *twl = lds + LD;  // w_L^i = w^(i D), i < L
    double *off = reinterpret_cast<double *>(twl + L);
    const double py = fft_pixel(row, N, r[FB_DY], r[FB_YC]);
    for (int e = threadIdx.x; e < ne; e += T)
        off[e] = __builtin_fma(r[FB_EDGE + 3 * e + 1], py, r[FB_EDGE + 3 * e + 2]);
    {
        static_assert(L % T == 0, "whole table rows per thread");
        float2 t[L / T];  // all loads in flight before the first LDS write
        for (int k = 0; k < L / T; ++k)
            t[k] = tw[(threadIdx.x + k * T) * D];
        for (int k = 0; k < L / T; ++k)
            twl[threadIdx.x + k * T] = t[k];
    }
    __syncthreads();
    // the rasterised window in first-stage order (element p = j + e NB is
    // column j0 + ((p - j0) mod L)); y = raster w^(col s), wr = w^col
    float2 y[PB][R], wr[PB][R];
    for (int b = 0; b < PB; ++b) {
        const int j = (threadIdx.x + b * T) % NB;  // lanes past the end repeat
        for (int e = 0; e < R; ++e) {
            const int col = j0 + ((j + e * NB - j0) & (L - 1));
            wr[b][e] = tw[col];
            const double px = fft_pixel(col, N, dx, xc);
            bool in = true;
            for (int q = 0; q < ne; ++q)
                in = in && (__builtin_fma(r[FB_EDGE + 3 * q], px, off[q]) > 0.0);
            y[b][e] = make_float2(in ? 1.0f : 0.0f, 0.0f);
        }
    }
    float2 o[D][PB][R];
    auto none = [&](int, int) { return make_float2(0.0f, 0.0f); };
    auto keep = [&](int, int i, float2 x) { lds[fft_pad(i)] = x; };
#pragma unroll
    for (int s = 0; s < D; ++s) {
        for (int b = 0; b < PB; ++b) {
            const int j = threadIdx.x + b * T;
            if (j >= NB)
                break;
            float2 v[R];
            for (int e = 0; e < R; ++e) {
                v[e] = y[b][e];
                if (s + 1 < D)
                    y[b][e] = cmulf(y[b][e], wr[b][e]);
            }
            fft_dft<R>(v);
            for (int e = 0; e < R; ++e)
                lds[fft_pad(j * R + e)] = v[e];
        }
        __syncthreads();
        fft_stages<LOGL, 1, T, false, 1, LR, LOGL>(lds, LD, twl, none, keep);
        for (int b = 0; b < PB; ++b)
            for (int e = 0; e < R; ++e)
                o[s][b][e] = lds[fft_pad((threadIdx.x + b * T) % NB + e * NB)];
        __syncthreads();
    }
    float2 *g = grid + ((size_t)beam * N + row) * N;
    for (int b = 0; b < PB; ++b) {
        const int j = threadIdx.x + b * T;
        if (j >= NB)
            break;
        for (int e = 0; e < R; ++e) {
            float2 *dst = g + D * (j + e * NB);
            if constexpr (D == 1) {
                dst[0] = o[0][b][e];
            } else {
                for (int s = 0; s < D; s += 2)
                    *reinterpret_cast<float4 *>(dst + s) =
                        make_float4(o[s][b][e].x, o[s][b][e].y, o[s + 1][b][e].x, o[s + 1][b][e].y);
            }
        }
    }
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
