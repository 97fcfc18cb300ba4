// upsample_kernels.hip -- band-limited (Fourier) upsampling of Mueller grids
// on the device: the interpolation step of the reference's angular FFT path
// (fft_upsample_phi_arr2d and fft_phi_pad_kernel, src/cuda/
// GpuDiffraction.cu:3195-3330 and :794-814, cuFFT there; the theta form
// fft_upsample_theta_arr2d, :3343-3420).
//
// A "line" is one sequence of grid samples: the n_az azimuths of one theta
// row (periodic), or the 2 n_zen samples of one azimuth column extended
// evenly about both ends of the zenith range (rows t and 2 n_zen - t).  Two
// of the 16 Mueller elements travel together as the real and imaginary
// parts of one complex line, which the transform keeps apart because the
// padded spectrum stays Hermitian for real data: an even count's Nyquist
// term is split evenly between +n/2 and -n/2 (for real data this equals the
// real part of the reference's one-sided placement).  Per line: forward DFT
// of n_in points, zero padding to n_out, scale 1/n_in, inverse DFT of n_out
// points.
//
// The DFTs are mixed-radix Stockham autosort FFTs (Govindaraju et al.,
// SC08, generalised to mixed radices) in LDS, FP64, one workgroup per line.
// A stage of radix R over Ns already-combined points maps output y =
// (jq Ns R + k Ns + jm) to
//   sum_r x[jq Ns + jm + r n/R] * w^(r (jm n/(Ns R) + k n/R)),  w = e^(-2 pi i/n),
// so every twiddle is a power of the n-th root, read from a table made once
// per length with sincospi.  Radices 4, 2, 3, 5, 7 first, then any other
// prime factor directly (O(n p) for a factor p).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

constexpr int kUpMaxN = 4096;      // longest line: two LDS buffers of up to 64 KB
constexpr int kUpMaxFactors = 24;  // 2^12 needs 6 radix-4 stages; primes fewer
constexpr int kUpThreads = 256;

struct UpFactors {
    int n, count;
    int r[kUpMaxFactors];
};

// w[q] = e^(-2 pi i q / n), q < n, in FP64.
__global__ void up_roots_kernel(int n, double2 *__restrict__ w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    double s, c;
    sincospi(2.0 * (double)q / (double)n, &s, &c);
    w[q] = make_double2(c, -s);
}

// One transform of length f.n in LDS, src -> (returned) buffer; `inv`
// conjugates the roots.  Ends with a barrier.
__device__ double2 *up_fft(const UpFactors &f, const double2 *__restrict__ roots, bool inv,
                           double2 *src, double2 *dst)
{
    const int n = f.n;
    int ns = 1;
    for (int s = 0; s < f.count; ++s) {
        const int R = f.r[s];
        const int nr = n / R;
        const int step = n / (ns * R);
        for (int y = threadIdx.x; y < n; y += blockDim.x) {
            const int jm = y % ns;
            const int rem = y / ns;
            const int k = rem % R;
            const int j = (rem / R) * ns + jm;
            const int e = jm * step + k * nr;  // exponent of term r is r * e
            double re = 0.0, im = 0.0;
            int ex = 0;
            for (int r = 0; r < R; ++r) {
                const double2 x = src[j + r * nr];
                const double2 w = roots[ex];
                const double wi = inv ? -w.y : w.y;
                re = fma(x.x, w.x, fma(-x.y, wi, re));
                im = fma(x.x, wi, fma(x.y, w.x, im));
                ex += e;
                if (ex >= n)
                    ex -= n;
            }
            dst[y] = make_double2(re, im);
        }
        __syncthreads();
        double2 *t = src;
        src = dst;
        dst = t;
        ns *= R;
    }
    return src;
}

// Line geometry of one upsampling pass.
//   along = 0: azimuth lines.  Line l = row t (l / 8) and element pair l % 8;
//     sample p of the input at cell (p, t) of a grid with n_zen1 rows; output
//     sample q at cell (q, t) of a grid with n_zen1 rows.
//   along = 1: zenith lines.  Line l = azimuth p (l / 8); input sample s at
//     row (s <= nz_in ? s : 2 nz_in - s) of a grid with nz_in + 1 rows,
//     n_in = 2 nz_in; output rows q <= nz_out of a grid with nz_out + 1 rows,
//     n_out = 2 nz_out.
struct UpLines {
    int along;
    int n_in, n_out;
    int nbuf;           // LDS buffer length, max(n_in, n_out)
    int n_zen1;         // azimuth lines: rows of both grids
    int nz_in, nz_out;  // zenith lines
};

__device__ __forceinline__ size_t up_src_index(const UpLines &g, int line, int s)
{
    const int o = line >> 3;
    if (g.along == 0)
        return ((size_t)s * g.n_zen1 + o) * 16;
    const int row = s <= g.nz_in ? s : 2 * g.nz_in - s;
    return ((size_t)o * (g.nz_in + 1) + row) * 16;
}

__device__ __forceinline__ size_t up_dst_index(const UpLines &g, int line, int q)
{
    const int o = line >> 3;
    if (g.along == 0)
        return ((size_t)q * g.n_zen1 + o) * 16;
    return ((size_t)o * (g.nz_out + 1) + q) * 16;
}

__global__ __launch_bounds__(kUpThreads) void upsample_lines_kernel(
    UpLines g, UpFactors fin, UpFactors fout, const double2 *__restrict__ roots_in,
    const double2 *__restrict__ roots_out, const double *__restrict__ src,
    double *__restrict__ dst)
{
    extern __shared__ double2 up_lds[];
    double2 *b0 = up_lds;
    double2 *b1 = up_lds + g.nbuf;
    const int line = blockIdx.x;
    const int e0 = 2 * (line & 7);
    for (int s = threadIdx.x; s < g.n_in; s += blockDim.x) {
        const double *c = src + up_src_index(g, line, s) + e0;
        b0[s] = make_double2(c[0], c[1]);
    }
    __syncthreads();
    double2 *x = up_fft(fin, roots_in, false, b0, b1);
    double2 *y = x == b0 ? b1 : b0;
    // padded spectrum, scaled by 1/n_in
    const int n_in = g.n_in, n_out = g.n_out;
    const double scale = 1.0 / (double)n_in;
    const int pos = (n_in + 1) / 2;  // 0 .. pos-1 keep their index
    const int neg = (n_in - 1) / 2;  // n_in-neg .. n_in-1 move to the top
    const bool even = (n_in & 1) == 0;
    for (int k = threadIdx.x; k < n_out; k += blockDim.x) {
        double2 v = make_double2(0.0, 0.0);
        if (k < pos) {
            v = x[k];
        } else if (k >= n_out - neg) {
            v = x[k - n_out + n_in];
        } else if (even && (k == n_in / 2 || k == n_out - n_in / 2)) {
            v = x[n_in / 2];
            v.x *= 0.5;
            v.y *= 0.5;
        }
        y[k] = make_double2(v.x * scale, v.y * scale);
    }
    __syncthreads();
    double2 *r = up_fft(fout, roots_out, true, y, x);
    const int n_keep = g.along == 0 ? n_out : g.nz_out + 1;
    for (int q = threadIdx.x; q < n_keep; q += blockDim.x) {
        double *c = dst + up_dst_index(g, line, q) + e0;
        const double2 v = r[q];
        c[0] = v.x;
        c[1] = v.y;
    }
}

}  // namespace mbs
