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
//      one edge line 0; the others are integrated exactly: the polygon
//      clipped to each quadrant of the support, and the bilinear tent
//      integrated over the clipped polygon from its moments;
//   2. deapodisation: each raster value divided by the transform of the
//      interpolation kernel at its position (below);
//   3. the 2-D FFT (the row and column passes of the aperture FFT,
//      fft_kernels.hip), N = 2 x the padded extent (oversampling 2);
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
    SP_STRIDE = SP_EDGE + 4 * kSpecMaxNv
};

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
// rule on 1024 nodes over [0, W/2] (the integrand is even).
__global__ void spectrum_deap_kernel(int n, double *__restrict__ deap)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    const double s = ((double)j - 0.5 * n + 0.5) / n;
    const int q = 1024;
    const double du = 0.5 * kSpecW / q;
    double acc = 0.0;
    for (int i = 0; i < q; ++i) {
        const double u = (i + 0.5) * du;
        const double z = 2.0 * u / kSpecW;
        acc += exp(kSpecBeta * (sqrt(1.0 - z * z) - 1.0)) * cospi(2.0 * u * s);
    }
    deap[j] = 1.0 / (2.0 * du * acc);
}

// Sutherland-Hodgman clip of a polygon (counter-clockwise, local
// coordinates) to u >= lo (axis 0: u, axis 1: v; dir +1) or u <= hi
// (dir -1).  Returns the new vertex count.
__device__ __forceinline__ int spec_clip(const double *iu, const double *iv, int n, double *ou,
                                         double *ov, int axis, double lim, double dir)
{
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const int j = i + 1 < n ? i + 1 : 0;
        const double pi_ = axis ? iv[i] : iu[i], pj = axis ? iv[j] : iu[j];
        const bool ini = dir * (pi_ - lim) >= 0.0, inj = dir * (pj - lim) >= 0.0;
        if (ini) {
            ou[m] = iu[i];
            ov[m] = iv[i];
            ++m;
        }
        if (ini != inj) {
            const double t = (lim - pi_) / (pj - pi_);
            ou[m] = __builtin_fma(t, iu[j] - iu[i], iu[i]);
            ov[m] = __builtin_fma(t, iv[j] - iv[i], iv[i]);
            ++m;
        }
    }
    return m;
}

// Tent-weighted integral of the polygon over the support of the pixel at
// (px, py), over h^2 (1 for a pixel whose support is inside): the polygon
// clipped to each quadrant of [-h, h]^2 around the pixel, the weight
// (h - qx u)(h - qy v) / h^2 integrated from the clipped polygon's moments
// (area, first moments and the uv moment, by Green's theorem).
__device__ double spec_tent_exact(const double *__restrict__ r, double px, double py, double h)
{
    const int nv = (int)r[SP_NV];
    double bu[2][kSpecMaxNv + 8], bv[2][kSpecMaxNv + 8];
    double tot = 0.0;
    for (int q = 0; q < 4; ++q) {
        const double qx = (q & 1) ? 1.0 : -1.0, qy = (q & 2) ? 1.0 : -1.0;
        for (int e = 0; e < nv; ++e) {
            bu[0][e] = r[SP_VX + e] - px;
            bv[0][e] = r[SP_VY + e] - py;
        }
        int n = nv;
        // quadrant box: u between 0 and qx h, v between 0 and qy h
        n = spec_clip(bu[0], bv[0], n, bu[1], bv[1], 0, 0.0, qx);
        if (n)
            n = spec_clip(bu[1], bv[1], n, bu[0], bv[0], 0, qx * h, -qx);
        if (n)
            n = spec_clip(bu[0], bv[0], n, bu[1], bv[1], 1, 0.0, qy);
        if (n)
            n = spec_clip(bu[1], bv[1], n, bu[0], bv[0], 1, qy * h, -qy);
        if (n < 3)
            continue;
        double a = 0.0, mu = 0.0, mv = 0.0, muv = 0.0;
        for (int i = 0; i < n; ++i) {
            const int j = i + 1 < n ? i + 1 : 0;
            const double u0 = bu[0][i], v0 = bv[0][i], u1 = bu[0][j], v1 = bv[0][j];
            const double c = u0 * v1 - u1 * v0;
            a += c;
            mu += (u0 + u1) * c;
            mv += (v0 + v1) * c;
            muv += (u0 * v1 + 2.0 * u0 * v0 + 2.0 * u1 * v1 + u1 * v0) * c;
        }
        a *= 0.5;
        mu *= 1.0 / 6.0;
        mv *= 1.0 / 6.0;
        muv *= 1.0 / 24.0;
        tot += h * h * a - h * qx * mu - h * qy * mv + qx * qy * muv;
    }
    return tot / (h * h * h * h);
}

// Raster value of pixel (px, py): 1 where the tent's support lies inside
// every edge line, 0 where it lies outside one, else the exact tent
// integral.
__device__ __forceinline__ double spec_pixel(const double *__restrict__ r, double px, double py)
{
    const int nv = (int)r[SP_NV];
    bool inside = true;
    for (int e = 0; e < nv; ++e) {
        const double *l = r + SP_EDGE + 4 * e;
        const double v = __builtin_fma(l[0], px, __builtin_fma(l[1], py, l[2]));
        if (v <= -l[3])
            return 0.0;  // the whole support is outside this edge's line
        inside = inside && v > l[3];
    }
    return inside ? 1.0 : spec_tent_exact(r, px, py, r[SP_H]);
}

// Raster + deapodisation + row FFTs of the spectrum beams of one group (all
// of grid size 2^LOGN): as fft_rows_kernel, with the tent raster.
template <int LOGN>
__global__ void __launch_bounds__(FFT_ROW_THREADS) spectrum_rows_kernel(
    int n_sb, const double *__restrict__ rec, const double *__restrict__ deap,
    const float2 *__restrict__ tw, float2 *__restrict__ grid)
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
    float2 *g = grid + ((size_t)beam * N + row0) * N;
    auto raster = [&](int seq, int i) {
        const int row = row0 + seq;
        const double px = fft_pixel(i, N, h, xc), py = fft_pixel(row, N, h, yc);
        const double v = spec_pixel(r, px, py);
        return make_float2((float)(v * deap[row] * deap[i]), 0.0f);
    };
    auto out = [&](int seq, int i, float2 x) { g[(size_t)seq * N + i] = x; };
    fft_stages<LOGN, FFT_ROWS, FFT_ROW_THREADS, false, 0>(lds, fft_ld(N), tw, raster, out);
}

// e^{i pi (n - 1) j / n}, the phase that centres bin j of an n-point
// transform of pixel-centred samples, with the angle reduced exactly in
// integers: (n - 1) j mod 2n.
__device__ __forceinline__ float2 spec_centre_phase(int j, int n)
{
    const long long m = (((long long)(n - 1) * j) % (2LL * n) + 2LL * n) % (2LL * n);
    float s, c;
    sincospif((float)((double)m / n), &s, &c);
    return make_float2(c, s);
}

// F of spectrum beam i at every cell: thread per (i, cell).  spec[i] is the
// offset of beam i's spectrum, logn[i] its grid size.  fbuf: [n_sb][cells].
__global__ void __launch_bounds__(256) spectrum_sample_kernel(
    int n_sb, int cells, const int32_t *__restrict__ beam, const int64_t *__restrict__ spec,
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
    const double d0 = tab[(size_t)(DT_DIR + 0) * cells + c];
    const double d1 = tab[(size_t)(DT_DIR + 1) * cells + c];
    const double d2 = tab[(size_t)(DT_DIR + 2) * cells + c];
    const double A = h[H_A0] - (d0 * h[H_HOR] + d1 * h[H_HOR + 1] + d2 * h[H_HOR + 2]);
    const double B = h[H_B0] - (d0 * h[H_VER] + d1 * h[H_VER + 1] + d2 * h[H_VER + 2]);
    const double psi = d0 * h[H_KC] + d1 * h[H_KC + 1] + d2 * h[H_KC + 2];
    const double px_ = r[SP_H];
    const double inv2pi = 0.15915494309189533577;
    const double fx = -k * A * inv2pi, fy = -k * B * inv2pi;
    // positions in bins; the W taps nearest each
    const double ux = fx * n * px_, uy = fy * n * px_;
    const int jx0 = (int)ceil(ux - 0.5 * kSpecW), jy0 = (int)ceil(uy - 0.5 * kSpecW);
    // kernel weights in FP32 (relative error ~1e-7, below the spectrum's)
    float2 wx[kSpecW], wy[kSpecW];
    for (int t = 0; t < kSpecW; ++t) {
        const float zx = (float)(2.0 * (ux - (jx0 + t)) / kSpecW);
        const float zy = (float)(2.0 * (uy - (jy0 + t)) / kSpecW);
        const float beta = (float)kSpecBeta;
        const float ex = fabsf(zx) < 1.0f ? expf(beta * (sqrtf(1.0f - zx * zx) - 1.0f)) : 0.0f;
        const float ey = fabsf(zy) < 1.0f ? expf(beta * (sqrtf(1.0f - zy * zy) - 1.0f)) : 0.0f;
        const float2 px = spec_centre_phase(jx0 + t, n), py = spec_centre_phase(jy0 + t, n);
        wx[t] = make_float2(ex * px.x, ex * px.y);
        wy[t] = make_float2(ey * py.x, ey * py.y);
    }
    const float2 *g = spectra + spec[i];
    float sr = 0.0f, si = 0.0f;
    for (int ty = 0; ty < kSpecW; ++ty) {
        const int iy = (((jy0 + ty) % n) + n) % n;
        float rr = 0.0f, ri = 0.0f;
        for (int tx = 0; tx < kSpecW; ++tx) {
            const int ix = (((jx0 + tx) % n) + n) % n;
            const float2 z = g[(size_t)iy * n + ix];
            rr = __builtin_fmaf(z.x, wx[tx].x, __builtin_fmaf(-z.y, wx[tx].y, rr));
            ri = __builtin_fmaf(z.x, wx[tx].y, __builtin_fmaf(z.y, wx[tx].x, ri));
        }
        sr = __builtin_fmaf(rr, wy[ty].x, __builtin_fmaf(-ri, wy[ty].y, sr));
        si = __builtin_fmaf(rr, wy[ty].y, __builtin_fmaf(ri, wy[ty].x, si));
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
