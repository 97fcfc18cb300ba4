// po_math.h -- device math for the PO diffraction kernels (gfx950).
//
// Phasors e^{i x} are evaluated here rather than through the device library
// so that the hot loop's cost and accuracy are known: a table-driven
// reduction against a 256-entry LDS table of e^{2 pi i j / 256} and short
// Taylor polynomials (see sincos_table below).  Phases reach ~1e4 rad for the
// largest particles; the reduction is exact to an ulp for |x| < 2^20.
// The FP32 variant used by the mixed-precision kernels reduces in FP64 the
// same way against a 512-entry table and needs only first- and second-order
// corrections in FP32 (error ~1 ulp of FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

constexpr double kPi = 3.141592653589793116;            // double(pi) = M_PI
// Small-phase switch: every vertex phase within 1e-3 rad
// (HandlerPO_fast.h:161-216).
constexpr double kSmallPhase = 1e-3;
// The mixed-precision kernels' switch.  Their edge sum differences FP32
// phasors (error ~1e-7 each) whose sum cancels to O(phase^2) of a single
// term, so its relative error grows like 1e-7 / phase^2 -- 1e-1 at 1e-3 rad.
// They take the moment form up to 0.25 rad instead, with polygon moments to
// fifth order (truncation below max|phase|^6 / 6! = 3.4e-7 there), where the
// edge sum's error has fallen to ~1e-5.
constexpr double kSmallPhaseMixed = 0.25;

// Slope stored for an edge that is not valid in that direction (|dx| or
// |dy| below the reference's geometric tolerance).  Its coefficient
// A + s B (or s A + B) is then at least 1e300 times the lane's larger
// projection, so the mixed kernels' FP32 reciprocal of it is exactly zero
// and the edge drops out of the sum with no per-lane validity test; only a
// lane with both projections below 1e-300 (every vertex phase tiny, so it
// takes the moment form) sees a small coefficient.  The FP64 kernels mask
// invalid edges by their flags and never use the value.
constexpr double kInvalidSlope = 1e300;
// One Horner step p*z + c as a single three-operand v_fma_f64 with the
// coefficient in an SGPR pair.  Written as inline assembly because the
// compiler otherwise keeps loop-invariant FP64 coefficients in VGPRs and
// emits v_mov_b64 + v_fmac_f64 (two instructions) for every step.
__device__ __forceinline__ double horner(double p, double z, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(p), "v"(z), "s"(c));
    return r;
}

// ---------------------------------------------------------------------------
// Table-driven phasor e^{i x}: x = j (2 pi/256) + d with j = rint(x 256/2pi),
// |d| <= pi/256.  e^{i j 2pi/256} comes from a 256-entry table in LDS (filled
// once per workgroup from the device library's sincos), e^{i d} from Taylor
// polynomials (sin to d^7, cos to d^6, truncation < 1e-20), and one complex
// product joins them.  The reduction is exact for |j| < 2^20 with
// kTabStep1 = double(pi)/128 (an exact scaling of double(pi)): both x and
// j*kTabStep1 are multiples of 2^-58 and |x - j*kTabStep1| < 2^-6.
// ---------------------------------------------------------------------------
constexpr int kTabN = 256;
constexpr double kTabInvStep = 40.743665431525205956;  // 256 / (2 pi)
constexpr double kTabStep1 = 3.141592653589793116 / 128.0;
constexpr double kTabStep2 = 1.2246467991473532e-16 / 128.0;

// j = rint(x 256/2pi) by the addition of 1.5 2^52, whose sum carries j in its
// low 32 bits (two's complement) and is j + 1.5 2^52 exactly for |j| < 2^51:
// one FMA for the product and the rounding, one subtraction for j as a double.
constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52
__device__ __forceinline__ uint32_t low_bits(double t)
{
    return (uint32_t)__builtin_bit_cast(uint64_t, t);
}

// In two halves, as for the FP32 variant below: reduction and table read,
// then polynomials and product.
__device__ __forceinline__ void sincos_reduce(double x, const double2 *__restrict__ tab,
                                              double &d, double2 &e)
{
    const double t = __builtin_fma(x, kTabInvStep, kRoundMagic);
    const double j = t - kRoundMagic;
    d = __builtin_fma(-j, kTabStep1, x);
    d = __builtin_fma(-j, kTabStep2, d);
    e = tab[low_bits(t) & (kTabN - 1)];
}

__device__ __forceinline__ void sincos_finish(double d, double2 e, double &s, double &c)
{
    const double z = d * d;
    double ps = __builtin_fma(z, -1.0 / 5040.0, 1.0 / 120.0);
    ps = horner(ps, z, -1.0 / 6.0);
    const double sd = __builtin_fma(d * z, ps, d);
    double pc = __builtin_fma(z, -1.0 / 720.0, 1.0 / 24.0);
    pc = horner(pc, z, -0.5);
    const double cd = __builtin_fma(pc, z, 1.0);
    s = __builtin_fma(e.y, cd, e.x * sd);
    c = __builtin_fma(e.x, cd, -(e.y * sd));
}

__device__ __forceinline__ void sincos_table(double x, const double2 *__restrict__ tab,
                                             double &s, double &c)
{
    double d;
    double2 e;
    sincos_reduce(x, tab, d, e);
    sincos_finish(d, e, s, c);
}

// FP32 result from an FP64 argument against a 512-entry FP32 table of
// e^{2 pi i j / 512}, in two halves so that callers can issue the table
// reads of several arguments before using any of them: the FP64 reduction
// and table read, then the FP32 correction and the product with the table
// entry.  One FMA reduces: kTab32Step = double(pi)/256 differs from 2 pi/512
// by 4.8e-19, so x - j kTab32Step is off by j 4.8e-19 = |x| 3.9e-17 at most,
// below half an ulp of x itself (the FP64 kernels add the second term of
// the split; an FP32 result cannot see it).  Against the 256-entry table the
// index costs one more instruction (a mask and a shift instead of one shift
// of the low byte) and the correction two fewer.
//
// An entry holds (cos, sin, cos/2, sin/2) of its angle (16 bytes, one LDS
// read): with the halves at hand the correction and the product are four
// FMAs (see sincos_finish_f32).
constexpr int kTabN32 = 512;
constexpr double kTab32InvStep = 81.487330863050411912;  // 512 / (2 pi)
constexpr double kTab32Step = 3.141592653589793116 / 256.0;

// Entry j of the FP32 table.
__device__ __forceinline__ float4 tab32_entry(int j)
{
    double ts, tc;
    sincospi((double)j / (kTabN32 / 2), &ts, &tc);
    return make_float4((float)tc, (float)ts, (float)(0.5 * tc), (float)(0.5 * ts));
}

__device__ __forceinline__ void sincos_reduce_f32(double x, const float4 *__restrict__ tab,
                                                  float &d, float4 &e)
{
    const double t = __builtin_fma(x, kTab32InvStep, kRoundMagic);
    const double j = t - kRoundMagic;
    d = (float)__builtin_fma(-j, kTab32Step, x);
    e = tab[low_bits(t) & (kTabN32 - 1)];
}

// |d| <= pi/512 = 6.14e-3, so e^{i d} = (1 - d^2/2) + i d to FP32 accuracy:
// the dropped terms are below 3.9e-8 (d^3/6) and 6e-11 (d^4/24), within an
// ulp of FP32 near 1 (6e-8) -- the phasor rounding the mixed kernels carry
// anyway.  With T = (c, s) the entry,
//   Im T e^{id} = s + d (c - d s/2),   Re T e^{id} = c - d (s + d c/2),
// two FMAs each on the stored halves.
__device__ __forceinline__ void sincos_finish_f32(float d, float4 e, float &s, float &c)
{
    s = __builtin_fmaf(d, __builtin_fmaf(-d, e.w, e.x), e.y);
    c = __builtin_fmaf(-d, __builtin_fmaf(d, e.z, e.y), e.x);
}

__device__ __forceinline__ void sincos_table_f32(double x, const float4 *__restrict__ tab,
                                                 float &s, float &c)
{
    float d;
    float4 e;
    sincos_reduce_f32(x, tab, d, e);
    sincos_finish_f32(d, e, s, c);
}

// 1/x to full FP64 precision: hardware estimate + two Newton steps.
__device__ __forceinline__ double rcp_f64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    return r;
}

// 1/x from the hardware estimate and one Newton step: relative error below
// (estimate error)^2, about 1e-14 -- used for the edge coefficients, whose
// quotients carry the rounding of ~1e2 rad phase arguments anyway.
__device__ __forceinline__ double rcp_f64_fast(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}

__device__ __forceinline__ float rcp_f32(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}

// sin(x)/x from its Taylor series through x^10; for |x| <= 0.25 the first
// omitted term, x^12/13!, is below 4e-18.
__device__ __forceinline__ double sinc_series(double x)
{
    const double z = x * x;
    double p = __builtin_fma(z, -1.0 / 39916800.0, 1.0 / 362880.0);
    p = __builtin_fma(z, p, -1.0 / 5040.0);
    p = __builtin_fma(z, p, 1.0 / 120.0);
    p = __builtin_fma(z, p, -1.0 / 6.0);
    return __builtin_fma(z, p, 1.0);
}

// add_stable_edge_quotient (src/handler/HandlerPO_fast.h:235-264): the exact
// half-angle form of exp(i phaseStart)(exp(i k c delta) - 1)/c for an edge
// coefficient c at or near zero.  Returns the increment (real part in .x,
// imaginary in .y) so the running sums stay in registers; only lanes at a
// near-singular edge execute it.  The reference evaluates sin(x)/x by a
// short series below |x| = 1e-4 and by the library sine above; here the
// series through x^10 covers |x| <= 0.25 (z = k c delta is below ~1e-4 in
// FP64 and ~0.02 in mixed precision for 50 um beams), and beyond it the plain
// quotient of the two vertex phasors (pr, pi) -> (nr, ni), whose relative
// rounding there is below 1e-12, needs no library call.
__device__ __forceinline__ double2 stable_edge_quotient(double pr, double pi, double nr,
                                                        double ni, double delta, double cc,
                                                        double k)
{
    const double z = k * cc * delta;
    if (__builtin_fabs(z) <= 0.25) {
        const double sh = sinc_series(0.5 * z);
        const double lr = -0.5 * k * k * cc * delta * delta * sh * sh;
        const double li = k * delta * sinc_series(z);
        return make_double2(pr * lr - pi * li, pr * li + pi * lr);
    }
    const double r = rcp_f64(cc);
    return make_double2((nr - pr) * r, (ni - pi) * r);
}

// AddMuellerFromJones (src/math/Mueller.hpp:37-87): out += w * Mueller(J),
// J given as 8 reals (00r 00i 01r 01i 10r 10i 11r 11i).
template <typename T>
__device__ __forceinline__ void add_mueller(double *m, const T *j, double w)
{
    const double j00r = j[0], j00i = j[1], j01r = j[2], j01i = j[3];
    const double j10r = j[4], j10i = j[5], j11r = j[6], j11i = j[7];
    const double a11 = j00r * j00r + j00i * j00i, a12 = j01r * j01r + j01i * j01i;
    const double a21 = j10r * j10r + j10i * j10i, a22 = j11r * j11r + j11i * j11i;
    double A1 = a11 + a21, A2 = a12 + a22;
    m[0] += ((A1 + A2) * 0.5) * w;
    m[1] += ((A1 - A2) * 0.5) * w;
    A1 = a11 - a21;
    A2 = a12 - a22;
    m[4] += ((A1 + A2) * 0.5) * w;
    m[5] += ((A1 - A2) * 0.5) * w;
    double c1r = j00r * j01r + j00i * j01i, c1i = j00i * j01r - j00r * j01i;
    double c2r = j11r * j10r + j11i * j10i, c2i = j11i * j10r - j11r * j10i;
    m[2] += (-c1r - c2r) * w;
    m[3] += (c2i - c1i) * w;
    m[6] += (c2r - c1r) * w;
    m[7] += (-c1i - c2i) * w;
    c1r = j00r * j10r + j00i * j10i;
    c1i = j00i * j10r - j00r * j10i;
    c2r = j11r * j01r + j11i * j01i;
    c2i = j11i * j01r - j11r * j01i;
    m[8] += (-c1r - c2r) * w;
    m[9] += (c2r - c1r) * w;
    m[12] += (c1i - c2i) * w;
    m[13] += (c2i + c1i) * w;
    c1r = j00r * j11r + j00i * j11i;
    c1i = j00i * j11r - j00r * j11i;
    c2r = j01r * j10r + j01i * j10i;
    c2i = j01i * j10r - j01r * j10i;
    m[10] += (c1r + c2r) * w;
    m[11] += (c1i - c2i) * w;
    m[14] += (-c1i - c2i) * w;
    m[15] += (c1r - c2r) * w;
}

}  // namespace mbs
