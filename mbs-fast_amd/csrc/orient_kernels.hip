// orient_kernels.hip -- orientation sets of the orientation average, on the
// device: one lane per orientation.
//
//   MBS_ORIENT_SOBOL       Owen-scrambled 2-D Sobol points (src/math/Sobol.cpp:
//                          1-106) mapped to beta = acos(1 - (1 - cos beta_sym) u),
//                          gamma = gamma_sym v, weight 1/N
//                          (src/scattering/TracerPOTotal.cpp:5465-5492)
//   MBS_ORIENT_HAMMERSLEY  symmetry-quotient Hammersley points: u = (i + 1/2)/N,
//                          v = radical inverse base 2 of i
//                          (TracerPOTotal.cpp:5347-5426, without the paired
//                          mirror audit)
//   MBS_ORIENT_SHOEMAKE    full-SO(3) Hammersley quaternions, Shoemake's map of
//                          (u1, u2, u3) = ((i+1/2)/N, RI2(i), RI3(i))
//                          (TracerPOTotal.cpp:5428-5463)
// Each row: beta, gamma (NaN for quaternion sets), weight, qx, qy, qz, qw, and
// the particle rotation matrix R (row-major), built as Particle::
// SetRotateMatrix(beta, gamma, alpha = 0) or SetRotateMatrixFromQuaternion
// (src/particle/Particle.cpp:1593-1642).  For the Euler sets the quaternion is
// that of R (Shepperd's method, qw >= 0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

enum { OR_BETA = 0, OR_GAMMA = 1, OR_W = 2, OR_Q = 3, OR_R = 7, OR_STRIDE = 16 };

// Hash of one node of the Owen scrambling tree: a seed/dimension/level/prefix
// mix followed by a 32-bit avalanche finaliser; the constants are those of
// Sobol2D::hash (src/math/Sobol.cpp:19-33), which fix the point set.
__device__ __forceinline__ uint32_t owen_node_hash(uint32_t seed, uint32_t dim, uint32_t level,
                                                   uint32_t prefix)
{
    uint32_t h = seed ^ 0x9e3779b9u;
    h ^= dim + 0x85ebca6bu + (h << 6) + (h >> 2);
    h ^= level + 0xc2b2ae35u + (h << 6) + (h >> 2);
    h ^= prefix + 0x27d4eb2fu + (h << 6) + (h >> 2);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Base-2 nested Owen scrambling: bit b (from the most significant) is flipped
// by the low hash bit of the node addressed by the b preceding original bits.
__device__ __forceinline__ uint32_t owen_scramble(uint32_t value, uint32_t dim, uint32_t seed)
{
    uint32_t out = 0, prefix = 0;
    for (uint32_t level = 0; level < 32; ++level) {
        const uint32_t shift = 31u - level;
        const uint32_t bit = (value >> shift) & 1u;
        out |= (bit ^ (owen_node_hash(seed, dim, level, prefix) & 1u)) << shift;
        prefix = (prefix << 1) | bit;
    }
    return out;
}

// Unscrambled 2-D Sobol point number n >= 1 of the Gray-code sequence: the
// XOR of the direction numbers selected by the bits of gray(n).  Dimension 0
// is the van der Corput sequence (v_j = 2^(31-j)); dimension 1 has the
// degree-1 primitive polynomial x + 1 (v_0 = 2^31, v_j = v_{j-1} ^ v_{j-1}>>1).
__device__ __forceinline__ void sobol_point(uint32_t n, uint32_t &x0, uint32_t &x1)
{
    const uint32_t g = n ^ (n >> 1);
    uint32_t a = 0, b = 0, v = 1u << 31;
    for (int j = 0; j < 32; ++j) {
        if ((g >> j) & 1u) {
            a ^= 1u << (31 - j);
            b ^= v;
        }
        v ^= v >> 1;
    }
    x0 = a;
    x1 = b;
}

__device__ __forceinline__ double radical_inverse2(uint32_t v)
{
    return (double)__builtin_bitreverse32(v) / 4294967296.0;
}

__device__ __forceinline__ double radical_inverse3(uint32_t v)
{
    const double inv_base = 1.0 / 3.0;
    double inv = inv_base, r = 0.0;
    while (v > 0) {
        r += (double)(v % 3u) * inv;
        v /= 3u;
        inv *= inv_base;
    }
    return r;
}

// Rotation matrix of Euler angles (beta, gamma, alpha = 0).
__device__ __forceinline__ void euler_matrix(double beta, double gamma, double *R)
{
    double sB, cB, sG, cG;
    sincos(beta, &sB, &cB);
    sincos(gamma, &sG, &cG);
    const double cA = 1.0, sA = 0.0;
    const double cAcB = cA * cB, sAcG = sA * cG, sAsG = sA * sG;
    R[0] = cAcB * cG - sAsG;
    R[1] = -(cAcB * sG + sAcG);
    R[2] = cA * sB;
    R[3] = sAcG * cB + cA * sG;
    R[4] = cA * cG - sAsG * cB;
    R[5] = sA * sB;
    R[6] = -sB * cG;
    R[7] = sB * sG;
    R[8] = cB;
}

__device__ __forceinline__ void quat_matrix(double qx, double qy, double qz, double qw, double *R)
{
    const double xx = qx * qx, yy = qy * qy, zz = qz * qz;
    const double xy = qx * qy, xz = qx * qz, yz = qy * qz;
    const double wx = qw * qx, wy = qw * qy, wz = qw * qz;
    R[0] = 1.0 - 2.0 * (yy + zz);
    R[1] = 2.0 * (xy - wz);
    R[2] = 2.0 * (xz + wy);
    R[3] = 2.0 * (xy + wz);
    R[4] = 1.0 - 2.0 * (xx + zz);
    R[5] = 2.0 * (yz - wx);
    R[6] = 2.0 * (xz - wy);
    R[7] = 2.0 * (yz + wx);
    R[8] = 1.0 - 2.0 * (xx + yy);
}

// Quaternion of a rotation matrix (Shepperd: pivot on the largest of
// trace, R00, R11, R22), sign fixed so that qw >= 0.
__device__ __forceinline__ void matrix_quat(const double *R, double *q)
{
    const double t = R[0] + R[4] + R[8];
    double x, y, z, w;
    if (t >= R[0] && t >= R[4] && t >= R[8]) {
        const double s = 2.0 * sqrt(fmax(0.0, 1.0 + t));
        w = 0.25 * s;
        x = (R[7] - R[5]) / s;
        y = (R[2] - R[6]) / s;
        z = (R[3] - R[1]) / s;
    } else if (R[0] >= R[4] && R[0] >= R[8]) {
        const double s = 2.0 * sqrt(fmax(0.0, 1.0 + R[0] - R[4] - R[8]));
        w = (R[7] - R[5]) / s;
        x = 0.25 * s;
        y = (R[1] + R[3]) / s;
        z = (R[2] + R[6]) / s;
    } else if (R[4] >= R[8]) {
        const double s = 2.0 * sqrt(fmax(0.0, 1.0 + R[4] - R[0] - R[8]));
        w = (R[2] - R[6]) / s;
        x = (R[1] + R[3]) / s;
        y = 0.25 * s;
        z = (R[5] + R[7]) / s;
    } else {
        const double s = 2.0 * sqrt(fmax(0.0, 1.0 + R[8] - R[0] - R[4]));
        w = (R[3] - R[1]) / s;
        x = (R[2] + R[6]) / s;
        y = (R[5] + R[7]) / s;
        z = 0.25 * s;
    }
    const double sg = w < 0.0 ? -1.0 : 1.0;
    q[0] = sg * x;
    q[1] = sg * y;
    q[2] = sg * z;
    q[3] = sg * w;
}

// kind: 0 Sobol, 1 Hammersley, 2 Shoemake.  cos_beta_sym is computed by the
// host (the reference evaluates it once on the host as well).
__global__ void __launch_bounds__(256) orientation_set_kernel(
    int kind, int n, uint32_t seed, double cos_beta_sym, double gamma_sym, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double *o = out + (size_t)i * OR_STRIDE;
    const double w = 1.0 / (double)n;
    double R[9], q[4];
    if (kind == 2) {
        const double u1 = ((double)i + 0.5) / (double)n;
        const double u2 = radical_inverse2((uint32_t)i);
        const double u3 = radical_inverse3((uint32_t)i);
        const double r1 = sqrt(fmax(0.0, 1.0 - u1));
        const double r2 = sqrt(u1);
        const double a = 2.0 * 3.141592653589793116 * u2;
        const double b = 2.0 * 3.141592653589793116 * u3;
        double sa, ca, sb, cb;
        sincos(a, &sa, &ca);
        sincos(b, &sb, &cb);
        q[0] = r1 * sa;
        q[1] = r1 * ca;
        q[2] = r2 * sb;
        q[3] = r2 * cb;
        const double nrm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        quat_matrix(q[0] / nrm, q[1] / nrm, q[2] / nrm, q[3] / nrm, R);
        o[OR_BETA] = __builtin_nan("");
        o[OR_GAMMA] = __builtin_nan("");
    } else {
        double u, v;
        if (kind == 0) {
            uint32_t x0, x1;
            sobol_point((uint32_t)i + 1u, x0, x1);
            u = (double)owen_scramble(x0, 0, seed) / 4294967296.0;
            v = (double)owen_scramble(x1, 1, seed) / 4294967296.0;
        } else {
            u = ((double)i + 0.5) / (double)n;
            v = radical_inverse2((uint32_t)i);
        }
        const double beta = acos(1.0 - (1.0 - cos_beta_sym) * u);
        const double gamma = gamma_sym * v;
        euler_matrix(beta, gamma, R);
        matrix_quat(R, q);
        o[OR_BETA] = beta;
        o[OR_GAMMA] = gamma;
    }
    o[OR_W] = w;
    for (int k = 0; k < 4; ++k)
        o[OR_Q + k] = q[k];
    for (int k = 0; k < 9; ++k)
        o[OR_R + k] = R[k];
}

}  // namespace mbs
