/*
 * po_oracle.c -- FP64 CPU restatement of the reference's PO diffraction hot
 * path.  TEST INFRASTRUCTURE ONLY: it is the checker the HIP engine is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it; the engine never calls it and has no CPU fallback.
 *
 * Every function follows the reference's CPU double path line by line in its
 * arithmetic (file:line given per function), with one deliberate change: the
 * sincos of phases uses libm sincos() instead of the reference's 14-digit
 * polynomial (src/handler/HandlerPO_fast.h:11-146), so the oracle is at least
 * as accurate as the path it restates.
 *
 * Parity pins (tests/test_oracle.py): the reference fixtures written by
 * tests/golden/make_fixtures.py (fixed orientation, absorbing fixed
 * orientation, an Owen-Sobol orientation average, each including the
 * reference's own .dat output), plus closed forms: the Fourier transform of a
 * rectangle and a triangle aperture, brute-force 2-D quadrature of a random
 * polygon, |F| = area/lambda at the beam's forward direction, the Mueller
 * identities of a single Jones matrix, the Mueller matrix of a rotated
 * linear polariser, and sum(2*pi*dcos) = 4*pi.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mbs_hip.h"

#define ORACLE_MAXV 32

/* Phase sincos of the oracle.  Mode 0 (default): libm sincos, correctly
 * rounded for practical purposes.  Mode 1: the reference CPU's own
 * polynomial fast_sincos (src/handler/HandlerPO_fast.h:116-146: reduction by
 * round(x/pi) with a two-part pi, degree-13/12 polynomials, sign flip for odd
 * multiples), used by the reference for grid tables, vertex phases and beam
 * centre phases; its cosine is off by up to 6.3e-9, which the fixtures carry.
 * Mode 1 exists to pin the oracle's algorithm to the fixtures tightly. */
static int g_ref_sincos = 0;

void mbs_oracle_set_ref_sincos(int mode) { g_ref_sincos = mode; }

#ifdef _OPENMP
#include <omp.h>
#endif
/* Number of OpenMP threads the oracle's parallel loops use. */
int mbs_oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void ref_poly_sincos(double x, double *s, double *c)
{
    const double n = __builtin_round(x * 0.31830988618379067);
    const double r = x - n * 3.1415926535897931 - n * 1.2246467991473532e-16;
    const double r2 = r * r;
    const double S = r2 * (-0.16666666666666632 + r2 * (0.0083333333333309497 +
                     r2 * (-0.00019841269836761127 + r2 * (2.7557316103728802e-06 +
                     r2 * (-2.5052106798274584e-08 + r2 * 1.5896230157654656e-10)))));
    const double C = r2 * (-0.49999999999999994 + r2 * (0.041666666666665589 +
                     r2 * (-0.0013888888888611484 + r2 * (2.4801587288851704e-05 +
                     r2 * (-2.7557314351390663e-07 + r2 * 2.0875723212981748e-09)))));
    double sv = r + r * S, cv = 1.0 + C;
    if (((int)n) & 1) {
        sv = -sv;
        cv = -cv;
    }
    *s = sv;
    *c = cv;
}

static void phase_sincos(double x, double *s, double *c)
{
    if (g_ref_sincos)
        ref_poly_sincos(x, s, c);
    else
        sincos(x, s, c);
}

/* Small-phase moment expansion of the polygon integral.
 * Follows small_phase_polygon_factor, src/handler/HandlerPO_fast.h:154-216. */
int mbs_oracle_small_phase_factor(const double *vx, const double *vy, int nv,
                                  double A, double B, double k,
                                  double *fr, double *fi)
{
    if (nv < 3)
        return 0;
    const double qx = k * A, qy = k * B;
    double maxPhase = 0.0;
    for (int i = 0; i < nv; ++i) {
        const double p = fabs(qx * vx[i] + qy * vy[i]);
        if (p > maxPhase)
            maxPhase = p;
    }
    if (maxPhase > 1e-3)
        return 0;
    double twiceArea = 0, fx6 = 0, fy6 = 0, sx12 = 0, sy12 = 0, sxy24 = 0;
    double minX = DBL_MAX, maxX = -DBL_MAX, minY = DBL_MAX, maxY = -DBL_MAX;
    for (int i = 0; i < nv; ++i) {
        const int j = (i + 1 < nv) ? i + 1 : 0;
        const double xi = vx[i], yi = vy[i], xj = vx[j], yj = vy[j];
        const double cross = xi * yj - xj * yi;
        twiceArea += cross;
        fx6 += (xi + xj) * cross;
        fy6 += (yi + yj) * cross;
        sx12 += (xi * xi + xi * xj + xj * xj) * cross;
        sy12 += (yi * yi + yi * yj + yj * yj) * cross;
        sxy24 += (2.0 * xi * yi + xi * yj + xj * yi + 2.0 * xj * yj) * cross;
        minX = fmin(minX, xi); maxX = fmax(maxX, xi);
        minY = fmin(minY, yi); maxY = fmax(maxY, yi);
    }
    const double areaScale = fmax((maxX - minX) * (maxY - minY), DBL_MIN);
    if (fabs(twiceArea) <= 0x1p-40 * areaScale) {
        *fr = 1.0;
        *fi = 0.0;
        return 1;
    }
    const double invArea = 2.0 / twiceArea;
    const double mX = (fx6 / 6.0) * invArea, mY = (fy6 / 6.0) * invArea;
    const double mXX = (sx12 / 12.0) * invArea, mYY = (sy12 / 12.0) * invArea;
    const double mXY = (sxy24 / 24.0) * invArea;
    const double p1 = qx * mX + qy * mY;
    const double p2 = qx * qx * mXX + 2.0 * qx * qy * mXY + qy * qy * mYY;
    *fr = 1.0 - 0.5 * p2;
    *fi = p1;
    return 1;
}

/* stable_sinc, src/handler/HandlerPO_fast.h:217-226. */
static double stable_sinc(double x)
{
    if (fabs(x) < 1e-4) {
        const double x2 = x * x;
        return 1.0 - x2 / 6.0 + x2 * x2 / 120.0;
    }
    return sin(x) / x;
}

/* add_stable_edge_quotient, src/handler/HandlerPO_fast.h:235-264. */
static void add_stable_edge_quotient(double pr, double pi, double delta,
                                     double c, double k, double *sr, double *si)
{
    const double z = k * c * delta;
    const double sh = stable_sinc(0.5 * z);
    const double lr = -0.5 * k * k * c * delta * delta * sh * sh;
    const double li = k * delta * stable_sinc(z);
    *sr += pr * lr - pi * li;
    *si += pr * li + pi * lr;
}

/* Aperture integral F(beam, direction) for given A, B and vertex phasors.
 * Edge loop of HandleBeamsToLocal, src/handler/HandlerPO.cpp:2690-2738, with
 * the small-phase branch of :2680-2688.  Writes F (complex) to fr/fi. */
static void aperture_integral(const double *vx, const double *vy,
                              const double *syx, const double *sxy,
                              const uint8_t *ef, int nv, double A, double B,
                              const double *vc, const double *vs,
                              const mbs_consts *c, double area,
                              double *fr, double *fi)
{
    const double k = c->wave_index;
    double spr, spi;
    if (mbs_oracle_small_phase_factor(vx, vy, nv, A, B, k, &spr, &spi)) {
        /* areaLimit = (legacy ? invComplWave : -invComplWave) * area */
        const double sgn = c->legacy_sign ? 1.0 : -1.0;
        const double ar = sgn * c->inv_compl_wave[0] * area;
        const double ai = sgn * c->inv_compl_wave[1] * area;
        *fr = ar * spr - ai * spi;
        *fi = ar * spi + ai * spr;
        return;
    }
    double sr = 0, si = 0;
    if (fabs(B) > fabs(A)) {
        for (int e = 0; e < nv; ++e) {
            if (!(ef[e] & MBS_EDGE_VALID_X))
                continue;
            const int en = (e + 1 < nv) ? e + 1 : 0;
            const double Ci = A + syx[e] * B;
            const double inv = (fabs(Ci) > c->eps1) ? 1.0 / Ci : 0.0;
            sr += (vc[en] - vc[e]) * inv;
            si += (vs[en] - vs[e]) * inv;
            if (fabs(Ci) <= c->eps1)
                add_stable_edge_quotient(vc[e], vs[e], vx[en] - vx[e], Ci, k,
                                         &sr, &si);
        }
        const double invB = 1.0 / B;
        sr *= invB;
        si *= invB;
    } else {
        for (int e = 0; e < nv; ++e) {
            if (!(ef[e] & MBS_EDGE_VALID_Y))
                continue;
            const int en = (e + 1 < nv) ? e + 1 : 0;
            const double Ei = A * sxy[e] + B;
            const double inv = (fabs(Ei) > c->eps1) ? 1.0 / Ei : 0.0;
            sr += (vc[en] - vc[e]) * inv;
            si += (vs[en] - vs[e]) * inv;
            if (fabs(Ei) <= c->eps1)
                add_stable_edge_quotient(vc[e], vs[e], vy[en] - vy[e], Ei, k,
                                         &sr, &si);
        }
        const double invNA = -1.0 / A;
        sr *= invNA;
        si *= invNA;
    }
    /* fresnel = complWave * (sr + i si) */
    *fr = c->compl_wave[0] * sr - c->compl_wave[1] * si;
    *fi = c->compl_wave[0] * si + c->compl_wave[1] * sr;
}

/* rotate_jones_inline, src/handler/HandlerPO_fast.h:580-683 (the exact
 * 1/sqrt normalisation of vt). */
static void rotate_jones(const double *b, const double *vf, const double *dir,
                         double *r)
{
    double vt[3] = {vf[1] * dir[2] - vf[2] * dir[1],
                    vf[2] * dir[0] - vf[0] * dir[2],
                    vf[0] * dir[1] - vf[1] * dir[0]};
    const double l2 = vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2];
    if (l2 > 1e-30) {
        const double inv = 1.0 / sqrt(l2);
        vt[0] *= inv; vt[1] *= inv; vt[2] *= inv;
    }
    const double *NT = b + MBS_B_NT, *NP = b + MBS_B_NP;
    const double *DT = b + MBS_B_NXDT, *DP = b + MBS_B_NXDP;
    const double dT = dir[0] * DT[0] + dir[1] * DT[1] + dir[2] * DT[2];
    const double dP = dir[0] * DP[0] + dir[1] * DP[1] + dir[2] * DP[2];
    const double cT[3] = {(dir[1] * NT[2] - dir[2] * NT[1]) + DT[0] - dir[0] * dT,
                          (dir[2] * NT[0] - dir[0] * NT[2]) + DT[1] - dir[1] * dT,
                          (dir[0] * NT[1] - dir[1] * NT[0]) + DT[2] - dir[2] * dT};
    const double cP[3] = {(dir[1] * NP[2] - dir[2] * NP[1]) + DP[0] - dir[0] * dP,
                          (dir[2] * NP[0] - dir[0] * NP[2]) + DP[1] - dir[1] * dP,
                          (dir[0] * NP[1] - dir[1] * NP[0]) + DP[2] - dir[2] * dP};
    r[0] = (cT[0] * vt[0] + cT[1] * vt[1] + cT[2] * vt[2]) * 0.5;
    r[1] = (cP[0] * vt[0] + cP[1] * vt[1] + cP[2] * vt[2]) * 0.5;
    r[2] = (cT[0] * vf[0] + cT[1] * vf[1] + cT[2] * vf[2]) * 0.5;
    r[3] = (cP[0] * vf[0] + cP[1] * vf[1] + cP[2] * vf[2]) * 0.5;
}

/* Validates one batch the way the reference GPU path does
 * (src/cuda/GpuDiffraction.cu:4297-4303). */
int mbs_oracle_check_batch(const mbs_beam_batch *bt)
{
    for (int b = 0; b < bt->n_beams; ++b) {
        const int nv = bt->beam_nv[b];
        if (nv < MBS_MIN_VERTICES || nv > MBS_MAX_VERTICES)
            return MBS_ERR_BAD_BEAM;
        if (bt->vert_offsets[b + 1] - bt->vert_offsets[b] != nv)
            return MBS_ERR_INVALID;
    }
    return MBS_OK;
}

/* Coherent Jones sum of one orientation over the whole grid.
 * Restates HandlerPO::HandleBeamsToLocal, src/handler/HandlerPO.cpp:2561-2796:
 * per beam, per phi row: ThetaCoeffs (HandlerPO_fast.h:504-560), vertex phases
 * sin(t)*psin + cos(t)*pcos + p0, direction phase -k(sin(t)*dp_sin +
 * cos(t)*dp_cos) for internal beams, aperture integral, Jones rotation, and
 * d = (F*dirPhase) R J_phased summed in beam order into j (8 reals per cell,
 * cell p*(n_zen+1)+t).  jns receives the sum over non-external beams (the
 * localJ_noshadow of :2780-2783); it may be NULL.  Cells are independent, so
 * phi rows run in parallel and the result does not depend on thread count. */
void mbs_oracle_add_mueller(double *out, const double *jj, double w);

/* With inc != NULL the beams' d are not summed: inc += w * Mueller(d) per
 * beam and cell instead (the incoherent sum), and j, jns are unused. */
static int orientation_core(const mbs_grid *g, const mbs_beam_batch *bt, const mbs_consts *c,
                            int orient, double *j, double *jns, double *inc, double w_inc)
{
    const int nAz = g->n_az, nZ = g->n_zen + 1;
    const int64_t b0 = bt->orient_offsets[orient], b1 = bt->orient_offsets[orient + 1];
    const double k = c->wave_index;
    if (!inc)
        memset(j, 0, sizeof(double) * 8 * (size_t)nAz * nZ);
    if (jns && !inc)
        memset(jns, 0, sizeof(double) * 8 * (size_t)nAz * nZ);
    double *st = (double *)malloc(sizeof(double) * 2 * nZ);
    double *ct = st + nZ;
    for (int t = 0; t < nZ; ++t)
        phase_sincos(g->theta[t], &st[t], &ct[t]);
    for (int64_t bi = b0; bi < b1; ++bi) {
        const double *s = bt->beam_scalars + bi * MBS_BEAM_SCALARS;
        const int nv = bt->beam_nv[bi];
        const int64_t vo = bt->vert_offsets[bi];
        const double *vx = bt->vx + vo, *vy = bt->vy + vo;
        const double *syx = bt->slope_yx + vo, *sxy = bt->slope_xy + vo;
        const uint8_t *ef = bt->edge_flags + vo;
        const int ext = bt->beam_external[bi];
        const double *hor = s + MBS_B_HOR, *ver = s + MBS_B_VER;
        const double *bd = s + MBS_B_DIR, *cen = s + MBS_B_CEN;
        const double *jp = s + MBS_B_JP;
#pragma omp parallel for schedule(static)
        for (int p = 0; p < nAz; ++p) {
            double sp, cp;
            phase_sincos(p * g->az_step, &sp, &cp);
            /* precompute_theta_coeffs, HandlerPO_fast.h:504-560 */
            const double a_sin = -cp * hor[0] + -sp * hor[1], a_cos = hor[2];
            const double a0 = bd[0] * hor[0] + bd[1] * hor[1] + bd[2] * hor[2];
            const double b_sin = -cp * ver[0] + -sp * ver[1], b_cos = ver[2];
            const double b0_ = bd[0] * ver[0] + bd[1] * ver[1] + bd[2] * ver[2];
            double psin[ORACLE_MAXV], pcos[ORACLE_MAXV], p0[ORACLE_MAXV];
            for (int v = 0; v < nv; ++v) {
                psin[v] = k * (a_sin * vx[v] + b_sin * vy[v]);
                pcos[v] = k * (a_cos * vx[v] + b_cos * vy[v]);
                p0[v] = k * (a0 * vx[v] + b0_ * vy[v]);
            }
            const double dp_sin = cp * cen[0] + sp * cen[1], dp_cos = -cen[2];
            for (int t = 0; t < nZ; ++t) {
                const double A = st[t] * a_sin + ct[t] * a_cos + a0;
                const double B = st[t] * b_sin + ct[t] * b_cos + b0_;
                double vc[ORACLE_MAXV], vs[ORACLE_MAXV];
                for (int v = 0; v < nv; ++v)
                    phase_sincos(st[t] * psin[v] + ct[t] * pcos[v] + p0[v], &vs[v], &vc[v]);
                double fr, fi;
                aperture_integral(vx, vy, syx, sxy, ef, nv, A, B, vc, vs, c,
                                  s[MBS_B_AREA], &fr, &fi);
                if (isnan(fr))
                    continue;
                double dpr = 1.0, dpi = 0.0;
                if (!ext)
                    phase_sincos(-k * (st[t] * dp_sin + ct[t] * dp_cos), &dpi, &dpr);
                const double dir[3] = {st[t] * cp, st[t] * sp, -ct[t]};
                const size_t cell = (size_t)p * nZ + t;
                double r[4];
                rotate_jones(s, g->vf + 3 * cell, dir, r);
                const double cpr = fr * dpr - fi * dpi, cpi = fr * dpi + fi * dpr;
                const double s00r = cpr * r[0], s00i = cpi * r[0];
                const double s01r = cpr * r[1], s01i = cpi * r[1];
                const double s10r = cpr * r[2], s10i = cpi * r[2];
                const double s11r = cpr * r[3], s11i = cpi * r[3];
                double d[8];
                d[0] = s00r * jp[0] - s00i * jp[1] + s01r * jp[4] - s01i * jp[5];
                d[1] = s00r * jp[1] + s00i * jp[0] + s01r * jp[5] + s01i * jp[4];
                d[2] = s00r * jp[2] - s00i * jp[3] + s01r * jp[6] - s01i * jp[7];
                d[3] = s00r * jp[3] + s00i * jp[2] + s01r * jp[7] + s01i * jp[6];
                d[4] = s10r * jp[0] - s10i * jp[1] + s11r * jp[4] - s11i * jp[5];
                d[5] = s10r * jp[1] + s10i * jp[0] + s11r * jp[5] + s11i * jp[4];
                d[6] = s10r * jp[2] - s10i * jp[3] + s11r * jp[6] - s11i * jp[7];
                d[7] = s10r * jp[3] + s10i * jp[2] + s11r * jp[7] + s11i * jp[6];
                if (inc) {
                    mbs_oracle_add_mueller(inc + 16 * cell, d, w_inc);
                    continue;
                }
                double *jc = j + 8 * cell;
                for (int q = 0; q < 8; ++q)
                    jc[q] += d[q];
                if (jns && !ext) {
                    double *jn = jns + 8 * cell;
                    for (int q = 0; q < 8; ++q)
                        jn[q] += d[q];
                }
            }
        }
    }
    free(st);
    return MBS_OK;
}

int mbs_oracle_orientation_jones(const mbs_grid *g, const mbs_beam_batch *bt,
                                 const mbs_consts *c, int orient,
                                 double *j, double *jns)
{
    return orientation_core(g, bt, c, orient, j, jns, NULL, 0.0);
}

/* Incoherent sum over a batch: m += sum_o w_o sum_b Mueller(d_b) per cell,
 * the Mueller matrices of single beams added without interference.  Its M11
 * element is what the reference GPU theta probe sums, 0.5 |d_b|^2 w_o per
 * beam (theta_m11_kernel, src/cuda/GpuDiffraction.cu:1909-1957). */
int mbs_oracle_incoherent_batch(const mbs_grid *g, const mbs_beam_batch *bt,
                                const mbs_consts *c, double *m)
{
    int rc = mbs_oracle_check_batch(bt);
    if (rc)
        return rc;
    for (int o = 0; o < bt->n_orient; ++o) {
        rc = orientation_core(g, bt, c, o, NULL, NULL, m, bt->weights[o]);
        if (rc)
            return rc;
    }
    return MBS_OK;
}

/* AddMuellerFromJones, src/math/Mueller.hpp:37-87: out += weight*Mueller(J)
 * for J = (j00, j01, j10, j11) given as 8 reals. */
void mbs_oracle_add_mueller(double *out, const double *jj, double w)
{
    const double j00r = jj[0], j00i = jj[1], j01r = jj[2], j01i = jj[3];
    const double j10r = jj[4], j10i = jj[5], j11r = jj[6], j11i = jj[7];
    const double a11 = j00r * j00r + j00i * j00i, a12 = j01r * j01r + j01i * j01i;
    const double a21 = j10r * j10r + j10i * j10i, a22 = j11r * j11r + j11i * j11i;
    double A1 = a11 + a21, A2 = a12 + a22;
    out[0] += ((A1 + A2) * 0.5) * w;
    out[1] += ((A1 - A2) * 0.5) * w;
    A1 = a11 - a21; A2 = a12 - a22;
    out[4] += ((A1 + A2) * 0.5) * w;
    out[5] += ((A1 - A2) * 0.5) * w;
    double c1r = j00r * j01r + j00i * j01i, c1i = j00i * j01r - j00r * j01i;
    double c2r = j11r * j10r + j11i * j10i, c2i = j11i * j10r - j11r * j10i;
    out[2] += (-c1r - c2r) * w;
    out[3] += (c2i - c1i) * w;
    out[6] += (c2r - c1r) * w;
    out[7] += (-c1i - c2i) * w;
    c1r = j00r * j10r + j00i * j10i; c1i = j00i * j10r - j00r * j10i;
    c2r = j11r * j01r + j11i * j01i; c2i = j11i * j01r - j11r * j01i;
    out[8] += (-c1r - c2r) * w;
    out[9] += (c2r - c1r) * w;
    out[12] += (c1i - c2i) * w;
    out[13] += (c2i + c1i) * w;
    c1r = j00r * j11r + j00i * j11i; c1i = j00i * j11r - j00r * j11i;
    c2r = j01r * j10r + j01i * j10i; c2i = j01i * j10r - j01r * j10i;
    out[10] += (c1r + c2r) * w;
    out[11] += (c1i - c2i) * w;
    out[14] += (-c1i - c2i) * w;
    out[15] += (c1r - c2r) * w;
}

/* Weighted Mueller sum over a batch of orientations, the CPU phase 2 of
 * TracerPOTotal::TraceWeightedOrientations (src/scattering/TracerPOTotal.cpp:
 * 5954-5995): m += sum_o w_o Mueller(J_o) per cell, orientations in order.
 * m_ns (may be NULL) gets the no-shadow sum.  per_orient (may be NULL)
 * receives w_o*Mueller(J_o) per orientation ([n_orient][cells][16]). */
int mbs_oracle_diffract_batch(const mbs_grid *g, const mbs_beam_batch *bt,
                              const mbs_consts *c, double *m, double *m_ns,
                              double *per_orient)
{
    int rc = mbs_oracle_check_batch(bt);
    if (rc)
        return rc;
    const size_t cells = (size_t)g->n_az * (g->n_zen + 1);
    double *j = (double *)malloc(sizeof(double) * 8 * cells);
    double *jn = m_ns ? (double *)malloc(sizeof(double) * 8 * cells) : NULL;
    for (int o = 0; o < bt->n_orient; ++o) {
        mbs_oracle_orientation_jones(g, bt, c, o, j, jn);
        const double w = bt->weights[o];
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)cells; ++q) {
            mbs_oracle_add_mueller(m + 16 * q, j + 8 * q, w);
            if (jn)
                mbs_oracle_add_mueller(m_ns + 16 * q, jn + 8 * q, w);
            if (per_orient) {
                double *po = per_orient + ((size_t)o * cells + q) * 16;
                memset(po, 0, 16 * sizeof(double));
                mbs_oracle_add_mueller(po, j + 8 * q, w);
            }
        }
    }
    free(j);
    free(jn);
    return MBS_OK;
}

/* 2*pi*dcos of theta row t with bin edges at the midpoints between rows, the
 * stable sine form of ScatteringRange::Compute2PiDcos,
 * src/handler/Handler.h:70-125. */
double mbs_oracle_2pi_dcos(const double *theta, int n_zen, int t)
{
    double lo, hi;
    if (t == 0) {
        lo = theta[0];
        hi = 0.5 * (theta[0] + theta[1]);
    } else if (t == n_zen) {
        lo = 0.5 * (theta[n_zen - 1] + theta[n_zen]);
        hi = theta[n_zen];
    } else {
        lo = 0.5 * (theta[t - 1] + theta[t]);
        hi = 0.5 * (theta[t] + theta[t + 1]);
    }
    return 2.0 * sin(0.5 * (lo + hi)) * sin(0.5 * (hi - lo)) * (2.0 * M_PI);
}

/* Azimuth (laboratory-alpha) average of an orientation-summed grid and the
 * pole forms: HandlerPOTotal::WriteMatricesToFile,
 * src/handler/HandlerPOTotal.cpp:23-122 with PoleMueller.h:25-55.
 * out[t*18 + 0] = theta (rad), out[t*18+1] = 2*pi*dcos,
 * out[t*18+2+ij] = mean_p M(p,t) L(-phi_p) (poles: plain mean + pole form). */
void mbs_oracle_azimuth_average(const mbs_grid *g, const double *m, double *out)
{
    const int nZ = g->n_zen + 1;
    const double tol = 256.0 * DBL_EPSILON;
    for (int t = 0; t < nZ; ++t) {
        const double th = g->theta[t];
        const int fwd = fabs(th) <= tol, bwd = fabs(th - M_PI) <= tol;
        double s[16] = {0};
        for (int p = 0; p < g->n_az; ++p) {
            const double *c = m + ((size_t)p * nZ + t) * 16;
            const double radAz = -p * g->az_step;
            const double cs = cos(2 * radAz), sn = sin(2 * radAz);
            for (int r = 0; r < 4; ++r) {
                const double *row = c + 4 * r;
                if (fwd || bwd) {
                    for (int q = 0; q < 4; ++q)
                        s[4 * r + q] += row[q];
                } else {
                    /* (m * Lp)[r][.]: Lp[1][1]=cs Lp[1][2]=sn Lp[2][1]=-sn Lp[2][2]=cs */
                    s[4 * r + 0] += row[0];
                    s[4 * r + 1] += row[1] * cs + row[2] * (-sn);
                    s[4 * r + 2] += row[1] * sn + row[2] * cs;
                    s[4 * r + 3] += row[3];
                }
            }
        }
        for (int q = 0; q < 16; ++q)
            s[q] /= g->n_az;
        if (bwd || fwd) {
            const double M00 = s[0], M33 = s[15], M03 = s[3];
            const double M11 = fwd ? 0.5 * (s[5] + s[10]) : 0.5 * (s[5] - s[10]);
            memset(s, 0, sizeof s);
            s[0] = M00;
            s[5] = M11;
            s[10] = fwd ? M11 : -M11;
            s[15] = M33;
            s[3] = M03;
            s[12] = M03;
        }
        double *o = out + (size_t)t * 18;
        o[0] = th;
        o[1] = mbs_oracle_2pi_dcos(g->theta, g->n_zen, t);
        for (int q = 0; q < 16; ++q)
            o[2 + q] = s[q];
    }
}

/* Polygon aperture integral F(A,B) for arbitrary A, B, from the same edge and
 * small-phase formulas (used by tests for closed-form checks). */
void mbs_oracle_aperture(const double *vx, const double *vy, const double *syx,
                         const double *sxy, const uint8_t *ef, int nv,
                         double A, double B, const mbs_consts *c, double area,
                         double *fr, double *fi)
{
    double vc[ORACLE_MAXV], vs[ORACLE_MAXV];
    for (int v = 0; v < nv; ++v)
        sincos(c->wave_index * (A * vx[v] + B * vy[v]), &vs[v], &vc[v]);
    aperture_integral(vx, vy, syx, sxy, ef, nv, A, B, vc, vs, c, area, fr, fi);
}

/* ---- orientation sets ----------------------------------------------------
 * Sobol2D restated in its sequential form (src/math/Sobol.cpp:1-106): the
 * state XORs direction number c = index of the lowest zero bit of the point
 * counter, then each coordinate is Owen-scrambled bit by bit with a hash of
 * (seed, dimension, level, prefix of original bits). */
static uint32_t sobol_hash(uint32_t seed, uint32_t dim, uint32_t level, uint32_t prefix)
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

static uint32_t sobol_owen(uint32_t seed, uint32_t value, uint32_t dim)
{
    uint32_t out = 0, prefix = 0;
    for (uint32_t level = 0; level < 32; ++level) {
        const uint32_t shift = 31u - level, bit = (value >> shift) & 1u;
        out |= (bit ^ (sobol_hash(seed, dim, level, prefix) & 1u)) << shift;
        prefix = (prefix << 1) | bit;
    }
    return out;
}

void mbs_oracle_sobol(int n, uint32_t seed, double *x, double *y)
{
    uint32_t dir[2][32];
    for (int i = 0; i < 32; ++i)
        dir[0][i] = 1u << (31 - i);
    dir[1][0] = 1u << 31;
    for (int i = 1; i < 32; ++i)
        dir[1][i] = dir[1][i - 1] ^ (dir[1][i - 1] >> 1);
    uint32_t s0 = 0, s1 = 0;
    for (uint32_t k = 0; k < (uint32_t)n; ++k) {
        uint32_t c = 0, v = k;
        while (v & 1u) {
            v >>= 1;
            ++c;
        }
        if (c >= 32)
            c = 0;
        s0 ^= dir[0][c];
        s1 ^= dir[1][c];
        x[k] = (double)sobol_owen(seed, s0, 0) / 4294967296.0;
        y[k] = (double)sobol_owen(seed, s1, 1) / 4294967296.0;
    }
}

static double radical_inverse(uint32_t v, uint32_t base)
{
    const double ib = 1.0 / (double)base;
    double inv = ib, r = 0.0;
    while (v > 0) {
        r += (double)(v % base) * inv;
        v /= base;
        inv *= ib;
    }
    return r;
}

/* Particle::SetRotateMatrix(beta, gamma, alpha) and
 * SetRotateMatrixFromQuaternion (src/particle/Particle.cpp:1593-1642). */
static void oracle_euler_matrix(double beta, double gamma, double alpha, double *R)
{
    double sA, cA, sB, cB, sG, cG;
    sincos(alpha, &sA, &cA);
    sincos(beta, &sB, &cB);
    sincos(gamma, &sG, &cG);
    R[0] = cA * cB * cG - sA * sG;
    R[1] = -(cA * cB * sG + sA * cG);
    R[2] = cA * sB;
    R[3] = sA * cG * cB + cA * sG;
    R[4] = cA * cG - sA * sG * cB;
    R[5] = sA * sB;
    R[6] = -sB * cG;
    R[7] = sB * sG;
    R[8] = cB;
}

static void oracle_quat_matrix(double x, double y, double z, double w, double *R)
{
    const double n = sqrt(x * x + y * y + z * z + w * w);
    if (n <= DBL_EPSILON) {
        x = y = z = 0.0;
        w = 1.0;
    } else {
        x /= n; y /= n; z /= n; w /= n;
    }
    R[0] = 1.0 - 2.0 * (y * y + z * z);
    R[1] = 2.0 * (x * y - w * z);
    R[2] = 2.0 * (x * z + w * y);
    R[3] = 2.0 * (x * y + w * z);
    R[4] = 1.0 - 2.0 * (x * x + z * z);
    R[5] = 2.0 * (y * z - w * x);
    R[6] = 2.0 * (x * z - w * y);
    R[7] = 2.0 * (y * z + w * x);
    R[8] = 1.0 - 2.0 * (x * x + y * y);
}

/* Orientation sets, rows of MBS_ORIENT_STRIDE doubles (see mbs_orientation_set
 * in include/mbs_hip.h): Sobol (TracerPOTotal.cpp:5465-5492), symmetry
 * Hammersley (:5347-5426, unpaired) and full-SO(3) Shoemake quaternions
 * (:5428-5463).  The quaternion of an Euler row is left zero here (the tests
 * check the engine's through its rotation matrix). */
void mbs_oracle_orientation_set(int kind, int n, uint32_t seed, double beta_sym,
                                double gamma_sym, double *out)
{
    double *x = malloc(sizeof(double) * (size_t)n), *y = malloc(sizeof(double) * (size_t)n);
    if (kind == MBS_ORIENT_SOBOL)
        mbs_oracle_sobol(n, seed, x, y);
    const double cbs = cos(beta_sym);
    for (int i = 0; i < n; ++i) {
        double *o = out + (size_t)i * MBS_ORIENT_STRIDE;
        memset(o, 0, sizeof(double) * MBS_ORIENT_STRIDE);
        o[2] = 1.0 / (double)n;
        if (kind == MBS_ORIENT_SHOEMAKE) {
            const double u1 = (i + 0.5) / n, u2 = radical_inverse((uint32_t)i, 2),
                         u3 = radical_inverse((uint32_t)i, 3);
            const double r1 = sqrt(fmax(0.0, 1.0 - u1)), r2 = sqrt(u1);
            const double a = 2.0 * M_PI * u2, b = 2.0 * M_PI * u3;
            o[0] = o[1] = NAN;
            o[3] = r1 * sin(a);
            o[4] = r1 * cos(a);
            o[5] = r2 * sin(b);
            o[6] = r2 * cos(b);
            oracle_quat_matrix(o[3], o[4], o[5], o[6], o + 7);
        } else {
            const double u = kind == MBS_ORIENT_SOBOL ? x[i] : (i + 0.5) / (double)n;
            const double v = kind == MBS_ORIENT_SOBOL ? y[i] : radical_inverse((uint32_t)i, 2);
            o[0] = acos(1.0 - (1.0 - cbs) * u);
            o[1] = gamma_sym * v;
            oracle_euler_matrix(o[0], o[1], 0.0, o + 7);
        }
    }
    free(x);
    free(y);
}

/* ---- aperture FFT raster -------------------------------------------------
 * The 0/1 raster of a convex aperture polygon on an n x n grid of pixel
 * centres spanning pad times its bounding box, as the device aperture FFT
 * builds it (the raster of tools/fft_aperture_probe.cu:130-169 for polygons):
 * pixel centre r = c + ((2j + 1 - n)/2) d per axis; inside iff a x + b y + c
 * > 0 on every edge line (oriented so that the inside is positive), every
 * product formed with the same explicit FMAs as the device.  grid: [n][n]
 * row-major (y, x); geom: xc, yc, dx, dy. */
void mbs_oracle_aperture_raster(const double *vx, const double *vy, int nv, int n, double pad,
                                double *grid, double *geom)
{
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300, area2 = 0.0;
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        x0 = fmin(x0, vx[e]);
        x1 = fmax(x1, vx[e]);
        y0 = fmin(y0, vy[e]);
        y1 = fmax(y1, vy[e]);
        area2 += vx[e] * vy[f] - vx[f] * vy[e];
    }
    const double xc = 0.5 * (x0 + x1), yc = 0.5 * (y0 + y1);
    const double dx = pad * (x1 - x0) / n, dy = pad * (y1 - y0) / n;
    const double sg = area2 >= 0.0 ? 1.0 : -1.0;
    double L[3 * ORACLE_MAXV];
    for (int e = 0; e < nv; ++e) {
        const int f = e + 1 < nv ? e + 1 : 0;
        const double a = -(vy[f] - vy[e]) * sg, b = (vx[f] - vx[e]) * sg;
        L[3 * e] = a;
        L[3 * e + 1] = b;
        L[3 * e + 2] = -fma(a, vx[e], b * vy[e]);
    }
    for (int y = 0; y < n; ++y) {
        const double py = fma((double)(2 * y + 1 - n), 0.5 * dy, yc);
        for (int x = 0; x < n; ++x) {
            const double px = fma((double)(2 * x + 1 - n), 0.5 * dx, xc);
            int in = 1;
            for (int e = 0; e < nv; ++e)
                in = in && (fma(L[3 * e], px, fma(L[3 * e + 1], py, L[3 * e + 2])) > 0.0);
            grid[(size_t)y * n + x] = in ? 1.0 : 0.0;
        }
    }
    geom[0] = xc;
    geom[1] = yc;
    geom[2] = dx;
    geom[3] = dy;
}
