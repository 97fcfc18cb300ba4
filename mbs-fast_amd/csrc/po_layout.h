// po_layout.h -- device-side layout of packed beams and grid tables.
#pragma once

#include <stdint.h>

namespace mbs {

// Beam header, FP64, HDR_STRIDE doubles per beam (256 B aligned records so a
// wave-uniform load of a header is a few 64-byte scalar loads).
enum {
    H_HOR = 0,   // aperture horizontal axis (3)
    H_A0 = 3,    // beamDir . hor
    H_VER = 4,   // aperture vertical axis (3)
    H_B0 = 7,    // beamDir . ver
    H_KC = 8,    // -k * centre (3), zero for external beams
    H_PHBM = 11, // the mixed kernels' bound for 0.25 rad (FP32, in the low word)
    H_AL = 12,   // small-phase area limit / complWave (complex)
    H_MOM = 14,  // k mX, k mY, k^2 mXX, 2 k^2 mXY, k^2 mYY (5)
    H_PHB = 19,  // max(|A|,|B|) above which some vertex phase exceeds 1e-3 rad
    H_T0 = 20,   // 1/2 (NT, nxDT)   (6)
    H_T1 = 26,   // 1/2 (NP, nxDP)   (6)
    H_JP = 32,   // complWave * J_phased (8)
    // moments of orders 3, 4, 5 for the mixed kernels' moment form: entry a
    // of order n multiplies A^a B^(n-a) and holds k^n C(n,a) <x^a y^(n-a)> / n!
    H_MOM3 = 40,  // (4)
    H_MOM4 = 44,  // (5)
    H_MOM5 = 49,  // (6)
    HDR_STRIDE = 56
};

// FP32 copy of the rotation data (T0, T1, complWave * J_phased) used by the
// vector-ALU form of the mixed kernel.
enum { H32_T0 = 0, H32_T1 = 6, H32_JP = 12, HDR32_STRIDE = 20 };

// MFMA B operand of one beam (mixed kernel), WB_STRIDE floats in lane order:
// lane l holds B[l/32][l%32] of the complex contraction
//   G[dir][n] += Re F(dir) * B[0][n] + Im F(dir) * B[1][n],
// with column n = 4m + 2c + r for rotation component m (0..5), Jones column c
// and r = 0 (real part) / 1 (imaginary part) of W[m][c] = T0[m] Jp[0][c] +
// T1[m] Jp[1][c]: B[0][n] = (Re W, Im W), B[1][n] = (-Im W, Re W); columns
// 24..31 are zero.
enum { WB_COLS = 24, WB_STRIDE = 64 };

// Vertex records, FP64, 32 bytes: k x, k y, slope dy/dx, slope dx/dy (the
// slopes of edge e = vertex e -> vertex e+1).  The edge deltas dx, dy, read
// only by the rare near-singular branch, live in a separate array.
enum { V_KX = 0, V_KY = 1, V_SYX = 2, V_SXY = 3, VTX_STRIDE = 4 };
enum { D_DX = 0, D_DY = 1, VDD_STRIDE = 2 };
// Per-vertex edge flags (as the C ABI's MBS_EDGE_VALID_X / _Y): the edge
// from this vertex has a usable dx (enters the |B| > |A| sum) / dy.
enum { EDGE_VX = 1, EDGE_VY = 2 };

// Grid tables, SoA [DT_COUNT][cells], cells in the device's theta-major order.
enum { DT_DIR = 0, DT_U0 = 3, DT_U1 = 9, DT_COUNT = 15 };
// FP32 copies of u0, u1 for the mixed kernels, [cells][DT32_COUNT].
enum { DT32_COUNT = 12 };

// A unit of work: orientations [o_begin, o_end) whole (Mueller mode), or the
// beam range [b_begin, b_end) of orientation o_begin (Jones mode).
struct WorkItem {
    int32_t o_begin;
    int32_t o_end;
    int64_t b_begin;
    int64_t b_end;
};

// Header of the device work queue (work mode MBS_WORK_QUEUE), written by
// plan_segments_kernel: the item count, then for each of the 8 XCDs the next
// unit to hand out and the end of its range of units (unit = item * n_tiles
// + tile).
enum { Q_ITEMS = 0, Q_NEXT = 1, Q_END = 9, Q_WORDS = 17 };

}  // namespace mbs
