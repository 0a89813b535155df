// ibm.hip — regularized-delta interpolation and force spreading
// (ZJS_VOLUMEFRACTION_DEM.F90:2-67, ZJS_INTERPOLATE_DEM.F90:2-17,
// interpolation.F90:2-33) with the multi-direct-forcing velocity update
// (fluid.F90:135-153) fused into the spread.
//
// Markers do not move during the fluid part of a step, so once per step:
//   - every marker gets a 64-byte record (position, cell; the interpolation
//     adds fdrag*vol to it each iteration), so the spread reads one line per
//     marker instead of three SoA arrays;
//   - the spread footprint is compacted into a sorted list of active 8x8x8
//     cell tiles, and each active tile gets the list of its source entries:
//     the markers of its 11x11x11 source cells in source order (kc,jc,ic)
//     and ascending id inside a cell, each packed as marker << 9 | image
//     code (3 bits per dimension: direct, periodic copy below or above,
//     mirror copy below or above).
//   - every marker's one-dimensional delta factors (per dimension:
//     staggered-face and cell-centre variants at cell offsets -1..+2) are
//     computed once and read by the interpolation and the spread of all
//     five iterations.
//
// Interpolation is marker-parallel and sums the taps in the reference's
// (kk,jj,ii) order, so it is bit-identical to the reference's 125-cell loop
// (zero-weight taps add an exact zero there).
//
// Spreading is deterministic: no two wavefronts ever add to one value. One
// wavefront owns a tile and keeps the tile's force (3 x 512 doubles) in LDS.
// It stages its entry list in chunks (delta factors, fdrag*vol and the cell
// of the marker or of its image) and takes the entries in list order: lane l
// adds the entry's contribution -(fdrag*vol)*w to cell l of the entry's
// 4x4x4 support (when that cell is in the tile) with an LDS add. The 64
// lanes of one entry add to 64 different cells, and the LDS applies a
// wavefront's operations in issue order, so every cell receives its
// contributions in entry-list order: the sum does not depend on timing. The reference's periodic ghost copies
// (x + coordShift) and wall-mirror copies (2*coordMirror - x, fdrag
// unnegated; ZJS_MPI_DEM.F90:173-269) enter as image entries with their
// factors evaluated at the image position. The wavefront then applies
// u += dt*(f - f_old) on the tile's cells that some entry reached; the
// others are skipped (their f is structurally zero in every iteration of
// the step).
#include "common.hpp"
#include "batched.hpp"
#include "scan.hpp"

#include <algorithm>
#include <mutex>

namespace {

using namespace bk;

constexpr int T = 256;
constexpr int CHUNK = 4096;   // flags per compaction block
static_assert(T == scan::SCAN_T && CHUNK == scan::SCAN_CHUNK, "the entry-count scan uses the shared chunk kernels");
constexpr int TL = IBM_TL;    // tile edge (cells)
constexpr int TC = IBM_TC;    // cells per tile
constexpr int SRC = TL + 3;   // source cells per tile edge
constexpr int NSRC = SRC * SRC * SRC;
// entries staged per pass: one pass's fixed costs (the staging, the wait for
// its records) are spread over 32 entries; 16 or 48 measured slower
constexpr int CAP = 32;
constexpr int WST2 = 13;      // LDS row stride of the 12 factor pairs (odd: fewer bank conflicts)
constexpr int SU = 4;         // entries whose terms are formed before their adds (2 or 8: slower)
static_assert(CAP % SU == 0, "entry groups tile the staged rows");
constexpr int SBITS = 9;      // image-code bits of an entry (3 per dimension)

// one-dimensional delta factors of a marker at xp for cell i: staggered
// face (x = (i-1)d) and cell centre (x = (i-1/2)d), GetWeight
// (ZJS_INTERPOLATE_DEM.F90:12-14); id = 1/d, used when d is a power of two
template <bool P2, bool BR = false>
__device__ __forceinline__ double wt_face(double xp, int i, double d, double id) {
    const double xc = ((double)i - 0.5) * d;
    const double dl = xp - xc;
    const double r = fabs(qdiv<P2>(dl + 0.5 * d, d, id));
    return BR ? phi1_branch(r) : phi1(r);
}
template <bool P2, bool BR = false>
__device__ __forceinline__ double wt_cent(double xp, int i, double d, double id) {
    const double xc = ((double)i - 0.5) * d;
    const double dl = xp - xc;
    const double r = fabs(qdiv<P2>(dl, d, id));
    return BR ? phi1_branch(r) : phi1(r);
}

// Active-tile slot of block b. The hardware hands block b to XCD b % 8; XCD x
// gets the contiguous slots [x*per, (x+1)*per) so tiles that share source
// markers share one L2. Returns -1 for blocks beyond the active count.
__device__ __forceinline__ int tile_slot(int b, int nact) {
    const int per = (nact + 7) >> 3;
    const int x = b & 7, q = b >> 3;
    if (q >= per) return -1;
    const int slot = x * per + q;
    return slot < nact ? slot : -1;
}

// per-step marker record: position, cell; zero fdrag; and the marker's
// one-dimensional delta factors, fac[d*4 + o] = (face, centre) for cell
// offset o - 1 of dimension d (the centre factor at offset +2 is always 0:
// the marker lies in its cell, so that centre is at least 3/2 cells away).
// Markers do not move during the five interpolate/spread iterations of a
// step, so these are computed once here and read by both.
template <bool P2>
__device__ __forceinline__ void records_body(const Blk &kb, int M, const GridDev &g, const double *__restrict__ mx,
                                             const int *__restrict__ mcell, double *__restrict__ rec,
                                             double *__restrict__ fac, double *__restrict__ fdrag) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M) return;
    double *r = rec + 8L * m;
    const double xp[3] = {mx[m], mx[(long)M + m], mx[2L * M + m]};
    const int cl = mcell[m];
    reinterpret_cast<double2 *>(r)[0] = make_double2(xp[0], xp[1]);
    reinterpret_cast<double2 *>(r)[1] = make_double2(xp[2], __longlong_as_double((long long)cl));
    for (int d = 0; d < 3; ++d) fdrag[(long)d * M + m] = 0.0;
    const int cc[3] = {cl % g.nx + 1, (cl / g.nx) % g.ny + 1, cl / (g.nx * g.ny) + 1};
    const double dd[3] = {g.dx, g.dy, g.dz};
    double2 *f = reinterpret_cast<double2 *>(fac + 24L * m);
    for (int d = 0; d < 3; ++d) {
        const double id = 1.0 / dd[d];
#pragma unroll
        for (int o = 0; o < 4; ++o)
            f[d * 4 + o] = make_double2(wt_face<P2, true>(xp[d], cc[d] + o - 1, dd[d], id),
                                        o < 3 ? wt_cent<P2, true>(xp[d], cc[d] + o - 1, dd[d], id) : 0.0);
    }
}
template <bool P2>
__global__ __launch_bounds__(T) void k_records(int M, GridDev g, const double *__restrict__ mx,
                                               const int *__restrict__ mcell, double *__restrict__ rec,
                                               double *__restrict__ fac, double *__restrict__ fdrag) {
    records_body<P2>(hw_blk(), M, g, mx, mcell, rec, fac, fdrag);
}

// Ghost rules of the four taps at offsets -1..+2 of one dimension around a
// marker's cell (the rules of vel_at): mapped 0-based index, and bit masks
// (bit q for offset q - 1) of tangential mirrors that negate, wall-normal
// zeros, inflow constants and outflow-plane reads.
struct TapMap4 {
    int idx[4];
    unsigned mir, wall, cst, pln;
    __device__ __forceinline__ void init(int i, int n, int lo, int hi) {
        mir = wall = cst = pln = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int j = i - 1 + q;
            if (j <= 1 && lo == IBMDEM_BC_INFLOW) cst |= 1u << q;
            if (j < 1) {
                if (lo == IBMDEM_BC_PERIODIC) j += n;
                else {
                    j = 1 - j;
                    mir |= 1u << q;
                    if (lo == IBMDEM_BC_WALL) wall |= 1u << q;
                }
            } else if (j > n) {
                if (hi == IBMDEM_BC_PERIODIC) j -= n;
                else {
                    j = 2 * n + 1 - j;
                    if (hi == IBMDEM_BC_OUTFLOW) pln |= 1u << q;
                    else mir |= 1u << q;
                    if (hi == IBMDEM_BC_WALL) wall |= 1u << q;
                }
            }
            idx[q] = j - 1;
        }
    }
};

__device__ __forceinline__ long tap_off(const GridDev &g, const TapMap4 &mx, const TapMap4 &my, const TapMap4 &mz,
                                       int a, int b, int c) {
    return (long)mx.idx[a] + (long)g.nx * ((long)my.idx[b] + (long)g.ny * mz.idx[c]);
}

// One interpolation term of the reference, wt * u * dx * dy * dz with
// wt = (w1 w2 w3) / (dx dy dz) (delta3d, interpolation.F90:7-30), given
// q = w1 w2 w3. When the spacings and the cell volume are powers of two the
// divisions and multiplications by them are exact scalings, so the term is
// q * u rounded once, bit for bit (in the normal range of doubles).
template <bool P2>
__device__ __forceinline__ double interp_term(double q, double val, double dV, double idV, double dx, double dy,
                                              double dz) {
    if (P2) return q * val;
    return q / dV * val * dx * dy * dz;
}

// velocity component C at the taps (a, b, c) of three TapMap4 (vel_at's
// value: tested bit for bit against the oracle's halo arrays)
template <int C, bool OPEN>
__device__ __forceinline__ double tap4(const double *__restrict__ f, const GridDev &g, const TapMap4 &mx,
                                       const TapMap4 &my, const TapMap4 &mz, int a, int b, int c) {
    const bool zero = C == 0 ? (mx.wall >> a) & 1u : (C == 1 ? (my.wall >> b) & 1u : (mz.wall >> c) & 1u);
    const unsigned negbits = (C != 0 ? (mx.mir >> a) & 1u : 0u) ^ (C != 1 ? (my.mir >> b) & 1u : 0u) ^
                             (C != 2 ? (mz.mir >> c) & 1u : 0u);
    const long fo = (long)mx.idx[a] + (long)g.nx * ((long)my.idx[b] + (long)g.ny * mz.idx[c]);
    double val;
    if (OPEN) {
        const bool cst = C == 0 ? (mx.cst >> a) & 1u : (C == 1 ? (my.cst >> b) & 1u : (mz.cst >> c) & 1u);
        const bool pln = C == 0 ? (mx.pln >> a) & 1u : (C == 1 ? (my.pln >> b) & 1u : (mz.pln >> c) & 1u);
        const long po = C == 0 ? (long)my.idx[b] + (long)g.ny * mz.idx[c]
                               : (C == 1 ? (long)mx.idx[a] + (long)g.nx * mz.idx[c]
                                         : (long)mx.idx[a] + (long)g.nx * my.idx[b]);
        val = pln ? g.plane[C][po] : f[fo];
        val = negbits ? -val : val;
        return cst ? g.uin[C] : (zero ? 0.0 : val);
    }
    val = f[fo];
    val = negbits ? -val : val;
    return zero ? 0.0 : val;
}

// UpdateFluidVelocityOnLagrangianPoints: U_p = sum w u dx dy dz,
// fdrag += (U_p - v_p)/dt; also stores fdrag*vol in the marker record.
// Markers are taken in cell-bin order (a permutation of 0..M-1), so the 64
// markers of a wavefront are neighbours and share their velocity taps in
// L1/L2; each marker's arithmetic is independent of the order.
template <bool P2, bool OPEN>
__device__ __forceinline__ void interp_body(const Blk &kb, int M, GridDev g, const int *__restrict__ order,
                                              const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ w,
                                              double *__restrict__ rec, const double *__restrict__ mv,
                                              double *__restrict__ fdrag, const double *__restrict__ mvol, double dV,
                                              double idV, double dt) {
    // consecutive bin-ordered markers (one body's neighbourhood) stay on one
    // XCD, so the field taps they share come from one L2
    const int slot = tile_slot(kb.bx, (M + T - 1) / T);
    if (slot < 0) return;
    const int t = slot * T + kb.tx;
    if (t >= M) return;
    const int m = order[t];
    double *r = rec + 8L * m;
    const double2 r01 = reinterpret_cast<const double2 *>(r)[0];
    const double2 r23 = reinterpret_cast<const double2 *>(r)[1];
    const double xp[3] = {r01.x, r01.y, r23.x};
    const int cl = (int)__double_as_longlong(r23.y);
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    // the marker's delta factors, computed here rather than read from the
    // step's table (the same function as k_records, so the same values; 21
    // evaluations cost less than streaming the 192-byte row every
    // iteration); centre factors at offset +2 are 0 and not used below
    double fx[4], cx[3], fy[4], cy[3], fz[4], cz[3];
    auto factors = [&]() {
        const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            fx[o] = wt_face<P2, true>(xp[0], ci + o - 1, dx, ix);
            fy[o] = wt_face<P2, true>(xp[1], cj + o - 1, dy, iy);
            fz[o] = wt_face<P2, true>(xp[2], ck + o - 1, dz, iz);
            if (o < 3) {
                cx[o] = wt_cent<P2, true>(xp[0], ci + o - 1, dx, ix);
                cy[o] = wt_cent<P2, true>(xp[1], cj + o - 1, dy, iy);
                cz[o] = wt_cent<P2, true>(xp[2], ck + o - 1, dz, iz);
            }
        }
    };
    // the ghost rules of the taps -1..+2 around the cell, once per marker
    TapMap4 mx, my, mz;
    mx.init(ci, g.nx, g.bt[0], g.bt[1]);
    my.init(cj, g.ny, g.bt[2], g.bt[3]);
    mz.init(ck, g.nz, g.bt[4], g.bt[5]);
    double ulp = 0.0, vlp = 0.0, wlp = 0.0;
    // A marker whose taps need no mirror, zero or open-side value (all of
    // them away from walls and open sides; periodic wraps are only index
    // arithmetic) reads every tap directly.
    const unsigned special = mx.mir | mx.wall | my.mir | my.wall | mz.mir | mz.wall |
                             (OPEN ? (mx.cst | mx.pln | my.cst | my.pln | mz.cst | mz.pln) : 0u);
    const bool plain = special == 0u;
    if (plain) {
        // Each component's 36 taps are loaded before its sum starts, so a
        // wavefront keeps them in flight together; the sum then runs in the
        // reference's (kk, jj, ii) order. The factors are computed while the
        // u taps are in flight.
        // u: x-face taps -1..2, y/z centre taps -1..1 (the nonzero ranges)
        double t[36];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) t[(kk * 3 + jj) * 4 + ii] = u[tap_off(g, mx, my, mz, ii, jj, kk)];
        factors();
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the sum
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 4; ++ii)
                    ulp = ulp + interp_term<P2>(fx[ii] * cy[jj] * cz[kk], t[(kk * 3 + jj) * 4 + ii], dV, idV, dx, dy,
                                                dz);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) t[(kk * 4 + jj) * 3 + ii] = v[tap_off(g, mx, my, mz, ii, jj, kk)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
                    vlp = vlp + interp_term<P2>(cx[ii] * fy[jj] * cz[kk], t[(kk * 4 + jj) * 3 + ii], dV, idV, dx, dy,
                                                dz);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) t[(kk * 3 + jj) * 3 + ii] = w[tap_off(g, mx, my, mz, ii, jj, kk)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
                    wlp = wlp + interp_term<P2>(cx[ii] * cy[jj] * fz[kk], t[(kk * 3 + jj) * 3 + ii], dV, idV, dx, dy,
                                                dz);
    } else {
        factors();
        for (int kk = 0; kk < 3; ++kk)
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 4; ++ii)
                    ulp = ulp + interp_term<P2>(fx[ii] * cy[jj] * cz[kk], tap4<0, OPEN>(u, g, mx, my, mz, ii, jj, kk),
                                                dV, idV, dx, dy, dz);
        for (int kk = 0; kk < 3; ++kk)
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
                    vlp = vlp + interp_term<P2>(cx[ii] * fy[jj] * cz[kk], tap4<1, OPEN>(v, g, mx, my, mz, ii, jj, kk),
                                                dV, idV, dx, dy, dz);
        for (int kk = 0; kk < 4; ++kk)
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
                    wlp = wlp + interp_term<P2>(cx[ii] * cy[jj] * fz[kk], tap4<2, OPEN>(w, g, mx, my, mz, ii, jj, kk),
                                                dV, idV, dx, dy, dz);
    }
    const double f0 = fdrag[m] + 1.0 * (ulp - mv[m]) / dt;
    const double f1 = fdrag[(long)M + m] + 1.0 * (vlp - mv[(long)M + m]) / dt;
    const double f2 = fdrag[2L * M + m] + 1.0 * (wlp - mv[2L * M + m]) / dt;
    fdrag[m] = f0;
    fdrag[(long)M + m] = f1;
    fdrag[2L * M + m] = f2;
    const double vol = mvol[m];
    reinterpret_cast<double2 *>(r)[2] = make_double2(f0 * vol, f1 * vol);
    r[6] = f2 * vol;
}
template <bool P2, bool OPEN>
__global__ __launch_bounds__(T) void k_interp(int M, GridDev g, const int *__restrict__ order,
                                              const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ w,
                                              double *__restrict__ rec, const double *__restrict__ mv,
                                              double *__restrict__ fdrag, const double *__restrict__ mvol, double dV,
                                              double idV, double dt) {
    interp_body<P2, OPEN>(hw_blk(), M, g, order, u, v, w, rec, mv, fdrag, mvol, dV, idV, dt);
}

struct TileGeom {
    int ntx, nty, ntz;
};

// mark the tiles holding an interior cell that a marker (or an image) can reach
__device__ __forceinline__ void mark_tiles_body(const Blk &kb, int M, const GridDev &g, TileGeom tg,
                                                const int *__restrict__ mcell, unsigned char *__restrict__ flag) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M) return;
    const int cl = mcell[m];
    const int cc[3] = {cl % g.nx + 1, (cl / g.nx) % g.ny + 1, cl / (g.nx * g.ny) + 1};
    const int n[3] = {g.nx, g.ny, g.nz};
    const int nt[3] = {tg.ntx, tg.nty, tg.ntz};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        int a = cc[d] - 1, b = cc[d] + 2;
        if (g.bt[2 * d] != IBMDEM_BC_PERIODIC) { a = a < 1 ? 1 : a; b = b > n[d] ? n[d] : b; }
        lo[d] = a; hi[d] = b;
    }
    for (int k = lo[2]; k <= hi[2]; ++k) {
        int kk = k < 1 ? k + n[2] : (k > n[2] ? k - n[2] : k);
        for (int j = lo[1]; j <= hi[1]; ++j) {
            int jj = j < 1 ? j + n[1] : (j > n[1] ? j - n[1] : j);
            for (int i = lo[0]; i <= hi[0]; ++i) {
                int ii = i < 1 ? i + n[0] : (i > n[0] ? i - n[0] : i);
                const long t = ((long)((kk - 1) / TL) * nt[1] + (jj - 1) / TL) * nt[0] + (ii - 1) / TL;
                flag[t] = 1;
            }
        }
    }
}
__global__ __launch_bounds__(T) void k_mark_tiles(int M, GridDev g, TileGeom tg, const int *__restrict__ mcell,
                                                  unsigned char *__restrict__ flag) {
    mark_tiles_body(hw_blk(), M, g, tg, mcell, flag);
}

// per-block counts of set flags (block b owns flags [b*CHUNK, (b+1)*CHUNK))
__device__ __forceinline__ void flag_count_body(const Blk &kb, long n, const unsigned char *__restrict__ flag,
                                                int *__restrict__ cnt) {
    __shared__ int red[T];
    const long lo = (long)kb.bx * CHUNK;
    int s = 0;
    for (long q = lo + kb.tx; q < min(n, lo + CHUNK); q += T) s += flag[q];
    red[kb.tx] = s;
    __syncthreads();
    for (int st = T / 2; st > 0; st >>= 1) {
        if (kb.tx < st) red[kb.tx] += red[kb.tx + st];
        __syncthreads();
    }
    if (kb.tx == 0) cnt[kb.bx] = red[0];
}
__global__ __launch_bounds__(T) void k_flag_count(long n, const unsigned char *__restrict__ flag, int *__restrict__ cnt) {
    flag_count_body(hw_blk(), n, flag, cnt);
}

// exclusive scan of nb ints in one block; total into a[nb] and *total
__device__ __forceinline__ void scan_small_body(const Blk &kb, int nb, int *__restrict__ a, int *__restrict__ total) {
    __shared__ int buf[1024];
    __shared__ int carry;
    const int t = kb.tx;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int q = base + t;
        int v = q < nb ? a[q] : 0;
        buf[t] = v;
        __syncthreads();
        for (int st = 1; st < 1024; st <<= 1) {
            int y = t >= st ? buf[t - st] : 0;
            __syncthreads();
            buf[t] += y;
            __syncthreads();
        }
        if (q < nb) a[q] = carry + buf[t] - v;
        __syncthreads();
        if (t == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (t == 0) { a[nb] = carry; *total = carry; }
}
__global__ __launch_bounds__(1024) void k_scan_small(int nb, int *__restrict__ a, int *__restrict__ total) {
    scan_small_body(hw_blk(), nb, a, total);
}

// ordered compaction of set flags into list; clears the flags
__device__ __forceinline__ void flag_compact_body(const Blk &kb, long n, unsigned char *__restrict__ flag,
                                                  const int *__restrict__ off, int *__restrict__ list) {
    __shared__ int buf[T];
    constexpr int PER = CHUNK / T;   // 16 consecutive flags per thread
    const int tx = kb.tx;
    const long lo = (long)kb.bx * CHUNK + (long)tx * PER;
    int s = 0;
    for (int q = 0; q < PER; ++q)
        if (lo + q < n) s += flag[lo + q];
    buf[tx] = s;
    __syncthreads();
    for (int st = 1; st < T; st <<= 1) {
        int t = tx >= st ? buf[tx - st] : 0;
        __syncthreads();
        buf[tx] += t;
        __syncthreads();
    }
    int pos = off[kb.bx] + buf[tx] - s;
    for (int q = 0; q < PER; ++q)
        if (lo + q < n && flag[lo + q]) {
            list[pos++] = (int)(lo + q);
            flag[lo + q] = 0;
        }
}
__global__ __launch_bounds__(T) void k_flag_compact(long n, unsigned char *__restrict__ flag,
                                                    const int *__restrict__ off, int *__restrict__ list) {
    flag_compact_body(hw_blk(), n, flag, off, list);
}

__global__ __launch_bounds__(T) void k_zero(long n, double *__restrict__ a) {
    for (long q = (long)blockIdx.x * T + threadIdx.x; q < n; q += (long)gridDim.x * T) a[q] = 0.0;
}

// one source-cell coordinate mapped to the interior with its image kind:
// 0 direct, 1 periodic (image = x + coordShift), 2 wall mirror
// (image = 2*coordMirror - x)
__device__ __forceinline__ bool src_map(int &i, int n, int lo, int hi, int &kind) {
    kind = 0;
    if (i < 1) {
        if (lo == IBMDEM_BC_PERIODIC) { i += n; kind = 1; return true; }
        i = 1 - i; kind = 2; return i >= 1 && i <= n;
    }
    if (i > n) {
        if (hi == IBMDEM_BC_PERIODIC) { i -= n; kind = 1; return true; }
        i = 2 * n + 1 - i; kind = 2; return i >= 1 && i <= n;
    }
    return true;
}

// 1-based unwrapped coordinates of source cell s of the tile with origin
// (ib, jb, kb) (the cell before the tile's first cell, per dimension)
__device__ __forceinline__ void src_coords(int s, int ib, int jb, int kb, int scc[3]) {
    scc[0] = ib + s % SRC;
    scc[1] = jb + (s / SRC) % SRC;
    scc[2] = kb + s / (SRC * SRC);
}

// image code of an unwrapped source cell, 3 bits per dimension: 0 the
// marker itself, 1/2 its periodic copy from below/above (x -/+ L), 3/4 its
// mirror copy below/above (-x, 2L - x)
__device__ __forceinline__ uint32_t image_code(const GridDev &g, const int scc[3]) {
    const int n3[3] = {g.nx, g.ny, g.nz};
    uint32_t code = 0;
    for (int d = 0; d < 3; ++d) {
        uint32_t c = 0;
        if (scc[d] < 1) c = g.bt[2 * d] == IBMDEM_BC_PERIODIC ? 1u : 3u;
        else if (scc[d] > n3[d]) c = g.bt[2 * d + 1] == IBMDEM_BC_PERIODIC ? 2u : 4u;
        code |= c << (3 * d);
    }
    return code;
}

// Per-source-cell bin ranges of a tile: first item of each of the NSRC
// source cells in s_gstart, the count prefix in s_off (s_off[NSRC] = total).
__device__ __forceinline__ void tile_bins(const GridDev &g, int ib, int jb, int kb, const int *__restrict__ bin_start,
                                          int *s_off, int *s_gstart) {
    const int lane = threadIdx.x;
    constexpr int SPL = (NSRC + 63) / 64;
    int sbin[SPL];
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        const int s = lane + 64 * r;
        sbin[r] = -1;
        if (s < NSRC) {
            int scc[3];
            src_coords(s, ib, jb, kb, scc);
            int ki, kj, kk;
            if (src_map(scc[0], g.nx, g.bt[0], g.bt[1], ki) && src_map(scc[1], g.ny, g.bt[2], g.bt[3], kj) &&
                src_map(scc[2], g.nz, g.bt[4], g.bt[5], kk))
                sbin[r] = (int)cidx(g, scc[0], scc[1], scc[2]);
        }
    }
    // all bin boundaries are loaded before any is used
    int sfirst[SPL], slast[SPL];
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        sfirst[r] = sbin[r] >= 0 ? bin_start[sbin[r]] : 0;
        slast[r] = sbin[r] >= 0 ? bin_start[sbin[r] + 1] : 0;
    }
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        const int s = lane + 64 * r;
        if (s < NSRC) {
            s_gstart[s] = sfirst[r];
            s_off[s + 1] = slast[r] - sfirst[r];
        }
    }
    __syncthreads();
    // prefix sum of the NSRC counts: PER consecutive per lane, then a
    // wavefront scan of the lane totals
    constexpr int PER = (NSRC + 63) / 64;
    const int b0 = lane * PER;
    int loc = 0;
    for (int q = 0; q < PER; ++q)
        if (b0 + q < NSRC) loc += s_off[b0 + q + 1];
    int inc = loc;
    for (int st = 1; st < 64; st <<= 1) {
        const int t = __shfl_up(inc, st, 64);
        if (lane >= st) inc += t;
    }
    int run = inc - loc;
    for (int q = 0; q < PER; ++q)
        if (b0 + q < NSRC) { run += s_off[b0 + q + 1]; s_off[b0 + q + 1] = run; }
    if (lane == 0) s_off[0] = 0;
    __syncthreads();
}

__device__ __forceinline__ void tile_origin(int tile, TileGeom tg, int &tx, int &ty, int &tz) {
    tx = tile % tg.ntx;
    ty = (tile / tg.ntx) % tg.nty;
    tz = tile / (tg.ntx * tg.nty);
}

// entries per active tile; slots nact..cap-1 get 0 (slot q is written by
// block q / 64 whatever that block's own tile is, so each slot has one writer)
// (64-thread blocks: tile_bins' lane is the thread index)
__device__ __forceinline__ void tile_count_body(const Blk &kb, const GridDev &g, TileGeom tg,
                                                const int *__restrict__ meta, const int *__restrict__ tiles,
                                                const int *__restrict__ bin_start, int *__restrict__ ecount,
                                                long cap) {
    __shared__ int s_off[NSRC + 1];
    __shared__ int s_gstart[NSRC];
    const int nact = meta[0];
    const long q = (long)kb.bx * 64 + kb.tx;
    if (q >= nact && q < cap) ecount[q] = 0;
    const int b = tile_slot(kb.bx, nact);
    if (b < 0) return;
    int tx, ty, tz;
    tile_origin(tiles[b], tg, tx, ty, tz);
    tile_bins(g, tx * TL - 1, ty * TL - 1, tz * TL - 1, bin_start, s_off, s_gstart);
    if (kb.tx == 0) ecount[b] = s_off[NSRC];
}
__global__ __launch_bounds__(64) void k_tile_count(GridDev g, TileGeom tg, const int *__restrict__ meta,
                                                   const int *__restrict__ tiles, const int *__restrict__ bin_start,
                                                   int *__restrict__ ecount, long cap) {
    tile_count_body(hw_blk(), g, tg, meta, tiles, bin_start, ecount, cap);
}

// write each active tile's entries at its scanned offset
__device__ __forceinline__ void tile_fill_body(const Blk &kb0, const GridDev &g, TileGeom tg,
                                               const int *__restrict__ meta, const int *__restrict__ tiles,
                                               const int *__restrict__ bin_start, const int *__restrict__ bin_items,
                                               const int *__restrict__ estart, uint32_t *__restrict__ ent,
                                               long ent_cap, int *__restrict__ flags) {
    __shared__ int s_off[NSRC + 1];
    __shared__ int s_gstart[NSRC];
    const int b = tile_slot(kb0.bx, meta[0]);
    if (b < 0) return;
    int tx, ty, tz;
    tile_origin(tiles[b], tg, tx, ty, tz);
    tile_bins(g, tx * TL - 1, ty * TL - 1, tz * TL - 1, bin_start, s_off, s_gstart);
    const long base = estart[b];
    if (base + s_off[NSRC] > ent_cap) {
        if (kb0.tx == 0) atomicOr(&flags[FL_SPREAD], 1);
        return;
    }
    const int ib = tx * TL - 1, jb = ty * TL - 1, kb = tz * TL - 1;
    // consecutive lanes write consecutive entries: entry e comes from the
    // last source cell s with s_off[s] <= e (a binary search over the LDS
    // offsets; empty cells share their offset with the next cell)
    const int total = s_off[NSRC];
    for (int e = kb0.tx; e < total; e += 64) {
        int lo = 0, hi = NSRC - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        int scc[3];
        src_coords(lo, ib, jb, kb, scc);
        ent[base + e] = ((uint32_t)bin_items[s_gstart[lo] + (e - s_off[lo])] << SBITS) | image_code(g, scc);
    }
}
__global__ __launch_bounds__(64) void k_tile_fill(GridDev g, TileGeom tg, const int *__restrict__ meta,
                                                  const int *__restrict__ tiles, const int *__restrict__ bin_start,
                                                  const int *__restrict__ bin_items, const int *__restrict__ estart,
                                                  uint32_t *__restrict__ ent, long ent_cap, int *__restrict__ flags) {
    tile_fill_body(hw_blk(), g, tg, meta, tiles, bin_start, bin_items, estart, ent, ent_cap, flags);
}

struct SpreadArgs {
    GridDev g;
    TileGeom tg;
    const int *meta;
    const int *tiles;
    const int *estart;
    const uint32_t *ent;
    long ent_cap;
    const double *rec;
    const double *fac;  // per marker: 12 (face, centre) factor pairs
    double *fact;      // 3 x (TC * cap_tiles), tile slot major
    long fstride;      // TC * cap_tiles
    unsigned long long *mask;
    double *u, *v, *w;
    double dV, idV, dt;
    int first;         // first MDF iteration of the step: f_old = 0
};

// The block is one wavefront: its lanes see each other's LDS writes in
// program order, so a compiler-level barrier replaces __syncthreads (which
// would also wait for the global loads the pipeline keeps in flight).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one lane's input for one staged entry. The packed entry id is kept raw
// until its records are loaded (one pipeline stage later), so issuing the
// load never waits for it.
struct EntryIn {
    double x0, x1, x2, f0, f1, f2;
    double2 fw[12]; // the marker's delta factors (k_records)
    int cell;       // the marker's cell (linear)
    uint32_t en;    // marker << SBITS | image code
    bool valid;     // the lane stages an entry
};

__device__ __forceinline__ void load_entry_id(const SpreadArgs &a, int e, bool valid, EntryIn &in) {
    in.valid = valid;
    in.en = valid ? a.ent[e] : 0u;
}
// The entry id is consumed here on every path, also where no lane of the
// wavefront is valid: the register then cannot look pending at the summing
// loop, which would otherwise wait for every load in flight before reusing
// it. (The empty asm only reads the register; it emits no instruction.)
__device__ __forceinline__ void load_entry_rec(const SpreadArgs &a, EntryIn &in) {
    asm volatile("" ::"v"(in.en));
    if (in.valid) {
        const long m = (long)(in.en >> SBITS);
        // each load has exactly the width of what is used: no dead
        // registers that the summing loop could reuse while a load is in
        // flight (that would make it wait for the next chunk's records)
        const double *r = a.rec + 8 * m;
        const double2 r01 = reinterpret_cast<const double2 *>(r)[0];
        const double2 r45 = reinterpret_cast<const double2 *>(r)[2];
        in.x0 = r01.x; in.x1 = r01.y; in.x2 = r[2];
        in.cell = reinterpret_cast<const int *>(r + 3)[0];   // low word of the cell slot
        in.f0 = r45.x; in.f1 = r45.y; in.f2 = r[6];
        const double2 *mf = reinterpret_cast<const double2 *>(a.fac + 24 * m);
#pragma unroll
        for (int q = 0; q < 12; ++q) in.fw[q] = mf[q];
    }
}

// What the summing loop needs of a staged entry that is the same for every
// lane: the support's corner relative to the tile (packed, 5 bits per
// dimension, biased by 8) and fdrag*vol. The staging lane keeps it in
// registers and the loop reads it with readlane, so it costs no LDS
// bandwidth.
struct EntryUni {
    int corner;
    double f0, f1, f2;
};

// delta factors of the lane's entry into LDS row `lane`; its cell and
// fdrag*vol into u
template <bool P2>
__device__ __forceinline__ void stage_entry(const SpreadArgs &a, const EntryIn &in, double2 (*s_w)[WST2],
                                            int i0, int j0, int k0, EntryUni &u) {
    const GridDev &g = a.g;
    const int lane = threadIdx.x;
    if (!in.valid) return;
    const uint32_t code = in.en & ((1u << SBITS) - 1);
    const double L3[3] = {g.L[0], g.L[1], g.L[2]};
    const double dd[3] = {g.dx, g.dy, g.dz};
    const double id3[3] = {1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz};
    const double xm[3] = {in.x0, in.x1, in.x2};
    const int cc[3] = {in.cell % g.nx + 1, (in.cell / g.nx) % g.ny + 1, in.cell / (g.nx * g.ny) + 1};
    // with a power-of-two cell volume the 1/dV of every weight is an exact
    // scaling, folded into fdrag*vol here once per entry (bitwise the same
    // products as fv * (w * (1/dV)))
    if (P2) {
        u.f0 = in.f0 * a.idV; u.f1 = in.f1 * a.idV; u.f2 = in.f2 * a.idV;
    } else {
        u.f0 = in.f0; u.f1 = in.f1; u.f2 = in.f2;
    }
    int cim[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t c = (code >> (3 * d)) & 7u;
        if (c == 0u) {
            // the marker itself: its factors from k_records
            cim[d] = cc[d];
#pragma unroll
            for (int o = 0; o < 4; ++o) s_w[lane][d * 4 + o] = in.fw[d * 4 + o];
            continue;
        }
        double xp = xm[d];
        if (c == 1u || c == 2u) xp = xp + (c == 1u ? -L3[d] : L3[d]);
        else xp = 2.0 * (c == 3u ? 0.0 : L3[d]) - xp;
        const int cimg = cell_of(xp, dd[d]);
        cim[d] = cimg;
        // the centre factor at offset +2 is always 0 (see k_records)
        for (int o = 0; o < 4; ++o)
            s_w[lane][d * 4 + o] = make_double2(wt_face<P2>(xp, cimg + o - 1, dd[d], id3[d]),
                                                o < 3 ? wt_cent<P2>(xp, cimg + o - 1, dd[d], id3[d]) : 0.0);
    }
    // support cells cim - 1 .. cim + 2 reach the tile, so each corner offset
    // lies in [-3, 7]; clamping to [-4, 8] changes nothing (such a support
    // misses the tile either way) and keeps the packed fields in range
    const int o3[3] = {i0, j0, k0};
    u.corner = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) u.corner |= (min(max(cim[d] - 1 - o3[d], -4), 8) + 8) << (5 * d);
}

// One wavefront per active tile (block b takes tile slot tile_slot(b)).
// LDS position of tile cell (x, y, z): z planes 68 doubles apart, so the 32
// lanes of a half-wavefront (support offsets 4 x 4 x 2) hit 32 different
// 8-byte bank pairs
constexpr int ACC_Z = TL * TL + 4, ACC_N = ACC_Z * TL;
// s += v as one LDS add without return (ds_add_f64: round to nearest even,
// the same result as s - p for v = -p). The LDS executes a wavefront's
// operations in issue order, so successive adds to one cell land in program
// order and the lane does not wait for the old value.
__device__ __forceinline__ void lds_add(double *s, double v) {
    __hip_atomic_fetch_add(s, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ int acc_idx(int q) { return q + 4 * (q >> 6); }
// lane e's double, on every lane (e uniform)
__device__ __forceinline__ double readlane_f64(double v, int e) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, e);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), e);
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

// End of a tile's spread: store the reached-cell mask (first iteration),
// then the multi-direct-forcing update u += dt*(f - f_old) on the reached
// cells: the lane's 8 cells (one per z plane of the tile) in two passes,
// every load of a pass issued before any is used.
__device__ __forceinline__ void tile_epilogue(const SpreadArgs &a, int slot, double (*s_acc)[ACC_N],
                                              const unsigned *s_touch, int i0, int j0, int k0) {
    const GridDev &g = a.g;
    const int lane = threadIdx.x;
    if (a.first && lane < IBM_MASK_WORDS)
        a.mask[(long)slot * IBM_MASK_WORDS + lane] =
            ((unsigned long long)s_touch[2 * lane + 1] << 32) | (unsigned long long)s_touch[2 * lane];
    // multi-direct-forcing update u += dt*(f - f_old) on the reached cells:
    // the lane's 8 cells (one per z plane of the tile) in two passes, every
    // load of a pass issued before any is used
    double *vel[3] = {a.u, a.v, a.w};
    const long fbase = (long)slot * TC;
    const int li = lane & (TL - 1), lj = lane >> 3;
    const int i = i0 + li, j = j0 + lj;
    constexpr int HALF = TC / 64 / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        bool has[HALF];
        double fo[HALF][3], vo[HALF][3];
        long o[HALF];
#pragma unroll
        for (int r = 0; r < HALF; ++r) {
            const int q = lane + 64 * (h * HALF + r), k = k0 + h * HALF + r;
            has[r] = ((s_touch[q >> 5] >> (q & 31)) & 1u) && i <= g.nx && j <= g.ny && k <= g.nz;
            o[r] = has[r] ? cidx(g, i, j, k) : 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                fo[r][c] = (has[r] && !a.first) ? a.fact[c * a.fstride + fbase + q] : 0.0;
                vo[r][c] = has[r] ? vel[c][o[r]] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < HALF; ++r) {
            if (!has[r]) continue;
            const int q = lane + 64 * (h * HALF + r), k = k0 + h * HALF + r;
            const int tgt[3] = {i, j, k};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double fn = s_acc[c][acc_idx(q)];
                a.fact[c * a.fstride + fbase + q] = fn;
                // the low boundary faces (wall, inflow) keep their fixed values
                const bool face = tgt[c] == 1 && g.bt[2 * c] != IBMDEM_BC_PERIODIC;
                if (!face && !(fn == 0.0 && fo[r][c] == 0.0)) vel[c][o[r]] = vo[r][c] + a.dt * (fn - fo[r][c]);
            }
        }
    }
}

template <bool P2>
__device__ __forceinline__ void spread_tiles_body(const Blk &kb, const SpreadArgs &a) {
    __shared__ double s_acc[3][ACC_N];         // the tile's force
    __shared__ unsigned s_touch[TC / 32];      // cells some entry reached
    // per staged entry: (face, centre) factor pairs for 3 dimensions x 4
    // offsets (padded to an odd number of 16-byte pairs per entry); its
    // support corner and fdrag*vol stay in the staging lane's registers
    __shared__ double2 s_w[CAP][WST2];
    const GridDev &g = a.g;
    const int lane = kb.tx;
    const int slot = tile_slot(kb.bx, a.meta[0]);
    if (slot < 0) return;
    const int elo = a.estart[slot], ehi = a.estart[slot + 1];
    if (ehi > a.ent_cap) return;   // the fill flagged a capacity error
    int tx, ty, tz;
    tile_origin(a.tiles[slot], a.tg, tx, ty, tz);
    const int i0 = tx * TL + 1, j0 = ty * TL + 1, k0 = tz * TL + 1;
    for (int q = lane; q < ACC_N; q += 64) {
        s_acc[0][q] = 0.0;
        s_acc[1][q] = 0.0;
        s_acc[2][q] = 0.0;
    }
    // the reached cells are found in the first iteration of a step; the
    // markers (so the entries and their supports) do not move until the next
    if (lane < TC / 32)
        s_touch[lane] = a.first ? 0u : (unsigned)(a.mask[(long)slot * IBM_MASK_WORDS + (lane >> 1)] >> (32 * (lane & 1)));
    // the lane's offset in an entry's 4x4x4 support
    const int oa = lane & 3, ob = (lane >> 2) & 3, oc = lane >> 4;
    const int loff = oa + TL * ob + ACC_Z * oc;   // its accumulator offset from the support corner
    EntryUni uni{};
    wave_sync();
    // Two-stage prefetch: while chunk c is summed, the records and factors
    // of chunk c+1 and the entry ids of chunk c+2 are in flight.
    EntryIn in, nx;
    load_entry_id(a, elo + lane, lane < min(CAP, ehi - elo), in);
    load_entry_rec(a, in);
    load_entry_id(a, elo + CAP + lane, lane < min(CAP, ehi - elo - CAP), nx);
    for (int e0 = elo; e0 < ehi; e0 += CAP) {
        const int ne = min(CAP, ehi - e0);
        stage_entry<P2>(a, in, s_w, i0, j0, k0, uni);
        wave_sync();
        in = nx;
        load_entry_rec(a, in);
        load_entry_id(a, e0 + 2 * CAP + lane, lane < min(CAP, ehi - e0 - 2 * CAP), nx);
        // SU entries at a time: their factor rows are read and their terms
        // formed first, then their adds are issued back to back in entry
        // order (rows at or past ne hold stale data and are not added)
        for (int g0 = 0; g0 < ne; g0 += SU) {
            double t0[SU], t1[SU], t2[SU];
            int qa[SU], qt[SU];
            bool hit[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int e = g0 + u;
                // the entry's uniform values from its staging lane (e is the
                // same on every lane, so these are scalars)
                const int cp = __builtin_amdgcn_readlane(uni.corner, e);
                const int bx = (cp & 31) - 8, by = ((cp >> 5) & 31) - 8, bz = ((cp >> 10) & 31) - 8;
                const int lx = bx + oa, ly = by + ob, lz = bz + oc;
                hit[u] = e < ne && (unsigned)lx < (unsigned)TL && (unsigned)ly < (unsigned)TL &&
                         (unsigned)lz < (unsigned)TL;
                qa[u] = (bx + TL * by + ACC_Z * bz) + loff;
                qt[u] = lx + TL * (ly + TL * lz);
                const double f0 = readlane_f64(uni.f0, e), f1 = readlane_f64(uni.f1, e),
                             f2 = readlane_f64(uni.f2, e);
                const double2 wx = s_w[e][oa], wy = s_w[e][4 + ob], wz = s_w[e][8 + oc];
                // (face, centre): u takes the x face, v the y face, w the z face
                const double pu = wx.x * wy.y * wz.y, pv = wx.y * wy.x * wz.y, pw = wx.y * wy.y * wz.x;
                const double wu = P2 ? pu : pu / a.dV, wv = P2 ? pv : pv / a.dV, ww = P2 ? pw : pw / a.dV;
                t0[u] = f0 * wu;
                t1[u] = f1 * wv;
                t2[u] = f2 * ww;
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                if (!hit[u]) continue;
                lds_add(&s_acc[0][qa[u]], -t0[u]);
                lds_add(&s_acc[1][qa[u]], -t1[u]);
                lds_add(&s_acc[2][qa[u]], -t2[u]);
                if (a.first) atomicOr(&s_touch[qt[u] >> 5], 1u << (qt[u] & 31));
            }
        }
        wave_sync();   // the rows are rewritten by the next chunk
    }
    tile_epilogue(a, slot, s_acc, s_touch, i0, j0, k0);
}
template <bool P2>
__global__ __launch_bounds__(64) void k_spread_tiles(SpreadArgs a) {
    spread_tiles_body<P2>(hw_blk(), a);
}

__global__ __launch_bounds__(64) void k_scatter_force(const int *__restrict__ meta, const int *__restrict__ tiles,
                                                      TileGeom tg, GridDev g, const double *__restrict__ fact,
                                                      long fstride, const unsigned long long *__restrict__ mask,
                                                      double *__restrict__ fx, double *__restrict__ fy,
                                                      double *__restrict__ fz) {
    const int b = tile_slot(blockIdx.x, meta[0]);
    if (b < 0) return;
    int tx, ty, tz;
    tile_origin(tiles[b], tg, tx, ty, tz);
    for (int q = threadIdx.x; q < TC; q += 64) {
        if (!((mask[(long)b * IBM_MASK_WORDS + q / 64] >> (q & 63)) & 1ull)) continue;
        const int i = tx * TL + 1 + (q & (TL - 1)), j = ty * TL + 1 + ((q / TL) & (TL - 1)),
                  k = tz * TL + 1 + q / (TL * TL);
        if (i > g.nx || j > g.ny || k > g.nz) continue;
        const long slot = (long)b * TC + q, o = cidx(g, i, j, k);
        fx[o] = fact[slot];
        fy[o] = fact[fstride + slot];
        fz[o] = fact[2 * fstride + slot];
    }
}

// The IBM force of the cells with k in [k0, k1) (1-based) into a slab
// buffer: component q of cell (i, j, k) at q * slab + cidx(i, j, k) - (k0 - 1)
// * nx * ny; cells outside the slab are skipped.
__global__ __launch_bounds__(64) void k_scatter_force_slab(const int *__restrict__ meta, const int *__restrict__ tiles,
                                                           TileGeom tg, GridDev g, const double *__restrict__ fact,
                                                           long fstride, const unsigned long long *__restrict__ mask,
                                                           int k0, int k1, long slab, double *__restrict__ out) {
    const int b = tile_slot(blockIdx.x, meta[0]);
    if (b < 0) return;
    int tx, ty, tz;
    tile_origin(tiles[b], tg, tx, ty, tz);
    if (tz * TL + TL < k0 || tz * TL + 1 >= k1) return;   // the tile misses the slab
    const long first = (long)(k0 - 1) * g.nx * g.ny;
    for (int q = threadIdx.x; q < TC; q += 64) {
        if (!((mask[(long)b * IBM_MASK_WORDS + q / 64] >> (q & 63)) & 1ull)) continue;
        const int i = tx * TL + 1 + (q & (TL - 1)), j = ty * TL + 1 + ((q / TL) & (TL - 1)),
                  k = tz * TL + 1 + q / (TL * TL);
        if (i > g.nx || j > g.ny || k < k0 || k >= k1) continue;
        const long sl = (long)b * TC + q, o = cidx(g, i, j, k) - first;
        out[o] = fact[sl];
        out[slab + o] = fact[fstride + sl];
        out[2 * slab + o] = fact[2 * fstride + sl];
    }
}

// cells with at least one contribution, summed over the active tiles
__global__ __launch_bounds__(1024) void k_mask_count(const int *__restrict__ meta,
                                                     const unsigned long long *__restrict__ mask,
                                                     long long *__restrict__ out) {
    __shared__ long long red[1024];
    const int nact = meta[0];
    long long s = 0;
    for (long b = threadIdx.x; b < (long)nact * IBM_MASK_WORDS; b += 1024) s += __popcll(mask[b]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

TileGeom tile_geom(const GridDev &g) { return TileGeom{div_up(g.nx, TL), div_up(g.ny, TL), div_up(g.nz, TL)}; }

// dx, dy, dz and their product are powers of two: every division by them
// is an exact multiplication by the reciprocal
bool pow2_grid(const ibmdem_ctx *c) {
    return is_pow2(c->g.dx) && is_pow2(c->g.dy) && is_pow2(c->g.dz) && is_pow2(c->dV);
}

}  // namespace

long ibm_tile_count(const GridDev &g) {
    TileGeom t = tile_geom(g);
    return (long)t.ntx * t.nty * t.ntz;
}

int ibm_begin_step(ibmdem_ctx *c) {
    const int M = c->M;
    c->first_mdf = true;
    c->spread_valid = false;
    if (M == 0) return IBMDEM_OK;
    hipStream_t s = c->stream;
    const TileGeom tg = tile_geom(c->g);
    const long nt = (long)tg.ntx * tg.nty * tg.ntz;
    if (pow2_grid(c))
        hipLaunchKernelGGL(k_records<true>, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mx, c->mcell, c->mrec,
                           c->mfac, c->fdrag);
    else
        hipLaunchKernelGGL(k_records<false>, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mx, c->mcell, c->mrec,
                           c->mfac, c->fdrag);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_mark_tiles, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, tg, c->mcell, c->act_flag);
    CHECK_LAUNCH(c);
    const int nb = div_up(nt, CHUNK);
    int *cnt = c->scan_blocks;   // sized >= n/CHUNK + 2 at create
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(T), 0, s, nt, c->act_flag, cnt);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, nb, cnt, c->act_meta);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_flag_compact, dim3(nb), dim3(T), 0, s, nt, c->act_flag, cnt, c->act_list);
    CHECK_LAUNCH(c);
    // the active count stays on the device: every tile kernel is launched
    // act_grid wide and its blocks beyond the count return at once
    hipLaunchKernelGGL(k_tile_count, dim3(c->act_grid), dim3(64), 0, s, c->g, tg, c->act_meta, c->act_list,
                       c->bin_start, c->tile_ecount, c->act_cap);
    CHECK_LAUNCH(c);
    int rc = scan_counts(c, c->tile_ecount, c->tile_estart, c->act_cap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tile_fill, dim3(c->act_grid), dim3(64), 0, s, c->g, tg, c->act_meta, c->act_list,
                       c->bin_start, c->bin_items, c->tile_estart, c->tile_ent, c->ent_cap, c->flags);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_interpolate(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    prof_begin(c, IBMDEM_K_INTERP);
#define LAUNCH_INTERP(P2, OP)                                                                                      \
    hipLaunchKernelGGL((k_interp<P2, OP>), dim3(8 * div_up(div_up(c->M, T), 8)), dim3(T), 0, c->stream, c->M, c->g, \
                       c->bin_items, c->u[cur],                                                                       \
                       c->v[cur], c->w[cur], c->mrec, c->mv, c->fdrag, c->mvol, c->dV, 1.0 / c->dV, c->cs.dt)
    const bool open = grid_open(c->g);
    if (pow2_grid(c)) { if (open) LAUNCH_INTERP(true, true); else LAUNCH_INTERP(true, false); }
    else { if (open) LAUNCH_INTERP(false, true); else LAUNCH_INTERP(false, false); }
#undef LAUNCH_INTERP
    // compulsory bytes: 72 per marker (its position, velocity and the
    // updated fdrag) and one read of each footprint velocity (cell,
    // component); the per-step delta-factor table is not compulsory
    prof_end(c, IBMDEM_K_INTERP, 72.0 * c->M + 8.0 * 3.0 * (double)c->footprint_cells);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_spread(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    SpreadArgs a;
    a.g = c->g;
    a.tg = tile_geom(c->g);
    a.meta = c->act_meta;
    a.tiles = c->act_list;
    a.estart = c->tile_estart;
    a.ent = c->tile_ent;
    a.ent_cap = c->ent_cap;
    a.rec = c->mrec;
    a.fac = c->mfac;
    a.fact = c->fact;
    a.fstride = (long)TC * c->act_cap;
    a.mask = c->tile_mask;
    a.u = c->u[cur]; a.v = c->v[cur]; a.w = c->w[cur];
    a.dV = c->dV;
    a.idV = 1.0 / c->dV;
    a.dt = c->cs.dt;
    a.first = c->first_mdf ? 1 : 0;
    prof_begin(c, IBMDEM_K_SPREAD);
    const dim3 grid(c->act_grid);
    if (pow2_grid(c))
        hipLaunchKernelGGL((k_spread_tiles<true>), grid, dim3(64), 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_spread_tiles<false>), grid, dim3(64), 0, c->stream, a);
    prof_end(c, IBMDEM_K_SPREAD, 0.0);
    CHECK_LAUNCH(c);
    if (c->prof.kernel == IBMDEM_K_SPREAD || c->prof.kernel == IBMDEM_K_INTERP) {
        if (a.first) {
            // profiling only: count the written cells once per step (host wait)
            long long *cnt = reinterpret_cast<long long *>(c->act_meta + 2);
            hipLaunchKernelGGL(k_mask_count, dim3(1), dim3(1024), 0, c->stream, c->act_meta, c->tile_mask, cnt);
            CHECK_LAUNCH(c);
            long long h = 0;
            HIPX(c, hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
            HIPX(c, hipStreamSynchronize(c->stream));
            c->footprint_cells = (long)h;
        }
        // compulsory bytes: 60 per marker (position, fdrag, volume, cell) and
        // per written (cell, component) u read and written, f written and
        // f_old read (not in the first iteration); the per-step delta-factor
        // table and the records' re-reads by neighbouring tiles are not
        const double per = a.first ? 24.0 : 32.0;
        if (c->prof.kernel == IBMDEM_K_SPREAD)
            c->prof.bytes_total += 60.0 * c->M + per * 3.0 * (double)c->footprint_cells;
    }
    c->first_mdf = false;
    c->spread_valid = true;
    return IBMDEM_OK;
}

int ibm_force_field(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    hipStream_t s = c->stream;
    const long n = c->n;
    double *f[3] = {fx, fy, fz};
    for (int q = 0; q < 3; ++q) {
        hipLaunchKernelGGL(k_zero, dim3(4096), dim3(T), 0, s, n, f[q]);
        CHECK_LAUNCH(c);
    }
    if (c->M > 0 && c->spread_valid) {
        hipLaunchKernelGGL(k_scatter_force, dim3(c->act_grid), dim3(64), 0, s, c->act_meta, c->act_list,
                           tile_geom(c->g), c->g, c->fact, (long)TC * c->act_cap, c->tile_mask, fx, fy, fz);
        CHECK_LAUNCH(c);
    }
    return IBMDEM_OK;
}

int ibm_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    const long n = c->n;
    double *buf = nullptr;
    HIPX(c, hipMalloc((void **)&buf, 3 * n * sizeof(double)));
    hipStream_t s = c->stream;
    int rc = IBMDEM_OK;
    if (hipMemsetAsync(buf, 0, 3 * n * sizeof(double), s) != hipSuccess) rc = IBMDEM_ERR_HIP;
    if (!rc && c->M > 0 && c->spread_valid) {
        hipLaunchKernelGGL(k_scatter_force, dim3(c->act_grid), dim3(64), 0, s, c->act_meta, c->act_list,
                           tile_geom(c->g), c->g, c->fact, (long)TC * c->act_cap, c->tile_mask, buf, buf + n, buf + 2 * n);
        if (hipGetLastError() != hipSuccess) rc = IBMDEM_ERR_HIP;
    }
    double *dst[3] = {fx, fy, fz};
    for (int q = 0; q < 3 && !rc; ++q)
        if (dst[q] && hipMemcpyAsync(dst[q], buf + q * n, n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = IBMDEM_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) rc = IBMDEM_ERR_HIP;
    (void)hipFree(buf);
    return rc ? set_error(c, rc, "force download failed") : IBMDEM_OK;
}

// ---------------------------------------------------------------------------
// ibmdem_download_force without device allocation. ibmdem_download_force
// stages the force fields in a device buffer it allocates per call; this
// variant stages them, one slab of z planes at a time, in a fixed scratch
// array of IBM_SCRATCH_DOUBLES doubles (16 MiB) that every GPU holds from the
// first call until the process ends. The host code never uses the array's
// name as an address: the per-device address comes from hipGetSymbolAddress.
// Calls on one GPU share its array, so calls that overlap on one GPU (two
// threads, two streams) race; each slab is cleared by a fill kernel before
// use, so no call sees another's contents.
// ---------------------------------------------------------------------------
constexpr long IBM_SCRATCH_DOUBLES = 1L << 21;
static __device__ __align__(256) double scratch[IBM_SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mutex;
double *scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// the scratch array's address on the context's device (looked up once per
// device, with that device current)
int scratch_on_device(ibmdem_ctx *c, double **out) {
    if (c->device < 0 || c->device >= SCRATCH_MAX_DEVICES)
        return set_error(c, IBMDEM_ERR_UNSUPPORTED, "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[c->device]) {
        int cur = -1;
        HIPX(c, hipGetDevice(&cur));
        if (cur != c->device) HIPX(c, hipSetDevice(c->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != c->device) HIPX(c, hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return set_error(c, IBMDEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[c->device] = static_cast<double *>(p);
    }
    *out = scratch_ptr[c->device];
    return IBMDEM_OK;
}

__global__ __launch_bounds__(256) void k_fill_zero(long n, double *__restrict__ a) {
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) a[q] = 0.0;
}
}  // namespace

extern "C" int ibmdem_download_force_noalloc(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    if (!c) return IBMDEM_ERR_ARG;
    const long plane = (long)c->g.nx * c->g.ny;
    const long per = IBM_SCRATCH_DOUBLES / 3;   // cells per slab
    if (plane > per)
        return set_error(c, IBMDEM_ERR_UNSUPPORTED,
                         "a z plane of " + std::to_string(plane) + " cells exceeds the no-allocation scratch (at most " +
                             std::to_string(per) + " cells per plane)");
    double *buf = nullptr;
    int rc = scratch_on_device(c, &buf);
    if (rc) return rc;
    const int nz = c->g.nz;
    const int kstep = (int)std::min<long>(nz, per / plane);
    hipStream_t s = c->stream;
    double *dst[3] = {fx, fy, fz};
    for (int k0 = 1; k0 <= nz; k0 += kstep) {
        const int k1 = std::min(nz + 1, k0 + kstep);
        const long slab = (long)(k1 - k0) * plane;
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<long>(div_up(3 * slab, 256), 2048)), dim3(256), 0, s,
                           3 * slab, buf);
        CHECK_LAUNCH(c);
        if (c->M > 0 && c->spread_valid) {
            hipLaunchKernelGGL(k_scatter_force_slab, dim3(c->act_grid), dim3(64), 0, s, c->act_meta, c->act_list,
                               tile_geom(c->g), c->g, c->fact, (long)TC * c->act_cap, c->tile_mask, k0, k1, slab, buf);
            CHECK_LAUNCH(c);
        }
        for (int q = 0; q < 3; ++q)
            if (dst[q])
                HIPX(c, hipMemcpyAsync(dst[q] + (long)(k0 - 1) * plane, buf + q * slab, slab * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
        // the next slab reuses the scratch: its copies must have landed
        HIPX(c, hipStreamSynchronize(s));
    }
    return IBMDEM_OK;
}

// ---------------------------------------------------------------------------
// The multi-direct-forcing loop of a batched step: the n_mdf iterations of
// interpolation and spread of every case of a group, each as one launch per
// kernel variant (power-of-two spacing, open boundaries) over all cases.
// Both kernels map their blocks to XCDs themselves (tile_slot), so a case's
// physical blocks are its own launch's blocks, starting on XCD 0.
// ---------------------------------------------------------------------------
namespace {
struct InterpArg {
    BGrid g;
    int M;
    GridDev grid;
    const int *order;
    const double *u, *v, *w;
    double *rec;
    const double *mv;
    double *fdrag;
    const double *mvol;
    double dV, idV, dt;
};
template <bool P2, bool OPEN>
__global__ __launch_bounds__(T) void kb_interp(BTable<InterpArg> tab) {
    batch_for(tab, [&](const InterpArg &a, const Blk &kb) {
        interp_body<P2, OPEN>(kb, a.M, a.grid, glb(a.order), glb(a.u), glb(a.v), glb(a.w), glb(a.rec), glb(a.mv),
                              glb(a.fdrag), glb(a.mvol), a.dV, a.idV, a.dt);
    });
}
struct SpreadArg {
    BGrid g;
    SpreadArgs s;
};
// uniform argument: first (the step's first iteration)
template <bool P2>
__global__ __launch_bounds__(64) void kb_spread(BTable<SpreadArg> tab, int first) {
    batch_for(tab, [&](const SpreadArg &a, const Blk &kb) {
        SpreadArgs s = a.s;
        s.meta = glb(s.meta); s.tiles = glb(s.tiles); s.estart = glb(s.estart); s.ent = glb(s.ent);
        s.rec = glb(s.rec); s.fac = glb(s.fac); s.fact = glb(s.fact); s.mask = glb(s.mask);
        s.u = glb(s.u); s.v = glb(s.v); s.w = glb(s.w);
        s.first = first;
        spread_tiles_body<P2>(kb, s);
    });
}
template <bool P2, bool OPEN>
void launch_interp(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL((kb_interp<P2, OPEN>), dim3(o.blocks), dim3(T), 0, s, table<InterpArg>(o));
}
template <bool P2>
void launch_spread(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_spread<P2>, dim3(o.blocks), dim3(64), 0, s, table<SpreadArg>(o), o.slot);
}

SpreadArgs spread_args(ibmdem_ctx *c) {
    const int cur = c->vcur;
    SpreadArgs a;
    a.g = c->g;
    a.tg = tile_geom(c->g);
    a.meta = c->act_meta;
    a.tiles = c->act_list;
    a.estart = c->tile_estart;
    a.ent = c->tile_ent;
    a.ent_cap = c->ent_cap;
    a.rec = c->mrec;
    a.fac = c->mfac;
    a.fact = c->fact;
    a.fstride = (long)TC * c->act_cap;
    a.mask = c->tile_mask;
    a.u = c->u[cur]; a.v = c->v[cur]; a.w = c->w[cur];
    a.dV = c->dV;
    a.idV = 1.0 / c->dV;
    a.dt = c->cs.dt;
    a.first = c->first_mdf ? 1 : 0;
    return a;
}
}  // namespace

// the n_mdf iterations of ibm_interpolate and ibm_spread of every context
// with markers (all share n_mdf; each at its current velocity parity)
int ibm_plan_mdf(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<InterpArg> ia[4];   // by (pow2, open)
    std::vector<SpreadArg> sa[2];   // by pow2
    for (ibmdem_ctx *c : cs) {
        if (c->M == 0) continue;
        const int cur = c->vcur;
        const bool p2 = pow2_grid(c), open = grid_open(c->g);
        InterpArg a{};
        a.g = bgrid(dim3(8 * div_up(div_up(c->M, T), 8)), dim3(T));
        a.M = c->M;
        a.grid = c->g;
        a.order = c->bin_items;
        a.u = c->u[cur]; a.v = c->v[cur]; a.w = c->w[cur];
        a.rec = c->mrec;
        a.mv = c->mv;
        a.fdrag = c->fdrag;
        a.mvol = c->mvol;
        a.dV = c->dV;
        a.idV = 1.0 / c->dV;
        a.dt = c->cs.dt;
        ia[(p2 ? 2 : 0) + (open ? 1 : 0)].push_back(a);
        SpreadArg b{};
        b.g = bgrid(dim3(c->act_grid), dim3(64));
        b.s = spread_args(c);
        sa[p2 ? 1 : 0].push_back(b);
    }
    const BLaunch fi[4] = {launch_interp<false, false>, launch_interp<false, true>, launch_interp<true, false>,
                           launch_interp<true, true>};
    const BLaunch fs[2] = {launch_spread<false>, launch_spread<true>};
    const int n_mdf = cs.empty() ? 0 : cs[0]->cs.n_mdf;
    int rc;
    for (int it = 0; it < n_mdf; ++it) {
        for (int v = 0; v < 4; ++v)
            if ((rc = add_op(P.ar, P.ops, fi[v], T, ia[v], BOp(), err, true))) return rc;
        BOp o;
        o.slot = it == 0 ? 1 : 0;   // first
        for (int v = 0; v < 2; ++v)
            if ((rc = add_op(P.ar, P.ops, fs[v], 64, sa[v], o, err, true))) return rc;
    }
    return IBMDEM_OK;
}

// ---------------------------------------------------------------------------
// The IBM set-up of a batched step (ibm_begin_step of every case of a group):
// marker records, active tiles and their entry lists, each kernel one launch
// over the cases with markers. A case's arguments are one IbmSetup in the
// plan's arena; each launch's table holds the case's grid of blocks and its
// index there. The tile kernels map their blocks to XCDs themselves
// (tile_slot), so their physical blocks are the case's own.
// ---------------------------------------------------------------------------
namespace {
struct IbmSetup {
    int M, nb, enb;   // markers, flag chunks, entry-count chunks
    long nt, cap, ent_cap;
    GridDev g;
    TileGeom tg;
    const double *mx;
    const int *mcell, *bin_start, *bin_items;
    double *rec, *fac, *fdrag;
    unsigned char *flag;
    int *cnt, *meta, *list, *ecount, *estart, *flags;
    uint32_t *ent;
};
struct SArg {
    BGrid g;
    int q;   // the case's IbmSetup
};
using STab = BTable<SArg>;

template <bool P2>
__global__ __launch_bounds__(T) void kb_records(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        records_body<P2>(kb, c.M, c.g, glb(c.mx), glb(c.mcell), glb(c.rec), glb(c.fac), glb(c.fdrag));
    });
}
__global__ __launch_bounds__(T) void kb_mark_tiles(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        mark_tiles_body(kb, c.M, c.g, c.tg, glb(c.mcell), glb(c.flag));
    });
}
__global__ __launch_bounds__(T) void kb_flag_count(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        flag_count_body(kb, c.nt, glb((const unsigned char *)c.flag), glb(c.cnt));
    });
}
__global__ __launch_bounds__(1024) void kb_scan_small(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        scan_small_body(kb, c.nb, glb(c.cnt), glb(c.meta));
    });
}
__global__ __launch_bounds__(T) void kb_flag_compact(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        flag_compact_body(kb, c.nt, glb(c.flag), glb((const int *)c.cnt), glb(c.list));
    });
}
__global__ __launch_bounds__(64) void kb_tile_count(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        tile_count_body(kb, c.g, c.tg, glb((const int *)c.meta), glb((const int *)c.list), glb(c.bin_start),
                        glb(c.ecount), c.cap);
    });
}
__global__ __launch_bounds__(T) void kb_ecount_sum(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        scan::chunk_sum_body(kb, c.cap, glb((const int *)c.ecount), glb(c.cnt));
    });
}
__global__ __launch_bounds__(1024) void kb_ecount_scan_blocks(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        scan::scan_blocks_body(kb, c.enb, glb(c.cnt));
    });
}
__global__ __launch_bounds__(T) void kb_ecount_scan(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        scan::chunk_scan_body(kb, c.cap, glb((const int *)c.ecount), glb((const int *)c.cnt), glb(c.estart), c.enb);
    });
}
__global__ __launch_bounds__(64) void kb_tile_fill(STab tab, const IbmSetup *cs) {
    batch_for(tab, [&](const SArg &a, const Blk &kb) {
        const IbmSetup &c = *cst(cs + a.q);
        tile_fill_body(kb, c.g, c.tg, glb((const int *)c.meta), glb((const int *)c.list), glb(c.bin_start),
                       glb(c.bin_items), glb((const int *)c.estart), glb(c.ent), c.ent_cap, glb(c.flags));
    });
}

#define IBM_SETUP_LAUNCHER(name, kernel, threads)                                                     \
    void launch_##name(const BOp &o, hipStream_t s, int) {                                           \
        hipLaunchKernelGGL(kernel, dim3(o.blocks), dim3(threads), 0, s, table<SArg>(o),               \
                           (const IbmSetup *)o.aux);                                                 \
    }
IBM_SETUP_LAUNCHER(records_p2, kb_records<true>, T)
IBM_SETUP_LAUNCHER(records, kb_records<false>, T)
IBM_SETUP_LAUNCHER(mark_tiles, kb_mark_tiles, T)
IBM_SETUP_LAUNCHER(flag_count, kb_flag_count, T)
IBM_SETUP_LAUNCHER(scan_small, kb_scan_small, 1024)
IBM_SETUP_LAUNCHER(flag_compact, kb_flag_compact, T)
IBM_SETUP_LAUNCHER(tile_count, kb_tile_count, 64)
IBM_SETUP_LAUNCHER(ecount_sum, kb_ecount_sum, T)
IBM_SETUP_LAUNCHER(ecount_scan_blocks, kb_ecount_scan_blocks, 1024)
IBM_SETUP_LAUNCHER(ecount_scan, kb_ecount_scan, T)
IBM_SETUP_LAUNCHER(tile_fill, kb_tile_fill, 64)
#undef IBM_SETUP_LAUNCHER
}  // namespace

// ibm_begin_step of every context with markers, in its order
int ibm_plan_begin_step(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<IbmSetup> v;
    std::vector<int> idx;   // the contexts with markers
    for (int q = 0; q < (int)cs.size(); ++q) {
        ibmdem_ctx *c = cs[q];
        if (c->M == 0) continue;
        IbmSetup d{};
        d.M = c->M;
        d.g = c->g;
        d.tg = tile_geom(c->g);
        d.nt = (long)d.tg.ntx * d.tg.nty * d.tg.ntz;
        d.nb = div_up(d.nt, CHUNK);
        d.cap = c->act_cap;
        d.enb = div_up(c->act_cap, CHUNK);
        if (d.enb + 1 > c->scan_nblocks) {
            err = "batched IBM set-up: scan longer than its block buffer";
            return IBMDEM_ERR_CAPACITY;
        }
        d.ent_cap = c->ent_cap;
        d.mx = c->mx;
        d.mcell = c->mcell; d.bin_start = c->bin_start; d.bin_items = c->bin_items;
        d.rec = c->mrec; d.fac = c->mfac; d.fdrag = c->fdrag;
        d.flag = c->act_flag;
        d.cnt = c->scan_blocks; d.meta = c->act_meta; d.list = c->act_list;
        d.ecount = c->tile_ecount; d.estart = c->tile_estart; d.flags = c->flags;
        d.ent = c->tile_ent;
        idx.push_back(q);
        v.push_back(d);
    }
    if (v.empty()) return IBMDEM_OK;
    const size_t aux = P.ar.put(v.data(), v.size() * sizeof(IbmSetup));
    // one launch of `threads` per block over the cases (those with pow2
    // spacing or not, for the records) with grid(case)
    auto op = [&](BLaunch fn, int threads, auto grid, int p2, bool ident) {
        std::vector<SArg> args;
        for (int i = 0; i < (int)v.size(); ++i) {
            if (p2 >= 0 && (pow2_grid(cs[idx[i]]) ? 1 : 0) != p2) continue;
            args.push_back(SArg{bgrid(grid(cs[idx[i]], v[i]), dim3(threads)), i});
        }
        BOp o;
        o.aux_off = aux;
        return add_op(P.ar, P.ops, fn, threads, args, o, err, ident);
    };
    auto markers = [](ibmdem_ctx *c, const IbmSetup &) { return dim3(div_up(c->M, T)); };
    auto flags = [](ibmdem_ctx *, const IbmSetup &d) { return dim3(d.nb); };
    auto one = [](ibmdem_ctx *, const IbmSetup &) { return dim3(1); };
    auto tiles = [](ibmdem_ctx *c, const IbmSetup &) { return dim3(c->act_grid); };
    auto ent = [](ibmdem_ctx *, const IbmSetup &d) { return dim3(d.enb); };
    int rc;
    if ((rc = op(launch_records_p2, T, markers, 1, false))) return rc;
    if ((rc = op(launch_records, T, markers, 0, false))) return rc;
    if ((rc = op(launch_mark_tiles, T, markers, -1, false))) return rc;
    if ((rc = op(launch_flag_count, T, flags, -1, false))) return rc;
    if ((rc = op(launch_scan_small, 1024, one, -1, false))) return rc;
    if ((rc = op(launch_flag_compact, T, flags, -1, false))) return rc;
    if ((rc = op(launch_tile_count, 64, tiles, -1, true))) return rc;
    if ((rc = op(launch_ecount_sum, T, ent, -1, false))) return rc;
    if ((rc = op(launch_ecount_scan_blocks, 1024, one, -1, false))) return rc;
    if ((rc = op(launch_ecount_scan, T, ent, -1, false))) return rc;
    return op(launch_tile_fill, 64, tiles, -1, true);
}
