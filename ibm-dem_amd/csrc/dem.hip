// dem.hip — DEM cell lists, marker contact and rigid-body dynamics.
//
//   binning          CompressDEM + PutParcelsIntoCellsDEM
//                    (ZJS_UPDATE_LIST_DEM.F90:2-97): a marker that left
//                    through a periodic face is wrapped with x + coordShift,
//                    cell = floor((x - 0)/dx) + 1; markers are counting-
//                    sorted by cell with ascending id inside a cell.
//   contact          RebuildParcelCollisionListDEM + Collision +
//                    CalculateCollisionForce + CalculateSolidBoundary...
//                    (ZJS_UPDATE_LIST_DEM.F90:248-286, ZJS_COLLISION_DEM.F90:
//                    2-148). The candidate scan and the force use the same
//                    positions (nothing moves between the list rebuild and
//                    the next contact call), so one marker-parallel kernel
//                    scans the 125 cells in (kk,jj,ii) order, tests each
//                    candidate (other body, d^2 < (1.65h)^2, periodic images
//                    included, wall mirrors excluded) and adds its force in
//                    list order.
//   body assembly    AssembleLocalForceMS (ZJS_MS_DEM.F90:494-519): one
//                    work-group per body over its contiguous markers, fixed
//                    reduction tree.
//   Newton-Euler     EvolutionMS (couplesolverCylinder.F90:2-42), fused into
//                    the body reduction.
//   marker update    ResetVelocityAndPositionLagrangianPoints +
//                    AdjustParcelPosition (couplesolverCylinder.F90:44-74,
//                    ZJS_EVOLUTION_DEM.F90:121-150).
#include "common.hpp"
#include "batched.hpp"
#include "scan.hpp"

#include <cmath>

namespace {

using namespace bk;

using namespace scan;

constexpr int T = 256;
constexpr int CHUNK = 4096;
constexpr int BT = 256;   // threads per body work-group
static_assert(T == SCAN_T && CHUNK == SCAN_CHUNK, "the binning scans use the shared chunk kernels");

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
// The cell bins (bin_start, bin_items) are read only by a contact scan of a
// marker whose body has another body in reach, and by the IBM tables at the
// start of a step. A rebuild between two sub-steps of a step therefore
// takes a gate, near[B] of k_body_near for the next sub-step: where it is 0
// no marker is scanned, and only the cells and the periodic wraps are
// updated (the counts stay zero, the bins keep their last contents). A null
// gate always bins.
__device__ __forceinline__ bool bins_off(const int *gate) { return gate != nullptr && *gate == 0; }

__device__ __forceinline__ void wrap_cells_count_body(const Blk &kb, int M, GridDev g, double *__restrict__ mx,
                                                        int *__restrict__ mcell, int *__restrict__ bin_count,
                                                        int *__restrict__ flags, const int *gate) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M) return;
    const int n[3] = {g.nx, g.ny, g.nz};
    const double dd[3] = {g.dx, g.dy, g.dz};
    int c[3];
    // the three coordinates are loaded before a wrapped one is stored back
    const double x3[3] = {mx[m], mx[(long)M + m], mx[2L * M + m]};
    for (int d = 0; d < 3; ++d) {
        double x = x3[d];
        int ci = cell_of(x, dd[d]);
        if (ci > n[d] && g.bt[2 * d + 1] == IBMDEM_BC_PERIODIC) {
            x = x + (-g.L[d]);
            mx[(long)d * M + m] = x;
            ci = cell_of(x, dd[d]);
        } else if (ci < 1 && g.bt[2 * d] == IBMDEM_BC_PERIODIC) {
            x = x + g.L[d];
            mx[(long)d * M + m] = x;
            ci = cell_of(x, dd[d]);
        }
        // a marker through a wall, or a non-finite position, is an error;
        // it is still binned in the nearest boundary cell (a non-finite one
        // in cell 1) so that every list stays consistent until the caller
        // sees the status
        const bool finite = x - x == 0.0;
        if (!finite) {
            flags[FL_NAN] = 1;
            ci = 1;
        } else if (!(ci >= 1 && ci <= n[d])) {
            flags[FL_OUTSIDE] = 1;
            ci = ci < 1 ? 1 : n[d];
        }
        c[d] = ci;
    }
    const int cl = (int)cidx(g, c[0], c[1], c[2]);
    mcell[m] = cl;
    if (!bins_off(gate)) atomicAdd(&bin_count[cl], 1);
}
__global__ __launch_bounds__(T) void k_wrap_cells_count(int M, GridDev g, double *__restrict__ mx,
                                                        int *__restrict__ mcell, int *__restrict__ bin_count,
                                                        int *__restrict__ flags, const int *gate) {
    wrap_cells_count_body(hw_blk(), M, g, mx, mcell, bin_count, flags, gate);
}

// the counting scan's three kernels (scan.hpp); a closed gate ends every
// block before its first barrier
__global__ __launch_bounds__(T) void k_chunk_sum(long n, const int *__restrict__ a, int *__restrict__ out,
                                                 const int *gate) {
    if (bins_off(gate)) return;
    chunk_sum_body(hw_blk(), n, a, out);
}
__global__ __launch_bounds__(1024) void k_scan_blocks(int nb, int *__restrict__ a, const int *gate) {
    if (bins_off(gate)) return;
    scan_blocks_body(hw_blk(), nb, a);
}
__global__ __launch_bounds__(T) void k_chunk_scan(long n, const int *__restrict__ cnt, const int *__restrict__ off,
                                                  int *__restrict__ start, int nb, const int *gate) {
    if (bins_off(gate)) return;
    chunk_scan_body(hw_blk(), n, cnt, off, start, nb);
}

// unordered placement; the counts return to zero
__device__ __forceinline__ void bin_scatter_body(const Blk &kb, int M, const int *__restrict__ mcell, const int *__restrict__ start,
                                                   int *__restrict__ cnt, int *__restrict__ tmp, const int *gate) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M || bins_off(gate)) return;
    const int cl = mcell[m];
    const int slot = start[cl] + atomicSub(&cnt[cl], 1) - 1;
    tmp[slot] = m;
}
__global__ __launch_bounds__(T) void k_bin_scatter(int M, const int *__restrict__ mcell, const int *__restrict__ start,
                                                   int *__restrict__ cnt, int *__restrict__ tmp, const int *gate) {
    bin_scatter_body(hw_blk(), M, mcell, start, cnt, tmp, gate);
}

// ascending id inside each cell: position = start + #{smaller ids in cell}
__device__ __forceinline__ void bin_rank_body(const Blk &kb, int M, const int *__restrict__ mcell, const int *__restrict__ start,
                                                const int *__restrict__ tmp, int *__restrict__ items, const int *gate) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M || bins_off(gate)) return;
    const int cl = mcell[m];
    const int q0 = start[cl], q1 = start[cl + 1];
    int r = 0;
    for (int q = q0; q < q1; ++q) r += tmp[q] < m ? 1 : 0;
    items[q0 + r] = m;
}
__global__ __launch_bounds__(T) void k_bin_rank(int M, const int *__restrict__ mcell, const int *__restrict__ start,
                                                const int *__restrict__ tmp, int *__restrict__ items, const int *gate) {
    bin_rank_body(hw_blk(), M, mcell, start, tmp, items, gate);
}

// ---------------------------------------------------------------------------
// contact
// ---------------------------------------------------------------------------
struct ContactParams {
    double lim2, kn_over, eta;   // (cd(1+buf))^2, knCoeff/(0.1 di), eta
};

// neighbour source cell of the candidate scan: interior index and the
// periodic shift of its images; false for cells beyond a wall (mirror
// copies carry pType 1 and are never partners).
__device__ __forceinline__ bool nb_cell(int &i, int n, int lo, int hi, double L, double &sh, int &code) {
    sh = 0.0; code = 0;
    if (i < 1) { if (lo != IBMDEM_BC_PERIODIC) return false; i += n; sh = -L; code = -1; }
    else if (i > n) { if (hi != IBMDEM_BC_PERIODIC) return false; i -= n; sh = L; code = 1; }
    return true;
}

// max |x_m - X_b| over each body's markers (rigid: constant up to rounding)
__global__ __launch_bounds__(T) void k_body_rmax(int M, int B, GridDev g, const int *__restrict__ bstart,
                                                 const double *__restrict__ mx, const double *__restrict__ bx,
                                                 double *__restrict__ rmax) {
    __shared__ double red[T];
    const int b = blockIdx.x;
    double r = 0.0;
    for (int m = bstart[b] + threadIdx.x; m < bstart[b + 1]; m += T) {
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            double dd = mx[(long)d * M + m] - bx[(long)d * B + b];
            if (g.bt[2 * d] == IBMDEM_BC_PERIODIC) dd -= g.L[d] * rint(dd / g.L[d]);
            d2 += dd * dd;
        }
        r = fmax(r, sqrt(d2));
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int st = T / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) rmax[b] = red[0];
}

// The bodies c whose markers can come within the candidate distance of body
// b's: |X_b - X_c| < rmax_b + rmax_c + lim (+ a safety margin). near[b]
// counts them and list[b * NEAR_MAX + s] holds the first NEAR_MAX (in no
// particular order: the list only culls, see k_contact), and near[B] is 1
// if any body has one (the gate of the bins, above). Thread b of block
// (x, y) tests body b = 64x + t against the NEAR_CH bodies c of chunk y,
// staged in LDS (every block would otherwise read all B positions from L2);
// k_body_near_clear (between sub-steps, the body reduction before it) zeroed
// the counts and near[B]. The periodic minimum image takes
// rint(d * (1/L)): it can differ from rint(d / L) only where both images are
// L/2 away, within rounding, which the safety margin covers, so no body in
// reach is missed.
constexpr int NEAR_CH = 16;
constexpr int NEAR_MAX = 32;
__device__ __forceinline__ void body_near_clear_body(const Blk &kb, int B, int *__restrict__ near) {
    const int b = kb.bx * 64 + kb.tx;
    if (b <= B) near[b] = 0;
}
__global__ __launch_bounds__(64) void k_body_near_clear(int B, int *__restrict__ near) {
    body_near_clear_body(hw_blk(), B, near);
}
__device__ __forceinline__ void body_near_body(const Blk &kb, int B, GridDev g, const double *__restrict__ bx,
                                                  const double *__restrict__ rmax, double lim, int *__restrict__ near,
                                                  int *__restrict__ list) {
    __shared__ double sx[NEAR_CH], sy[NEAR_CH], sz[NEAR_CH], sr[NEAR_CH];
    const int c0 = kb.by * NEAR_CH, nc = min(NEAR_CH, B - c0);
    for (int q = kb.tx; q < nc; q += 64) {
        sx[q] = bx[c0 + q];
        sy[q] = bx[(long)B + c0 + q];
        sz[q] = bx[2L * B + c0 + q];
        sr[q] = rmax[c0 + q];
    }
    __syncthreads();
    const int b = kb.bx * 64 + kb.tx;
    if (b >= B) return;
    const double xb0 = bx[b], xb1 = bx[(long)B + b], xb2 = bx[2L * B + b];
    const double rb = rmax[b] + lim + 1e-6 * (g.dx + lim);
    const double iL[3] = {1.0 / g.L[0], 1.0 / g.L[1], 1.0 / g.L[2]};
    for (int q = 0; q < nc; ++q) {
        if (c0 + q == b) continue;
        double d[3] = {sx[q] - xb0, sy[q] - xb1, sz[q] - xb2};
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            if (g.bt[2 * k] == IBMDEM_BC_PERIODIC) d[k] -= g.L[k] * rint(d[k] * iL[k]);
            d2 += d[k] * d[k];
        }
        const double lim_bc = rb + sr[q];
        if (d2 < lim_bc * lim_bc) {
            const int slot = atomicAdd(&near[b], 1);
            if (slot < NEAR_MAX) list[(long)b * NEAR_MAX + slot] = c0 + q;
            near[B] = 1;
        }
    }
}
__global__ __launch_bounds__(64) void k_body_near(int B, GridDev g, const double *__restrict__ bx,
                                                  const double *__restrict__ rmax, double lim, int *__restrict__ near,
                                                  int *__restrict__ list) {
    body_near_body(hw_blk(), B, g, bx, rmax, lim, near, list);
}

// fcol of marker i (ZJS_COLLISION_DEM.F90:31-72, 93-146)
template <bool PAIRS>
__device__ __forceinline__ void contact_body(const Blk &kb, int M, GridDev g, const double *__restrict__ mx,
                                               const double *__restrict__ mv, const double *__restrict__ mmass,
                                               const double *__restrict__ mrad, const int *__restrict__ mbody,
                                               const int *__restrict__ mcell, const int *__restrict__ bin_start,
                                               const int *__restrict__ bin_items, double *__restrict__ fcol,
                                               ContactParams cp, int max_pairs, int *__restrict__ flags,
                                               int cap, int *__restrict__ pcount, int *__restrict__ ppartner,
                                               int *__restrict__ pshift, const int *__restrict__ bnear,
                                               const int *__restrict__ blist, const double *__restrict__ bx, int B,
                                               const double *__restrict__ brmax, double lim) {
    const int i = kb.bx * T + kb.tx;
    if (i >= M) return;
    const int cl = mcell[i];
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    const int bi = mbody[i];
    const int nnear = bnear[bi];
    // wall contact reaches the markers of the first/last interior layer
    bool walls = false;
    for (int s = 0; s < 6; ++s) walls = walls || g.bt[s] != IBMDEM_BC_PERIODIC;
    const bool layer = walls && (ci == 1 || ci == g.nx || cj == 1 || cj == g.ny || ck == 1 || ck == g.nz);
    if (nnear == 0 && !layer) {
        // no candidate pair and no wall: the sums below stay zero, so they
        // are written without loading the marker's state
        if (0 > max_pairs) flags[FL_CAPACITY] = 1;
        if (PAIRS) {
            pcount[i] = 0;
            return;
        }
        fcol[i] = 0.0;
        fcol[(long)M + i] = 0.0;
        fcol[2L * M + i] = 0.0;
        return;
    }
    const double xi0 = mx[i], xi1 = mx[(long)M + i], xi2 = mx[2L * M + i];
    const double vi0 = mv[i], vi1 = mv[(long)M + i], vi2 = mv[2L * M + i];
    const double mi = mmass[i], ri = mrad[i];
    double f0 = 0, f1 = 0, f2 = 0;
    int cnt = 0;
    // a marker with no other body in reach has no candidate pairs: skipping
    // its scan adds nothing, exactly as the scan would. A partner j of body c
    // within lim of x_i puts x_i within lim + rmax_c of c's centre (nearest
    // image), so the marker is scanned only if some body on its body's near
    // list is that close (with the body test's margin), or if the list
    // overflowed.
    bool scan = nnear > NEAR_MAX;
    if (nnear > 0 && !scan) {
        const double iL[3] = {1.0 / g.L[0], 1.0 / g.L[1], 1.0 / g.L[2]};
        const double xi[3] = {xi0, xi1, xi2};
        const double rl = lim + 1e-6 * (g.dx + lim);
        for (int s = 0; s < nnear && !scan; ++s) {
            const int cb = blist[(long)bi * NEAR_MAX + s];
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                double d = xi[k] - bx[(long)k * B + cb];
                if (g.bt[2 * k] == IBMDEM_BC_PERIODIC) d -= g.L[k] * rint(d * iL[k]);
                d2 += d * d;
            }
            const double r = rl + brmax[cb];
            scan = d2 < r * r;
        }
    }
    const int kstart = scan ? -2 : 3;
    for (int kk = kstart; kk <= 2; ++kk) {
        int kc = ck + kk, kcode; double ksh;
        if (!nb_cell(kc, g.nz, g.bt[4], g.bt[5], g.L[2], ksh, kcode)) continue;
        for (int jj = -2; jj <= 2; ++jj) {
            int jc = cj + jj, jcode; double jsh;
            if (!nb_cell(jc, g.ny, g.bt[2], g.bt[3], g.L[1], jsh, jcode)) continue;
            // the five x cells of this row; when none wraps they are
            // consecutive cells, so their sorted items form one range
            const bool row = ci - 2 >= 1 && ci + 2 <= g.nx;
            for (int ii = row ? 0 : -2; ii <= 2; ++ii) {
                int ic = ci + ii, icode; double ish;
                int q0, q1;
                if (row) {
                    icode = 0; ish = 0.0;
                    q0 = bin_start[cidx(g, ci - 2, jc, kc)];
                    q1 = bin_start[cidx(g, ci + 2, jc, kc) + 1];
                    ii = 2;
                } else {
                    if (!nb_cell(ic, g.nx, g.bt[0], g.bt[1], g.L[0], ish, icode)) continue;
                    const long b = cidx(g, ic, jc, kc);
                    q0 = bin_start[b];
                    q1 = bin_start[b + 1];
                }
                for (int q = q0; q < q1; ++q) {
                    const int j = bin_items[q];
                    if (mbody[j] == bi) continue;
                    double xj0 = mx[j], xj1 = mx[(long)M + j], xj2 = mx[2L * M + j];
                    if (icode) xj0 = xj0 + ish;
                    if (jcode) xj1 = xj1 + jsh;
                    if (kcode) xj2 = xj2 + ksh;
                    const double r1 = xi0 - xj0, r2 = xi1 - xj1, r3 = xi2 - xj2;
                    const double d2 = r1 * r1 + r2 * r2 + r3 * r3;
                    if (!(d2 < cp.lim2)) continue;
                    if (PAIRS) {
                        if (cnt < cap) {
                            ppartner[(long)i * cap + cnt] = j;
                            pshift[(long)i * cap + cnt] = (icode + 1) + 3 * (jcode + 1) + 9 * (kcode + 1);
                        }
                        ++cnt;
                        continue;
                    }
                    ++cnt;
                    const double dist = sqrt(d2);
                    const double n1 = r1 / dist, n2 = r2 / dist, n3 = r3 / dist;
                    const double mj = mmass[j];
                    const double meff = 2 * mi * mj / (mi + mj);
                    const double rij = ri + mrad[j];
                    const double deltan = rij - dist;
                    const double kn = cp.kn_over * meff;
                    const double etan = cp.eta * sqrt(meff * kn);
                    const double vnij = n1 * (vi0 - mv[j]) + n2 * (vi1 - mv[(long)M + j]) + n3 * (vi2 - mv[2L * M + j]);
                    double g1 = 0, g2 = 0, g3 = 0;
                    if (deltan > 0) {
                        g1 = n1 * (kn * deltan - etan * vnij);
                        g2 = n2 * (kn * deltan - etan * vnij);
                        g3 = n3 * (kn * deltan - etan * vnij);
                    }
                    f0 = f0 + g1;
                    f1 = f1 + g2;
                    f2 = f2 + g3;
                }
            }
        }
    }
    if (cnt > max_pairs) flags[FL_CAPACITY] = 1;
    if (PAIRS) {
        pcount[i] = cnt;
        return;
    }
    if (layer) {
        for (int s = 0; s < 6; ++s) {
            if (g.bt[s] == IBMDEM_BC_PERIODIC) continue;   // walls, inflow and outflow screens
            const double p1 = (s % 2 == 0) ? 0.0 : g.L[0];
            const double p2 = (s % 2 == 0) ? 0.0 : g.L[1];
            const double p3 = (s % 2 == 0) ? 0.0 : g.L[2];
            const double sg = (s % 2 == 0) ? 1.0 : -1.0;
            const double n1 = (s / 2 == 0) ? sg : 0.0, n2 = (s / 2 == 1) ? sg : 0.0, n3 = (s / 2 == 2) ? sg : 0.0;
            const double r1 = xi0 - p1, r2 = xi1 - p2, r3 = xi2 - p3;
            const double dist = r1 * n1 + r2 * n2 + r3 * n3;
            const double meff = mi;
            const double deltan = ri - dist;
            const double kn = cp.kn_over * meff;
            const double etan = cp.eta * sqrt(meff * kn);
            const double vnij = n1 * vi0 + n2 * vi1 + n3 * vi2;
            if (deltan > 0) {
                f0 = f0 + n1 * (kn * deltan - etan * vnij);
                f1 = f1 + n2 * (kn * deltan - etan * vnij);
                f2 = f2 + n3 * (kn * deltan - etan * vnij);
            }
        }
    }
    fcol[i] = f0;
    fcol[(long)M + i] = f1;
    fcol[2L * M + i] = f2;
}
template <bool PAIRS>
__global__ __launch_bounds__(T) void k_contact(int M, GridDev g, const double *__restrict__ mx,
                                               const double *__restrict__ mv, const double *__restrict__ mmass,
                                               const double *__restrict__ mrad, const int *__restrict__ mbody,
                                               const int *__restrict__ mcell, const int *__restrict__ bin_start,
                                               const int *__restrict__ bin_items, double *__restrict__ fcol,
                                               ContactParams cp, int max_pairs, int *__restrict__ flags,
                                               int cap, int *__restrict__ pcount, int *__restrict__ ppartner,
                                               int *__restrict__ pshift, const int *__restrict__ bnear,
                                               const int *__restrict__ blist, const double *__restrict__ bx, int B,
                                               const double *__restrict__ brmax, double lim) {
    contact_body<PAIRS>(hw_blk(), M, g, mx, mv, mmass, mrad, mbody, mcell, bin_start, bin_items, fcol, cp, max_pairs, flags, cap, pcount, ppartner, pshift, bnear, blist, bx, B, brmax, lim);
}

// ---------------------------------------------------------------------------
// rigid bodies
// ---------------------------------------------------------------------------
struct BodyState {
    double *x, *v, *vold, *om, *L, *Lold, *rot, *rotg, *I, *m, *rho;
    int *fixed;
    double *fdrag, *fmom, *fcol, *fmomcol, *fhydro, *fmomhydro;
    int *flags;     // FL_NAN is raised for a non-finite updated state
};

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double d[3]) {
    const double c1 = a[1] * b[2] - a[2] * b[1];
    const double c2 = -a[0] * b[2] + a[2] * b[0];
    const double c3 = a[0] * b[1] - a[1] * b[0];
    d[0] = c1; d[1] = c2; d[2] = c3;
}
__device__ __forceinline__ void tmul(const double *A, const double *B, double *D) {
    double C[9];
    C[0] = A[0] * B[0] + A[1] * B[3] + A[2] * B[6];
    C[1] = A[0] * B[1] + A[1] * B[4] + A[2] * B[7];
    C[2] = A[0] * B[2] + A[1] * B[5] + A[2] * B[8];
    C[3] = A[3] * B[0] + A[4] * B[3] + A[5] * B[6];
    C[4] = A[3] * B[1] + A[4] * B[4] + A[5] * B[7];
    C[5] = A[3] * B[2] + A[4] * B[5] + A[5] * B[8];
    C[6] = A[6] * B[0] + A[7] * B[3] + A[8] * B[6];
    C[7] = A[6] * B[1] + A[7] * B[4] + A[8] * B[7];
    C[8] = A[6] * B[2] + A[7] * B[5] + A[8] * B[8];
    for (int q = 0; q < 9; ++q) D[q] = C[q];
}

// nearest periodic image of body b's centre seen from position xm
__device__ __forceinline__ void body_image(const GridDev &g, const double *bx, int B, int b, const double xm[3],
                                           double xb[3]) {
    for (int d = 0; d < 3; ++d) {
        xb[d] = bx[(long)d * B + b];
        if (g.bt[2 * d] == IBMDEM_BC_PERIODIC) {
            const double dd = xm[d] - xb[d];
            if (dd > 0.5 * g.L[d]) xb[d] = xb[d] + g.L[d];
            else if (dd < -0.5 * g.L[d]) xb[d] = xb[d] + (-g.L[d]);
        }
    }
}

// EvolutionMS for one body (couplesolverCylinder.F90:14-40)
// One thread per body: every body value is loaded before the first store
// (the state arrays are written through the same pointers, so loads after a
// store would each wait for the one before), then the update, then the
// stores. The arithmetic is EvolutionMS's in its order.
__device__ void evolve_body(const BodyState &s, int B, int b, double dtD, double rhof, const double grav[3]) {
    if (s.fixed[b]) return;
    double v[3], fh[3], fc[3], om[3], L[3], fmh[3], fmc[3], x[3], Iv[9], rgl[9];
    const double rho = s.rho[b], mb = s.m[b];
    for (int d = 0; d < 3; ++d) {
        const long o = (long)d * B + b;
        v[d] = s.v[o]; fh[d] = s.fhydro[o]; fc[d] = s.fcol[o]; om[d] = s.om[o];
        L[d] = s.L[o]; fmh[d] = s.fmomhydro[o]; fmc[d] = s.fmomcol[o]; x[d] = s.x[o];
    }
    for (int q = 0; q < 9; ++q) {
        Iv[q] = s.I[9L * b + q];
        rgl[q] = s.rotg[9L * b + q];
    }
    const double ratio = (rho - rhof) / rho;
    for (int d = 0; d < 3; ++d) v[d] = v[d] + fh[d] * dtD + fc[d] / mb * dtD + ratio * grav[d] * dtD;
    // CalculateRotateMatrix(omega*dtDEM) (couplesolverCylinder.F90:76-97)
    const double th0 = om[0] * dtD, th1 = om[1] * dtD, th2 = om[2] * dtD;
    double mag = sqrt(th0 * th0 + th1 * th1 + th2 * th2);
    mag = mag > 1e-10 ? mag : 1e-10;
    const double ux = th0 / mag, uy = th1 / mag, uz = th2 / mag;
    const double cm = cos(mag), sm = sin(mag);
    double rot[9];
    rot[0] = (1 - cm) * ux * ux + cm;
    rot[1] = (1 - cm) * ux * uy - sm * uz;
    rot[2] = (1 - cm) * ux * uz + sm * uy;
    rot[3] = (1 - cm) * uy * ux + sm * uz;
    rot[4] = (1 - cm) * uy * uy + cm;
    rot[5] = (1 - cm) * uy * uz - sm * ux;
    rot[6] = (1 - cm) * uz * ux - sm * uy;
    rot[7] = (1 - cm) * uz * uy + sm * ux;
    rot[8] = (1 - cm) * uz * uz + cm;
    double ITInv[9];
    for (int q = 0; q < 9; ++q) ITInv[q] = Iv[q];
    ITInv[0] = 1.0 / ITInv[0]; ITInv[4] = 1.0 / ITInv[4]; ITInv[8] = 1.0 / ITInv[8];
    tmul(rgl, rot, rgl);
    const double RMInv[9] = {rgl[0], rgl[3], rgl[6], rgl[1], rgl[4], rgl[7], rgl[2], rgl[5], rgl[8]};
    double T1[9], TM[9];
    tmul(rgl, ITInv, T1);
    tmul(T1, RMInv, TM);
    for (int d = 0; d < 3; ++d) L[d] = L[d] + (fmh[d] + fmc[d]) * dtD;
    const double o0 = TM[0] * L[0] + TM[1] * L[1] + TM[2] * L[2];
    const double o1 = TM[3] * L[0] + TM[4] * L[1] + TM[5] * L[2];
    const double o2 = TM[6] * L[0] + TM[7] * L[1] + TM[8] * L[2];
    for (int d = 0; d < 3; ++d) x[d] = x[d] + v[d] * dtD;
    for (int q = 0; q < 9; ++q) {
        s.rot[9L * b + q] = rot[q];
        s.rotg[9L * b + q] = rgl[q];
    }
    const double omn[3] = {o0, o1, o2};
    double chk = 0.0;   // finite iff every updated value is
    for (int d = 0; d < 3; ++d) {
        const long o = (long)d * B + b;
        s.v[o] = v[d]; s.L[o] = L[d]; s.om[o] = omn[d]; s.x[o] = x[d];
        chk += (v[d] - v[d]) + (L[d] - L[d]) + (omn[d] - omn[d]) + (x[d] - x[d]);
    }
    if (!(chk == 0.0)) s.flags[FL_NAN] = 1;
}

// mode 0: fcol, fmomcol only; 1: all four sums (AssembleLocalForceMS);
// 2: all four + CloudEvolutionMS hydro terms (couplesolverCylinder.F90:194-211)
// integrate: EvolutionMS after the sums.
// ResetVelocityAndPositionLagrangianPoints with RotateLagrangianPoints
// (couplesolverCylinder.F90:44-74) and AdjustParcelPosition's clamp
// (ZJS_EVOLUTION_DEM.F90:121-150) for marker m, given its body's state after
// the sub-step
__device__ __forceinline__ void follow_marker(int M, const GridDev &g, double *__restrict__ mx,
                                                double *__restrict__ mv, int m, bool fixed, const double bxv[3],
                                                const double bv[3], const double om[3], const double mat[9],
                                                double dtD) {
    double xm[3] = {mx[m], mx[(long)M + m], mx[2L * M + m]};
    if (!fixed) {
        // body_image's rule on the loaded centre
        double xb[3];
        for (int d = 0; d < 3; ++d) {
            xb[d] = bxv[d];
            if (g.bt[2 * d] == IBMDEM_BC_PERIODIC) {
                const double dd = xm[d] - xb[d];
                if (dd > 0.5 * g.L[d]) xb[d] = xb[d] + g.L[d];
                else if (dd < -0.5 * g.L[d]) xb[d] = xb[d] + (-g.L[d]);
            }
        }
        for (int d = 0; d < 3; ++d) xm[d] = xm[d] + bv[d] * dtD;
        const double t0 = xm[0] - xb[0], t1 = xm[1] - xb[1], t2 = xm[2] - xb[2];
        xm[0] = mat[0] * t0 + mat[1] * t1 + mat[2] * t2;
        xm[1] = mat[3] * t0 + mat[4] * t1 + mat[5] * t2;
        xm[2] = mat[6] * t0 + mat[7] * t1 + mat[8] * t2;
        for (int d = 0; d < 3; ++d) xm[d] = xm[d] + xb[d];
        const double rr[3] = {xm[0] - xb[0], xm[1] - xb[1], xm[2] - xb[2]};
        double vv[3];
        cross3(om, rr, vv);
        for (int d = 0; d < 3; ++d) mv[(long)d * M + m] = vv[d] + bv[d];
    }
    const double e[3] = {g.dx * 1e-20, g.dy * 1e-20, g.dz * 1e-20};
    for (int d = 0; d < 3; ++d) {
        double x = xm[d];
        if (fabs(x - 0.0) < e[d]) x = 0.0 + e[d];
        if (fabs(x - g.L[d]) < e[d]) x = g.L[d] + e[d];
        mx[(long)d * M + m] = x;
    }
}

__device__ __forceinline__ void body_finish(const Blk &kb, int B, int b, BodyState s, int mode, int integrate,
                                              double rhof, double dt, double dtD, double g0, double g1, double g2,
                                              double (*red)[BT], int follow, double *sbody, int &sfixed);
// mx is written too when follow is set (so it is neither const nor restrict)
__device__ __forceinline__ void body_reduce_body(const Blk &kb, int M, int B, GridDev g, const int *__restrict__ bstart,
                                                    double *mx, const double *__restrict__ fdrag,
                                                    const double *__restrict__ fcol, const double *__restrict__ mvol,
                                                    BodyState s, int mode, int integrate, double rhof, double dt,
                                                    double dtD, double g0, double g1, double g2, int follow,
                                                    double *__restrict__ mv, int *__restrict__ near_clear) {
    __shared__ double red[12][BT];
    const int b = kb.bx;
    const int m0 = bstart[b], m1 = bstart[b + 1];
    double acc[12];
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    for (int m = m0 + kb.tx; m < m1; m += BT) {
        const double xm[3] = {mx[m], mx[(long)M + m], mx[2L * M + m]};
        double xb[3], rpc[3], fv[3], fc[3], mt[3], mc[3];
        body_image(g, s.x, B, b, xm, xb);
        for (int d = 0; d < 3; ++d) {
            rpc[d] = xm[d] - xb[d];
            fc[d] = fcol[(long)d * M + m];
        }
        cross3(rpc, fc, mc);
        for (int d = 0; d < 3; ++d) {
            acc[3 + d] += fc[d];
            acc[9 + d] += mc[d];
        }
        if (mode >= 1) {
            for (int d = 0; d < 3; ++d) fv[d] = fdrag[(long)d * M + m] * mvol[m];
            cross3(rpc, fv, mt);
            for (int d = 0; d < 3; ++d) {
                acc[d] += fdrag[(long)d * M + m] * mvol[m] * rhof;
                acc[6 + d] += mt[d];
            }
        }
    }
    for (int q = 0; q < 12; ++q) red[q][kb.tx] = acc[q];
    __syncthreads();
    for (int st = BT / 2; st > 0; st >>= 1) {
        if (kb.tx < st)
            for (int q = 0; q < 12; ++q) red[q][kb.tx] += red[q][kb.tx + st];
        __syncthreads();
    }
    // follow: the work-group then moves its body's markers (follow_body's
    // update, from the body's new state as thread 0 leaves it in LDS)
    __shared__ double sbody[18];
    __shared__ int sfixed;
    if (kb.tx == 0) {
        body_finish(kb, B, b, s, mode, integrate, rhof, dt, dtD, g0, g1, g2, red, follow, sbody, sfixed);
        // the next culling's counts (k_body_near_clear's work), when asked
        if (near_clear) {
            near_clear[b] = 0;
            if (b == 0) near_clear[B] = 0;
        }
    }
    if (!follow) return;
    __syncthreads();
    const bool fixed = sfixed != 0;
    double bxv[3], bv[3], om[3], mat[9];
    for (int d = 0; d < 3; ++d) {
        bxv[d] = sbody[d];
        bv[d] = sbody[3 + d];
        om[d] = sbody[6 + d];
    }
    for (int q = 0; q < 9; ++q) mat[q] = sbody[9 + q];
    for (int m = m0 + kb.tx; m < m1; m += BT)
        follow_marker(M, g, mx, mv, m, fixed, bxv, bv, om, mat, dtD);
}
// thread 0 of body_reduce_body: the body's sums stored, the hydro terms, the
// Newton-Euler update, and for follow the new state into LDS
__device__ __forceinline__ void body_finish(const Blk &kb, int B, int b, BodyState s, int mode, int integrate,
                                              double rhof, double dt, double dtD, double g0, double g1, double g2,
                                              double (*red)[BT], int follow, double *sbody, int &sfixed) {
    for (int d = 0; d < 3; ++d) {
        s.fcol[(long)d * B + b] = red[3 + d][0];
        s.fmomcol[(long)d * B + b] = red[9 + d][0];
        if (mode >= 1) {
            s.fdrag[(long)d * B + b] = red[d][0];
            s.fmom[(long)d * B + b] = red[6 + d][0];
        }
    }
    if (mode == 2) {
        // the sums just stored are still in red; the rest is loaded before
        // any store
        const double rho = s.rho[b], mb = s.m[b];
        double v[3], vo[3], L[3], Lo[3];
        for (int d = 0; d < 3; ++d) {
            const long o = (long)d * B + b;
            v[d] = s.v[o]; vo[d] = s.vold[o]; L[d] = s.L[o]; Lo[d] = s.Lold[o];
        }
        const double ratio = (rho - rhof) / rho;
        for (int d = 0; d < 3; ++d) {
            const long o = (long)d * B + b;
            const double fh = (v[d] - vo[d]) / dt * (1.0 - ratio);
            const double fmh = (L[d] - Lo[d]) / dt * (1.0 - ratio);
            s.vold[o] = v[d];
            s.Lold[o] = L[d];
            s.fhydro[o] = fh + red[d][0] / mb;
            s.fmomhydro[o] = fmh + red[6 + d][0];
        }
    }
    if (integrate) {
        const double grav[3] = {g0, g1, g2};
        evolve_body(s, B, b, dtD, rhof, grav);
    }
    if (follow) {
        // the values follow_body loads, as evolve_body left them
        for (int d = 0; d < 3; ++d) {
            sbody[d] = s.x[(long)d * B + b];
            sbody[3 + d] = s.v[(long)d * B + b];
            sbody[6 + d] = s.om[(long)d * B + b];
        }
        for (int q = 0; q < 9; ++q) sbody[9 + q] = s.rot[9L * b + q];
        sfixed = s.fixed[b];
    }
}
__global__ __launch_bounds__(BT) void k_body_reduce(int M, int B, GridDev g, const int *__restrict__ bstart,
                                                    double *mx, const double *__restrict__ fdrag,
                                                    const double *__restrict__ fcol, const double *__restrict__ mvol,
                                                    BodyState s, int mode, int integrate, double rhof, double dt,
                                                    double dtD, double g0, double g1, double g2, int follow,
                                                    double *__restrict__ mv, int *__restrict__ near_clear) {
    body_reduce_body(hw_blk(), M, B, g, bstart, mx, fdrag, fcol, mvol, s, mode, integrate, rhof, dt, dtD, g0, g1, g2,
                     follow, mv, near_clear);
}

__global__ void k_body_integrate(int B, BodyState s, double dtD, double rhof, double g0, double g1, double g2) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double grav[3] = {g0, g1, g2};
    evolve_body(s, B, b, dtD, rhof, grav);
}

// ResetVelocityAndPositionLagrangianPoints + AdjustParcelPosition
__device__ __forceinline__ void follow_body(const Blk &kb, int M, int B, GridDev g, double *__restrict__ mx,
                                              double *__restrict__ mv, const int *__restrict__ mbody, BodyState s,
                                              double dtD) {
    const int m = kb.bx * T + kb.tx;
    if (m >= M) return;
    const int b = mbody[m];
    // the body's values are loaded together, before the fixed test
    const bool fixed = s.fixed[b] != 0;
    double bxv[3], bv[3], om[3], mat[9];
    for (int d = 0; d < 3; ++d) {
        bxv[d] = s.x[(long)d * B + b];
        bv[d] = s.v[(long)d * B + b];
        om[d] = s.om[(long)d * B + b];
    }
    for (int q = 0; q < 9; ++q) mat[q] = s.rot[9L * b + q];
    follow_marker(M, g, mx, mv, m, fixed, bxv, bv, om, mat, dtD);
}
__global__ __launch_bounds__(T) void k_follow(int M, int B, GridDev g, double *__restrict__ mx,
                                              double *__restrict__ mv, const int *__restrict__ mbody, BodyState s,
                                              double dtD) {
    follow_body(hw_blk(), M, B, g, mx, mv, mbody, s, dtD);
}

BodyState body_state(ibmdem_ctx *c) {
    return BodyState{c->bx, c->bv, c->bvold, c->bom, c->bL, c->bLold, c->brot, c->brotg, c->bI, c->bm, c->brho,
                     c->bfixed, c->bfdrag, c->bfmom, c->bfcol, c->bfmomcol, c->bfhydro, c->bfmomhydro, c->flags};
}

ContactParams contact_params(ibmdem_ctx *c) {
    const double lim = c->cs.collision_dist * (1.0 + c->cs.collision_buffer);
    return ContactParams{lim * lim, c->cs.kn_coeff / (0.1 * c->g.dx), c->eta};
}

}  // namespace

int scan_counts(ibmdem_ctx *c, const int *cnt, int *start, long n) {
    hipStream_t s = c->stream;
    const int nb = div_up(n, CHUNK);
    if (nb + 1 > c->scan_nblocks) return set_error(c, IBMDEM_ERR_CAPACITY, "scan longer than its block buffer");
    const int *none = nullptr;
    hipLaunchKernelGGL(k_chunk_sum, dim3(nb), dim3(T), 0, s, n, cnt, c->scan_blocks, none);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, nb, c->scan_blocks, none);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_chunk_scan, dim3(nb), dim3(T), 0, s, n, cnt, c->scan_blocks, start, nb, none);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_rebuild(ibmdem_ctx *c, const int *gate) {
    const int M = c->M;
    if (M == 0) return IBMDEM_OK;
    hipStream_t s = c->stream;
    const long n = c->n;
    hipLaunchKernelGGL(k_wrap_cells_count, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mx, c->mcell,
                       c->bin_count, c->flags, gate);
    CHECK_LAUNCH(c);
    const int nb = div_up(n, CHUNK);
    hipLaunchKernelGGL(k_chunk_sum, dim3(nb), dim3(T), 0, s, n, c->bin_count, c->scan_blocks, gate);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, nb, c->scan_blocks, gate);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_chunk_scan, dim3(nb), dim3(T), 0, s, n, c->bin_count, c->scan_blocks, c->bin_start, nb,
                       gate);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_bin_scatter, dim3(div_up(M, T)), dim3(T), 0, s, M, c->mcell, c->bin_start, c->bin_count,
                       c->bin_tmp, gate);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_bin_rank, dim3(div_up(M, T)), dim3(T), 0, s, M, c->mcell, c->bin_start, c->bin_tmp,
                       c->bin_items, gate);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}
int dem_rebuild(ibmdem_ctx *c) { return dem_rebuild(c, nullptr); }

// the candidate distance of the contact scan, as the body culling uses it
double near_lim(const ibmdem_ctx *c) { return c->cs.collision_dist * (1.0 + c->cs.collision_buffer); }

// cleared: the counts and near[B] are already zero (the body reduction that
// ran just before cleared them)
int dem_near(ibmdem_ctx *c, bool cleared) {
    if (c->B == 0) return IBMDEM_OK;
    if (c->rmax_dirty) {
        hipLaunchKernelGGL(k_body_rmax, dim3(c->B), dim3(T), 0, c->stream, c->M, c->B, c->g, c->body_start, c->mx,
                           c->bx, c->brmax);
        CHECK_LAUNCH(c);
        c->rmax_dirty = false;
    }
    const double lim = near_lim(c);
    if (!cleared) {
        hipLaunchKernelGGL(k_body_near_clear, dim3(div_up(c->B + 1, 64)), dim3(64), 0, c->stream, c->B, c->bnear);
        CHECK_LAUNCH(c);
    }
    hipLaunchKernelGGL(k_body_near, dim3(div_up(c->B, 64), div_up(c->B, NEAR_CH)), dim3(64), 0, c->stream, c->B,
                       c->g, c->bx, c->brmax, lim, c->bnear, c->blist);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_near(ibmdem_ctx *c) { return dem_near(c, false); }

// the contact scan with the body culling of the last dem_near
int dem_contact_scan(ibmdem_ctx *c) {
    const int M = c->M;
    if (M == 0) return IBMDEM_OK;
    prof_begin(c, IBMDEM_K_CONTACT);
    hipLaunchKernelGGL((k_contact<false>), dim3(div_up(M, T)), dim3(T), 0, c->stream, M, c->g, c->mx, c->mv,
                       c->mmass, c->mrad, c->mbody, c->mcell, c->bin_start, c->bin_items, c->fcol,
                       contact_params(c), c->cs.max_pairs, c->flags, 0, nullptr, nullptr, nullptr, c->bnear, c->blist,
                       c->bx, c->B, c->brmax, near_lim(c));
    prof_end(c, IBMDEM_K_CONTACT, 0.0);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_contact(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    int rc;
    if ((rc = dem_near(c))) return rc;
    return dem_contact_scan(c);
}

int dem_pairs(ibmdem_ctx *c, int cap) {
    const int M = c->M;
    if (M == 0) return IBMDEM_OK;
    int rc;
    if (cap > c->pair_cap) {
        if ((rc = dev_alloc(c, &c->pair_count, M))) return rc;
        if ((rc = dev_alloc(c, &c->pair_partner, (size_t)M * cap))) return rc;
        if ((rc = dev_alloc(c, &c->pair_shift, (size_t)M * cap))) return rc;
        c->pair_cap = cap;
    }
    if ((rc = dem_near(c))) return rc;
    hipLaunchKernelGGL((k_contact<true>), dim3(div_up(M, T)), dim3(T), 0, c->stream, M, c->g, c->mx, c->mv,
                       c->mmass, c->mrad, c->mbody, c->mcell, c->bin_start, c->bin_items, c->fcol,
                       contact_params(c), c->cs.max_pairs, c->flags, c->pair_cap, c->pair_count, c->pair_partner,
                       c->pair_shift, c->bnear, c->blist, c->bx, c->B, c->brmax, near_lim(c));
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_body_assemble(ibmdem_ctx *c, bool all_terms) {
    if (c->B == 0) return IBMDEM_OK;
    hipLaunchKernelGGL(k_body_reduce, dim3(c->B), dim3(BT), 0, c->stream, c->M, c->B, c->g, c->body_start, c->mx,
                       c->fdrag, c->fcol, c->mvol, body_state(c), all_terms ? 1 : 0, 0, c->cs.rhof, c->cs.dt,
                       c->cs.dt / c->cs.couple_interval, c->cs.gravity[0], c->cs.gravity[1], c->cs.gravity[2], 0,
                       (double *)nullptr, (int *)nullptr);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_body_hydro(ibmdem_ctx *c) {
    if (c->B == 0) return IBMDEM_OK;
    hipLaunchKernelGGL(k_body_reduce, dim3(c->B), dim3(BT), 0, c->stream, c->M, c->B, c->g, c->body_start, c->mx,
                       c->fdrag, c->fcol, c->mvol, body_state(c), 2, 0, c->cs.rhof, c->cs.dt,
                       c->cs.dt / c->cs.couple_interval, c->cs.gravity[0], c->cs.gravity[1], c->cs.gravity[2], 0,
                       (double *)nullptr, (int *)nullptr);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_body_integrate(ibmdem_ctx *c) {
    if (c->B == 0) return IBMDEM_OK;
    hipLaunchKernelGGL(k_body_integrate, dim3(div_up(c->B, 64)), dim3(64), 0, c->stream, c->B, body_state(c),
                       c->cs.dt / c->cs.couple_interval, c->cs.rhof, c->cs.gravity[0], c->cs.gravity[1],
                       c->cs.gravity[2]);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dem_markers_follow(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    hipLaunchKernelGGL(k_follow, dim3(div_up(c->M, T)), dim3(T), 0, c->stream, c->M, c->B, c->g, c->mx, c->mv,
                       c->mbody, body_state(c), c->cs.dt / c->cs.couple_interval);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

// one CloudEvolutionStepMS with the body reduction and Newton-Euler fused.
// The body culling of the first sub-step of a sequence is computed here;
// each later sub-step's right after the bodies of the one before have moved
// (its contact scan sees those positions), so that it gates that sub-step's
// rebuild. The last sub-step's rebuild always bins: the IBM tables of the
// next step and the pair lists read the bins. The body reduction's
// work-groups also move their bodies' markers (the marker update needs only
// its own body's new state), in place of a launch of k_follow.
int dem_substep_fused(ibmdem_ctx *c, bool first, bool last) {
    int rc;
    const bool cull = c->M > 0 && c->B > 0;
    if (first && cull && (rc = dem_near(c))) return rc;
    if ((rc = dem_contact_scan(c))) return rc;
    if (c->B > 0) {
        hipLaunchKernelGGL(k_body_reduce, dim3(c->B), dim3(BT), 0, c->stream, c->M, c->B, c->g, c->body_start,
                           c->mx, c->fdrag, c->fcol, c->mvol, body_state(c), last ? 1 : 0, 1, c->cs.rhof,
                           c->cs.dt, c->cs.dt / c->cs.couple_interval, c->cs.gravity[0], c->cs.gravity[1],
                           c->cs.gravity[2], c->M > 0 ? 1 : 0, c->mv, !last && cull ? c->bnear : (int *)nullptr);
        CHECK_LAUNCH(c);
    } else if ((rc = dem_markers_follow(c))) {
        return rc;
    }
    if (!last && cull && (rc = dem_near(c, true))) return rc;
    return dem_rebuild(c, last || !cull || !c->bin_gate ? nullptr : c->bnear + c->B);
}

int dem_check_flags(ibmdem_ctx *c) {
    HIPX(c, hipMemcpyAsync(c->flags_host, c->flags, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    return dem_flags_status(c, c->flags_host);
}

int dem_setup(ibmdem_ctx *c) { (void)c; return IBMDEM_OK; }

// ---------------------------------------------------------------------------
// Per-body probe ring (the reference's PostProcessing probe files,
// fluid.F90:368-404): after each step one small kernel copies the listed
// bodies' position, velocity, angular velocity, hydrodynamic force and
// torque and contact force into slot (count mod capacity) of a device ring
// and advances the device-side count, so recording needs no host wait and
// can be captured with the step.
// ---------------------------------------------------------------------------
constexpr int PROBE_VALUES = 18;

namespace {
__device__ __forceinline__ void probe_append_body(const Blk &kb, int n, int cap, int B, const int *__restrict__ list, BodyState s,
                                                     double *__restrict__ ring, long long *__restrict__ count) {
    __shared__ long long slot;
    if (kb.tx == 0) slot = *count % cap;
    __syncthreads();
    for (int q = kb.tx; q < n * PROBE_VALUES; q += 64) {
        const int i = q / PROBE_VALUES, v = q % PROBE_VALUES, b = list[i];
        const int grp = v / 3, d = v % 3;
        const double *src = grp == 0 ? s.x : grp == 1 ? s.v : grp == 2 ? s.om : grp == 3 ? s.fdrag
                          : grp == 4 ? s.fmom : s.fcol;
        ring[(slot * n + i) * PROBE_VALUES + v] = src[(long)d * B + b];
    }
    __syncthreads();
    if (kb.tx == 0) *count = *count + 1;
}
__global__ __launch_bounds__(64) void k_probe_append(int n, int cap, int B, const int *__restrict__ list, BodyState s,
                                                     double *__restrict__ ring, long long *__restrict__ count) {
    probe_append_body(hw_blk(), n, cap, B, list, s, ring, count);
}
}  // namespace

int dem_probe_append(ibmdem_ctx *c) {
    if (c->probe_n == 0) return IBMDEM_OK;
    hipLaunchKernelGGL(k_probe_append, dim3(1), dim3(64), 0, c->stream, c->probe_n, c->probe_cap, c->B,
                       c->probe_list, body_state(c), c->probe_ring, c->probe_count);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

extern "C" int ibmdem_probe_bodies(ibmdem_ctx *c, int32_t n, const int32_t *bodies, int32_t capacity) {
    if (!c || n < 0 || (n > 0 && (!bodies || capacity < 1))) return IBMDEM_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (bodies[i] < 0 || bodies[i] >= c->B) return set_error(c, IBMDEM_ERR_ARG, "probed body index out of range");
    HIPX(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);   // the step's launches change
    c->probe_n = 0;
    if (n == 0) return IBMDEM_OK;
    const long ring = (long)n * capacity * PROBE_VALUES;
    if (ring > (1L << 30)) return set_error(c, IBMDEM_ERR_ARG, "probe ring larger than 8 GiB");
    int rc;
    if (!c->probe_count && (rc = dev_alloc(c, &c->probe_count, 1))) return rc;
    // buffers are kept and reused by a later, no larger, probe list
    if (n > c->probe_list_cap) {
        if ((rc = dev_alloc(c, &c->probe_list, n))) return rc;
        c->probe_list_cap = n;
    }
    if (ring > c->probe_ring_cap) {
        if ((rc = dev_alloc(c, &c->probe_ring, (size_t)ring))) return rc;
        c->probe_ring_cap = ring;
    }
    HIPX(c, hipMemcpyAsync(c->probe_list, bodies, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPX(c, hipMemsetAsync(c->probe_count, 0, sizeof(long long), c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    c->probe_n = n;
    c->probe_cap = capacity;
    return IBMDEM_OK;
}

extern "C" int ibmdem_probe_read(ibmdem_ctx *c, int32_t *n_steps, int64_t *first_step, double *records) {
    if (!c || !n_steps || !first_step) return IBMDEM_ERR_ARG;
    *n_steps = 0;
    *first_step = 0;
    if (c->probe_n == 0) return IBMDEM_OK;
    long long count = 0;
    HIPX(c, hipMemcpyAsync(&count, c->probe_count, sizeof count, hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    const long long kept = std::min<long long>(count, c->probe_cap);
    *n_steps = (int32_t)kept;
    *first_step = count - kept;
    if (!records || kept == 0) return IBMDEM_OK;
    // oldest first: slots (count - kept .. count - 1) mod capacity
    const size_t rec = (size_t)c->probe_n * PROBE_VALUES;
    const long long s0 = (count - kept) % c->probe_cap;
    const long long tail = std::min<long long>(kept, c->probe_cap - s0);
    HIPX(c, hipMemcpyAsync(records, c->probe_ring + s0 * rec, tail * rec * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    if (kept > tail)
        HIPX(c, hipMemcpyAsync(records + tail * rec, c->probe_ring, (kept - tail) * rec * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    return IBMDEM_OK;
}

// ---------------------------------------------------------------------------
// The DEM part of a batched step: every launch of the step's DEM phase
// (CloudEvolutionMS's hydrodynamic terms, the couple_interval sub-steps,
// the probe record) as one launch over all cases of a group. A case's
// arguments are one DemCase in the plan's arena; each launch's table holds
// the case's grid of blocks and its index there.
// ---------------------------------------------------------------------------
namespace {

struct DemCase {
    int M, B, nb, max_pairs;
    long n;
    GridDev g;
    double *mx, *mv;
    const double *mmass, *mrad, *mvol, *fdrag;
    const int *mbody, *body_start;
    int *mcell, *bin_start, *bin_count, *bin_tmp, *bin_items, *flags, *scan_blocks, *bnear, *blist;
    double *fcol;
    const double *brmax;
    BodyState s;
    ContactParams cp;
    double lim, rhof, dt, dtD, grav[3];
    int probe_n, probe_cap;
    const int *probe_list;
    double *probe_ring;
    long long *probe_count;
};

struct DArg {
    BGrid g;
    int q;   // the case's DemCase
};
using DTab = BTable<DArg>;

__device__ __forceinline__ BodyState glb_state(const BodyState &s) {
    return BodyState{glb(s.x), glb(s.v), glb(s.vold), glb(s.om), glb(s.L), glb(s.Lold), glb(s.rot), glb(s.rotg),
                     glb(s.I), glb(s.m), glb(s.rho), glb(s.fixed), glb(s.fdrag), glb(s.fmom), glb(s.fcol),
                     glb(s.fmomcol), glb(s.fhydro), glb(s.fmomhydro), glb(s.flags)};
}

// the bins' gate of a case (uniform argument gated: a rebuild between two
// sub-steps); the culling runs only where the case has markers and bodies
__device__ __forceinline__ const int *bin_gate(const DemCase &c, int gated) {
    return gated && c.M > 0 && c.B > 0 ? glb((const int *)c.bnear) + c.B : nullptr;
}
__global__ __launch_bounds__(T) void kb_wrap_cells_count(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        wrap_cells_count_body(kb, c.M, c.g, glb(c.mx), glb(c.mcell), glb(c.bin_count), glb(c.flags),
                              bin_gate(c, gated));
    });
}
__global__ __launch_bounds__(T) void kb_chunk_sum(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        if (bins_off(bin_gate(c, gated))) return;
        chunk_sum_body(kb, c.n, glb(c.bin_count), glb(c.scan_blocks));
    });
}
__global__ __launch_bounds__(1024) void kb_scan_blocks(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        if (bins_off(bin_gate(c, gated))) return;
        scan_blocks_body(kb, c.nb, glb(c.scan_blocks));
    });
}
__global__ __launch_bounds__(T) void kb_chunk_scan(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        if (bins_off(bin_gate(c, gated))) return;
        chunk_scan_body(kb, c.n, glb(c.bin_count), glb(c.scan_blocks), glb(c.bin_start), c.nb);
    });
}
__global__ __launch_bounds__(T) void kb_bin_scatter(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        bin_scatter_body(kb, c.M, glb(c.mcell), glb(c.bin_start), glb(c.bin_count), glb(c.bin_tmp),
                         bin_gate(c, gated));
    });
}
__global__ __launch_bounds__(T) void kb_bin_rank(DTab tab, const DemCase *cs, int gated) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        bin_rank_body(kb, c.M, glb(c.mcell), glb(c.bin_start), glb(c.bin_tmp), glb(c.bin_items),
                      bin_gate(c, gated));
    });
}
__global__ __launch_bounds__(64) void kb_body_near_clear(DTab tab, const DemCase *cs) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        body_near_clear_body(kb, c.B, glb(c.bnear));
    });
}
__global__ __launch_bounds__(64) void kb_body_near(DTab tab, const DemCase *cs) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        body_near_body(kb, c.B, c.g, glb((const double *)c.s.x), glb(c.brmax), c.lim, glb(c.bnear), glb(c.blist));
    });
}
__global__ __launch_bounds__(T) void kb_contact(DTab tab, const DemCase *cs) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        contact_body<false>(kb, c.M, c.g, glb((const double *)c.mx), glb((const double *)c.mv), glb(c.mmass),
                            glb(c.mrad), glb(c.mbody), glb((const int *)c.mcell), glb((const int *)c.bin_start),
                            glb((const int *)c.bin_items), glb(c.fcol), c.cp, c.max_pairs, glb(c.flags), 0,
                            (int *)nullptr, (int *)nullptr, (int *)nullptr, glb((const int *)c.bnear),
                            glb((const int *)c.blist), glb((const double *)c.s.x), c.B, glb(c.brmax), c.lim);
    });
}
// uniform arguments: mode, integrate, follow (1: move the markers; 2: also
// clear the next culling's counts)
__global__ __launch_bounds__(BT) void kb_body_reduce(DTab tab, const DemCase *cs, int mode, int integrate,
                                                     int follow) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        body_reduce_body(kb, c.M, c.B, c.g, glb(c.body_start), glb(c.mx), glb(c.fdrag),
                         glb((const double *)c.fcol), glb(c.mvol), glb_state(c.s), mode, integrate, c.rhof, c.dt,
                         c.dtD, c.grav[0], c.grav[1], c.grav[2], follow && c.M > 0 ? 1 : 0, glb(c.mv),
                         follow == 2 && c.M > 0 ? glb(c.bnear) : (int *)nullptr);
    });
}
__global__ __launch_bounds__(T) void kb_follow(DTab tab, const DemCase *cs) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        follow_body(kb, c.M, c.B, c.g, glb(c.mx), glb(c.mv), glb(c.mbody), glb_state(c.s), c.dtD);
    });
}
__global__ __launch_bounds__(64) void kb_probe_append(DTab tab, const DemCase *cs) {
    batch_for(tab, [&](const DArg &a, const Blk &kb) {
        const DemCase &c = *cst(cs + a.q);
        probe_append_body(kb, c.probe_n, c.probe_cap, c.B, glb(c.probe_list), glb_state(c.s), glb(c.probe_ring),
                          glb(c.probe_count));
    });
}

#define DEM_LAUNCHER(name, threads)                                                                   \
    void launch_##name(const BOp &o, hipStream_t s, int) {                                           \
        hipLaunchKernelGGL(kb_##name, dim3(o.blocks), dim3(threads), 0, s, table<DArg>(o),            \
                           (const DemCase *)o.aux);                                                  \
    }
// the rebuild's kernels take the gate switch as o.op
#define DEM_GATED_LAUNCHER(name, threads)                                                             \
    void launch_##name(const BOp &o, hipStream_t s, int) {                                           \
        hipLaunchKernelGGL(kb_##name, dim3(o.blocks), dim3(threads), 0, s, table<DArg>(o),            \
                           (const DemCase *)o.aux, o.op);                                            \
    }
DEM_GATED_LAUNCHER(wrap_cells_count, T)
DEM_GATED_LAUNCHER(chunk_sum, T)
DEM_GATED_LAUNCHER(scan_blocks, 1024)
DEM_GATED_LAUNCHER(chunk_scan, T)
DEM_GATED_LAUNCHER(bin_scatter, T)
DEM_GATED_LAUNCHER(bin_rank, T)
#undef DEM_GATED_LAUNCHER
DEM_LAUNCHER(body_near_clear, 64)
DEM_LAUNCHER(body_near, 64)
DEM_LAUNCHER(contact, T)
DEM_LAUNCHER(follow, T)
DEM_LAUNCHER(probe_append, 64)
#undef DEM_LAUNCHER
void launch_body_reduce(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_body_reduce, dim3(o.blocks), dim3(BT), 0, s, table<DArg>(o), (const DemCase *)o.aux,
                       o.slot, o.op, o.color);
}

// one launch over the cases for which grid(c) is not empty
template <class G>
int dem_op(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, BLaunch fn, int threads, G grid,
           BOp o, std::string &err) {
    std::vector<DArg> args;
    for (int q = 0; q < (int)cs.size(); ++q) {
        const dim3 gr = grid(cs[q]);
        if (gr.x == 0 || gr.y == 0 || gr.z == 0) continue;
        args.push_back(DArg{bgrid(gr, dim3(threads)), q});
    }
    o.aux_off = cases;
    return add_op(P.ar, P.ops, fn, threads, args, o, err);
}

}  // namespace

size_t dem_plan_cases(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs) {
    std::vector<DemCase> v;
    for (ibmdem_ctx *c : cs) {
        DemCase d{};
        d.M = c->M;
        d.B = c->B;
        d.n = c->n;
        d.nb = div_up(c->n, CHUNK);
        d.max_pairs = c->cs.max_pairs;
        d.g = c->g;
        d.mx = c->mx; d.mv = c->mv;
        d.mmass = c->mmass; d.mrad = c->mrad; d.mvol = c->mvol; d.fdrag = c->fdrag;
        d.mbody = c->mbody; d.body_start = c->body_start;
        d.mcell = c->mcell; d.bin_start = c->bin_start; d.bin_count = c->bin_count; d.bin_tmp = c->bin_tmp;
        d.bin_items = c->bin_items; d.flags = c->flags; d.scan_blocks = c->scan_blocks;
        d.bnear = c->bnear; d.blist = c->blist;
        d.fcol = c->fcol;
        d.brmax = c->brmax;
        d.s = body_state(c);
        d.cp = contact_params(c);
        d.lim = near_lim(c);
        d.rhof = c->cs.rhof;
        d.dt = c->cs.dt;
        d.dtD = c->cs.dt / c->cs.couple_interval;
        for (int k = 0; k < 3; ++k) d.grav[k] = c->cs.gravity[k];
        d.probe_n = c->probe_n; d.probe_cap = c->probe_cap; d.probe_list = c->probe_list;
        d.probe_ring = c->probe_ring; d.probe_count = c->probe_count;
        v.push_back(d);
    }
    return P.ar.put(v.data(), v.size() * sizeof(DemCase));
}

// dem_body_hydro of every case
int dem_plan_body_hydro(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    BOp o;
    o.slot = 2;   // mode
    o.op = 0;     // integrate
    return dem_op(P, cases, cs, launch_body_reduce, BT, [](ibmdem_ctx *c) { return dim3(c->B); }, o, err);
}

// dem_substep_fused of every case (the contact scan's body culling with the
// near lists, computed first in the first sub-step and after the body update
// in the others; the caller has made sure no case needs its body radii again)
static int dem_plan_near(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, std::string &err,
                         bool cleared = false) {
    int rc;
    auto near = [](ibmdem_ctx *c) { return (unsigned)(c->M > 0 && c->B > 0 ? div_up(c->B, 64) : 0); };
    if (!cleared && (rc = dem_op(P, cases, cs, launch_body_near_clear, 64,
                                 [&](ibmdem_ctx *c) { return dim3(near(c) ? div_up(c->B + 1, 64) : 0); }, BOp(),
                                 err)))
        return rc;
    return dem_op(P, cases, cs, launch_body_near, 64,
                  [&](ibmdem_ctx *c) { return dim3(near(c), near(c) ? div_up(c->B, NEAR_CH) : 0); }, BOp(), err);
}
int dem_plan_substep(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, bool first, bool last,
                     std::string &err) {
    int rc;
    auto mk = [](ibmdem_ctx *c) { return (unsigned)(c->M > 0 ? div_up(c->M, T) : 0); };
    // dem_contact_scan, after the culling
    if (first && (rc = dem_plan_near(P, cases, cs, err))) return rc;
    if ((rc = dem_op(P, cases, cs, launch_contact, T, [&](ibmdem_ctx *c) { return dim3(mk(c)); }, BOp(), err)))
        return rc;
    BOp o;
    o.slot = last ? 1 : 0;   // mode
    o.op = 1;                // integrate
    o.color = last ? 1 : 2;  // follow: the markers moved by their body's work-group (2: and the next
                             // culling's counts cleared)
    if ((rc = dem_op(P, cases, cs, launch_body_reduce, BT, [](ibmdem_ctx *c) { return dim3(c->B); }, o, err)))
        return rc;
    // cases without bodies move their markers (if any) by k_follow's rule
    if ((rc = dem_op(P, cases, cs, launch_follow, T,
                     [&](ibmdem_ctx *c) { return dim3(c->B > 0 ? 0u : mk(c)); }, BOp(), err)))
        return rc;
    if (!last && (rc = dem_plan_near(P, cases, cs, err, true))) return rc;
    // dem_rebuild, gated between sub-steps
    BOp gb;
    gb.op = last ? 0 : 1;
    for (ibmdem_ctx *c : cs) gb.op = c->bin_gate ? gb.op : 0;
    auto chunks = [](ibmdem_ctx *c) { return dim3(c->M > 0 ? div_up(c->n, CHUNK) : 0); };
    if ((rc = dem_op(P, cases, cs, launch_wrap_cells_count, T, [&](ibmdem_ctx *c) { return dim3(mk(c)); }, gb,
                     err)))
        return rc;
    if ((rc = dem_op(P, cases, cs, launch_chunk_sum, T, chunks, gb, err))) return rc;
    if ((rc = dem_op(P, cases, cs, launch_scan_blocks, 1024, [](ibmdem_ctx *c) { return dim3(c->M > 0 ? 1 : 0); },
                     gb, err)))
        return rc;
    if ((rc = dem_op(P, cases, cs, launch_chunk_scan, T, chunks, gb, err))) return rc;
    if ((rc = dem_op(P, cases, cs, launch_bin_scatter, T, [&](ibmdem_ctx *c) { return dim3(mk(c)); }, gb, err)))
        return rc;
    return dem_op(P, cases, cs, launch_bin_rank, T, [&](ibmdem_ctx *c) { return dim3(mk(c)); }, gb, err);
}

// dem_probe_append of every case that records probes
int dem_plan_probe(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    return dem_op(P, cases, cs, launch_probe_append, 64, [](ibmdem_ctx *c) { return dim3(c->probe_n > 0 ? 1 : 0); },
                  BOp(), err);
}

// the step's error flags as dem_check_flags reports them, from a host copy
int dem_flags_status(ibmdem_ctx *c, const int *f) {
    if (!(f[FL_NAN] || f[FL_OUTSIDE] || f[FL_CAPACITY] || f[FL_SPREAD])) return IBMDEM_OK;
    HIPX(c, hipMemsetAsync(c->flags, 0, FL_COUNT * sizeof(int), c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    if (f[FL_NAN]) return set_error(c, IBMDEM_ERR_NAN, "non-finite marker position or body state");
    if (f[FL_OUTSIDE]) return set_error(c, IBMDEM_ERR_OUTSIDE, "a marker left the domain through a wall");
    if (f[FL_CAPACITY]) return set_error(c, IBMDEM_ERR_CAPACITY, "contact candidates exceed maxCollisionNo");
    return set_error(c, IBMDEM_ERR_CAPACITY, "IBM spread entry list exceeds its capacity");
}
