// poisson.hip — pressure Poisson solve: multigrid-preconditioned CG.
//
// Replaces SetPressureVector + PoissonSolver (PETSc KSPSolve, GMRES(30) +
// BoomerAMG) + GetLocalArray (poisson.F90:107-186).
//
// Operator (poisson.F90:74-94, initialization.F90:105-173): the 7-point
// -lap_h with wall neighbours omitted (Neumann) and periodic neighbours
// wrapped; (Ax)_i = diag_i x_i - nbsum_i, nbsum = (xw+xe)/hx^2 +
// (xs+xn)/hy^2 + (xb+xt)/hz^2, wall neighbours contributing 0.
// The reference pins row P'(1,1,1) to the identity with b[0] = 0
// (poisson.F90:78,97-100,126); that system's solution is the solution of
// the singular system with b[0] := -sum_{i>0} b[i], shifted so that
// P'(1,1,1) = 0, which is what is solved here.
// Inflow sides are Neumann like walls. An outflow side (no counterpart in
// the reference) holds P' = 0 on the face: the ghost is -P', which adds
// 2/h^2 to the diagonal of the last cell; the operator is then
// non-singular, and there is no pinned row and no gauge shift.
//
// Preconditioner: one symmetric V-cycle. Smoother red-black Gauss-Seidel
// (pre: red then black, post: black then red); coarsest level: symmetric
// red-black sweeps in one work-group (damped Jacobi if a periodic dimension
// is odd); transfers: cell-centred trilinear prolongation (3/4, 1/4 per
// dimension) and restriction = prolongation^T / 8, with wall mirror and
// periodic wrap index rules. In the mixed form the whole V-cycle runs in
// FP32 on FP32 level arrays; the CG vectors, matrix-vector product, inner
// products and the stopping test stay FP64.
//
// Reductions have a fixed association (fixed block counts, LDS trees), so
// results are run-to-run identical.
#include "common.hpp"
#include "batched.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>

namespace {

using namespace bk;

constexpr int RB = 256;           // threads per block for level kernels
constexpr int RED_BLOCKS = 1024;
constexpr int RED_T = 256;
constexpr int COARSE_T = 1024;

struct LaunchShape {
    dim3 grid, block;
};

// Barrier of one wavefront over LDS: its LDS operations complete in program
// order, so a compiler-level fence suffices.
__device__ __forceinline__ void wave_barrier_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
LaunchShape shape_for(int nx, int ny, int nz);
// grid of a red-black half sweep (RB_ROWS rows per thread) on an nx*ny*nz level
dim3 half_grid(int nx, int ny, int nz);
// Blocks always hold RB threads (bx * by = RB, rows past the level masked),
// so a launch over a batch of cases can run every case's blocks with one
// block size and the same block decomposition as the case's own launch.
LaunchShape shape_for(int nx, int ny, int nz) {
    (void)ny;
    int bx = std::min(64, pow2ceil(nx));
    int by = RB / bx;
    return {dim3(div_up(nx, bx), div_up(ny, by), nz), dim3(bx, by, 1)};
}

// nbsum in the oracle's association: ((xw+xe)*ihx2 + (xs+xn)*ihy2) + (xb+xt)*ihz2,
// wall neighbours 0, periodic neighbours wrapped; x(o) is the value at
// storage position o
template <typename T, typename F>
__device__ __forceinline__ T nbsum_f(const F &x, const LevelDev &L, int i, int j, int k) {
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    const bool pxw = L.bt[0] == IBMDEM_BC_PERIODIC, pyw = L.bt[2] == IBMDEM_BC_PERIODIC,
               pzw = L.bt[4] == IBMDEM_BC_PERIODIC;
    const bool hw = i > 0 || pxw, he = i < L.nx - 1 || pxw;
    const bool hs = j > 0 || pyw, hn = j < L.ny - 1 || pyw;
    const bool hb = k > 0 || pzw, ht = k < L.nz - 1 || pzw;
    const int iw = i > 0 ? i - 1 : L.nx - 1, ie = i < L.nx - 1 ? i + 1 : 0;
    const int js = j > 0 ? j - 1 : L.ny - 1, jn = j < L.ny - 1 ? j + 1 : 0;
    const int kb = k > 0 ? k - 1 : L.nz - 1, kt = k < L.nz - 1 ? k + 1 : 0;
    const T xw = hw ? x(lidx(L, iw, j, k)) : T(0), xe = he ? x(lidx(L, ie, j, k)) : T(0);
    const T xs = hs ? x(lidx(L, i, js, k)) : T(0), xn = hn ? x(lidx(L, i, jn, k)) : T(0);
    const T xb = hb ? x(lidx(L, i, j, kb)) : T(0), xt = ht ? x(lidx(L, i, j, kt)) : T(0);
    const T sx = (xw + xe) * ihx2;
    const T syy = (xs + xn) * ihy2;
    const T szz = (xb + xt) * ihz2;
    return sx + syy + szz;
}
template <typename T>
__device__ __forceinline__ T nbsum(const T *__restrict__ x, const LevelDev &L, int i, int j, int k) {
    return nbsum_f<T>([&](int o) { return x[o]; }, L, i, j, k);
}

// per-dimension diagonal count: 1 per present neighbour, 2 for an outflow
// (Dirichlet) face
__device__ __forceinline__ int diag_count(int i, int n, int lo, int hi) {
    int c = 2 - ((i == 0 && lo != IBMDEM_BC_PERIODIC) ? 1 : 0) - ((i == n - 1 && hi != IBMDEM_BC_PERIODIC) ? 1 : 0);
    if (i == n - 1 && hi == IBMDEM_BC_OUTFLOW) c += 2;
    return c;
}
// the diagonal table entry of a cell with both neighbours in all three
// dimensions (counts 2, 2, 2)
constexpr int DIAG_INTERIOR = 2 * 16 + 2 * 4 + 2;
__device__ __forceinline__ int diag_index(const LevelDev &L, int i, int j, int k) {
    return diag_count(i, L.nx, L.bt[0], L.bt[1]) * 16 + diag_count(j, L.ny, L.bt[2], L.bt[3]) * 4 +
           diag_count(k, L.nz, L.bt[4], L.bt[5]);
}

// The x-pair (2p, 2p+1) of row (j, k) in colour-split storage: both cells
// sit at position pos of their colour halves (the even cell in half ca), and
// every neighbour of either cell lies in the other cell's half, at pos or pos
// plus one of these offsets (wrapped across a periodic side; the h* flags
// are false across any other side). With these the stencils below evaluate
// nbsum's expression without per-neighbour index arithmetic.
struct PairGeo {
    int pos, ca;
    int dw, de, ds, dn, db, dt;   // even cell's west, odd cell's east, then y and z
    bool hw, he, hs, hn, hb, ht;
    int cx0, cx1, cyz;            // diagonal counts: x of the even / odd cell, cy*4 + cz
};
__device__ __forceinline__ PairGeo pair_geo(const LevelDev &L, int p, int j, int k) {
    const int rl = L.nx >> 1, pl = rl * L.ny;
    const bool px = L.bt[0] == IBMDEM_BC_PERIODIC, py = L.bt[2] == IBMDEM_BC_PERIODIC,
               pz = L.bt[4] == IBMDEM_BC_PERIODIC;
    PairGeo g;
    g.pos = (k * L.ny + j) * rl + p;
    g.ca = (j + k) & 1;
    g.hw = p > 0 || px;          g.dw = p > 0 ? -1 : rl - 1;
    g.he = p < rl - 1 || px;     g.de = p < rl - 1 ? 1 : 1 - rl;
    g.hs = j > 0 || py;          g.ds = j > 0 ? -rl : (L.ny - 1) * rl;
    g.hn = j < L.ny - 1 || py;   g.dn = j < L.ny - 1 ? rl : -(L.ny - 1) * rl;
    g.hb = k > 0 || pz;          g.db = k > 0 ? -pl : (L.nz - 1) * pl;
    g.ht = k < L.nz - 1 || pz;   g.dt = k < L.nz - 1 ? pl : -(L.nz - 1) * pl;
    g.cx0 = diag_count(2 * p, L.nx, L.bt[0], L.bt[1]);
    g.cx1 = diag_count(2 * p + 1, L.nx, L.bt[0], L.bt[1]);
    g.cyz = diag_count(j, L.ny, L.bt[2], L.bt[3]) * 4 + diag_count(k, L.nz, L.bt[4], L.bt[5]);
    return g;
}
// nbsum of the pair's even (h = 0) or odd (h = 1) cell; q(d) is the value
// at offset d from the pair's position in the other cell's colour half
template <typename T, typename F>
__device__ __forceinline__ T nbsum_pair_f(const F &q, const PairGeo &g, int h, const LevelDev &L) {
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    T xw, xe;
    if (h == 0) {
        xw = g.hw ? q(g.dw) : T(0);
        xe = q(0);
    } else {
        xw = q(0);
        xe = g.he ? q(g.de) : T(0);
    }
    const T xs = g.hs ? q(g.ds) : T(0), xn = g.hn ? q(g.dn) : T(0);
    const T xb = g.hb ? q(g.db) : T(0), xt = g.ht ? q(g.dt) : T(0);
    const T sx = (xw + xe) * ihx2;
    const T syy = (xs + xn) * ihy2;
    const T szz = (xb + xt) * ihz2;
    return sx + syy + szz;
}
// nbsum_pair_f's value from an array. Every neighbour is loaded without a
// branch (an absent one from offset 0, the pair's own position, and then
// replaced by 0), so a thread's loads can all be in flight at once.
template <typename T>
__device__ __forceinline__ T nbsum_pair(const T *__restrict__ oth, const PairGeo &g, int h, const LevelDev &L) {
    const T *q = oth + g.pos;
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    const bool pw = h == 0 ? g.hw : true, pe = h == 0 ? true : g.he;
    const T vw = q[h == 0 && g.hw ? g.dw : 0], ve = q[h != 0 && g.he ? g.de : 0];
    const T vs = q[g.hs ? g.ds : 0], vn = q[g.hn ? g.dn : 0];
    const T vb = q[g.hb ? g.db : 0], vt = q[g.ht ? g.dt : 0];
    const T xw = pw ? vw : T(0), xe = pe ? ve : T(0);
    const T xs = g.hs ? vs : T(0), xn = g.hn ? vn : T(0);
    const T xb = g.hb ? vb : T(0), xt = g.ht ? vt : T(0);
    const T sx = (xw + xe) * ihx2;
    const T syy = (xs + xn) * ihy2;
    const T szz = (xb + xt) * ihz2;
    return sx + syy + szz;
}

// nbsum_pair from the loaded values: the pair value, the other x neighbour
// (west for h = 0, east for h = 1), south, north, below, above (each read
// at offset 0 when absent, and replaced by 0 here)
template <typename T>
__device__ __forceinline__ T nbsum_pair_vals(const PairGeo &g, int h, T part, T xo, T ys, T yn, T zb, T zt,
                                             const LevelDev &L) {
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    const bool pw = h == 0 ? g.hw : true, pe = h == 0 ? true : g.he;
    const T vw = h == 0 ? xo : part, ve = h == 0 ? part : xo;
    const T xw = pw ? vw : T(0), xe = pe ? ve : T(0);
    const T xs = g.hs ? ys : T(0), xn = g.hn ? yn : T(0);
    const T xb = g.hb ? zb : T(0), xt = g.ht ? zt : T(0);
    const T sx = (xw + xe) * ihx2;
    const T syy = (xs + xn) * ihy2;
    const T szz = (xb + xt) * ihz2;
    return sx + syy + szz;
}

// one red-black half sweep; each thread owns the cell of the colour in a
// pair of x cells, in RB_ROWS = 2 adjacent rows (independent cells, so
// their loads are in flight together, two of them shared). ZI: the level's
// x is implicitly 0 (first sweep of a V-cycle visit); the other colour is
// written 0. DOT (the last two level-0 sweeps of the preconditioner, whose
// b is the CG residual r and whose x is z = M^-1 r): also one partial sum of
// r.z over the block's updated cells into partial[pofs + block]; in the
// mixed form (z64 set) r is read from the FP64 r64 and the sweep also
// writes z64 = double(x), the FP64 z of its cells.
// RES = 1 (the last pre-smoothing sweep): also the residual of the updated
// cells, r = b - (diag x - nbsum) with the sweep's own nbsum (the other
// colour does not change in between); RES = 2: only that residual, for the
// cells of `color`, x unchanged. Together they give r = b - A x.
constexpr int RB_ROWS = 2;
static_assert(RB_ROWS == 2, "the half sweep shares loads between its two adjacent rows");
// Two-value accesses of a colour half (8 bytes in FP32, 16 in FP64).
template <typename T>
struct Vec2;
template <>
struct Vec2<float> { using type = float2; };
template <>
struct Vec2<double> { using type = double2; };
template <typename T>
__device__ __forceinline__ void ld2(const T *p, T &v0, T &v1) {
    const typename Vec2<T>::type v = *reinterpret_cast<const typename Vec2<T>::type *>(p);
    v0 = v.x;
    v1 = v.y;
}
template <typename T>
__device__ __forceinline__ void st2(T *p, T v0, T v1) {
    typename Vec2<T>::type v;
    v.x = v0;
    v.y = v1;
    *reinterpret_cast<typename Vec2<T>::type *>(p) = v;
}

// A half sweep with each thread on two adjacent x-pairs (pc, pc + 1, pc
// even) of its RB_ROWS rows, on split levels whose rows hold an even number
// of pairs: the two pairs' cells are adjacent in every colour half (and
// every row starts at an even position), so each load and store of the
// thread moves both values at once, half the memory instructions per cell
// of rb_half_body. Each cell's neighbours, their association and the
// update are rb_half_body's (nbsum_pair_vals), so x, r and z are the same
// bit for bit; the r.z partial sums (DOT) follow this kernel's blocks.
template <typename T, bool DOT, int RES, bool ZI>
__device__ __forceinline__ void rb_half2_body(const Blk &kb, const LevelDev &L, T *__restrict__ x,
                                              const T *__restrict__ b, int color, double *__restrict__ partial,
                                              long pofs, T *__restrict__ r, const double *__restrict__ r64,
                                              double *__restrict__ z64) {
    const T id_int = T(L.idiag[DIAG_INTERIOR]), dg_int = T(L.diag[DIAG_INTERIOR]);
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    const int rl = L.nx >> 1, pl = rl * L.ny;
    const int p0 = 2 * (kb.bx * kb.dx + kb.tx), k = kb.bz;
    const int j0 = (kb.by * kb.dy + kb.ty) * RB_ROWS;
    const bool px = L.bt[0] == IBMDEM_BC_PERIODIC, py = L.bt[2] == IBMDEM_BC_PERIODIC,
               pz = L.bt[4] == IBMDEM_BC_PERIODIC;
    // pairs past the row's end load from the last two (valid addresses)
    // and are dropped at the stores
    const int pc = min(p0, rl - 2);
    T *own = x + (long)color * L.half;
    T *oth = x + (long)(1 - color) * L.half;
    // x neighbours outside the two pairs: the odd cell of pair pc - 1 (west
    // of an even cell of pair pc) and the even cell of pair pc + 2 (east of
    // an odd cell of pair pc + 1), as offsets from pair pc
    const bool hw = pc > 0 || px, he = pc + 2 < rl || px;
    const int dw = pc > 0 ? -1 : rl - 1, de = pc + 2 < rl ? 2 : 2 - rl;
    const bool hb = k > 0 || pz, ht = k < L.nz - 1 || pz;
    const int db = k > 0 ? -pl : (L.nz - 1) * pl, dt = k < L.nz - 1 ? pl : -(L.nz - 1) * pl;
    const int cz = diag_count(k, L.nz, L.bt[4], L.bt[5]);
    bool ok[RB_ROWS], hs[RB_ROWS], hn[RB_ROWS];
    int pos[RB_ROWS], h[RB_ROWS], cyz[RB_ROWS];
    T bo[RB_ROWS][2], av[RB_ROWS][2], xo[RB_ROWS], ys[RB_ROWS][2], yn[RB_ROWS][2], zb[RB_ROWS][2],
        zt[RB_ROWS][2], ov[RB_ROWS][2];
    const bool adj = j0 + 1 < L.ny;   // both rows exist: they share their y loads
#pragma unroll
    for (int rr = 0; rr < RB_ROWS; ++rr) {
        const int j = min(j0 + rr, L.ny - 1);
        ok[rr] = j0 + rr < L.ny && p0 < rl;
        pos[rr] = (k * L.ny + j) * rl + pc;
        h[rr] = ((j + k) & 1) == color ? 0 : 1;
        hs[rr] = j > 0 || py;
        hn[rr] = j < L.ny - 1 || py;
        cyz[rr] = diag_count(j, L.ny, L.bt[2], L.bt[3]) * 4 + cz;
        const T *q = oth + pos[rr];
        ld2(b + (long)color * L.half + pos[rr], bo[rr][0], bo[rr][1]);
        if (RES == 2) ld2(own + pos[rr], ov[rr][0], ov[rr][1]);
        if (!ZI || RES == 2) {
            ld2(q, av[rr][0], av[rr][1]);
            xo[rr] = h[rr] == 0 ? q[hw ? dw : 0] : q[he ? de : 0];
            ld2(q + (hb ? db : 0), zb[rr][0], zb[rr][1]);
            ld2(q + (ht ? dt : 0), zt[rr][0], zt[rr][1]);
        }
    }
    T sn[RB_ROWS][2];
    if (ZI && RES != 2) {
#pragma unroll
        for (int rr = 0; rr < RB_ROWS; ++rr) sn[rr][0] = sn[rr][1] = T(0);
    } else {
        // row 0's north neighbours are row 1's pair values and row 1's
        // south neighbours row 0's (when both rows exist)
        const T *q0 = oth + pos[0], *q1 = oth + pos[1];
        const int jr0 = min(j0, L.ny - 1), jr1 = min(j0 + 1, L.ny - 1);   // the rows' (clamped) indices
        const int ds0 = jr0 > 0 ? -rl : (L.ny - 1) * rl;
        const int dn1 = jr1 < L.ny - 1 ? rl : -(L.ny - 1) * rl;
        ld2(q0 + (hs[0] ? ds0 : 0), ys[0][0], ys[0][1]);
        ld2(q1 + (hn[1] ? dn1 : 0), yn[1][0], yn[1][1]);
        if (adj) {
            yn[0][0] = av[1][0]; yn[0][1] = av[1][1];
            ys[1][0] = av[0][0]; ys[1][1] = av[0][1];
        } else {
            const int dn0 = jr0 < L.ny - 1 ? rl : -(L.ny - 1) * rl;
            ld2(q0 + (hn[0] ? dn0 : 0), yn[0][0], yn[0][1]);
            ys[1][0] = ys[0][0]; ys[1][1] = ys[0][1];   // row 1 does not exist
        }
#pragma unroll
        for (int rr = 0; rr < RB_ROWS; ++rr)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                // nbsum_pair_vals of the cell of pair pc + c: its pair value
                // and its other x neighbour (inside the two pairs but for
                // the outer ones)
                const bool even = h[rr] == 0;
                const T part = av[rr][c];
                const T oxv = even ? (c == 0 ? xo[rr] : av[rr][0]) : (c == 0 ? av[rr][1] : xo[rr]);
                const bool pw = even ? (c == 0 ? hw : true) : true;
                const bool pe = even ? true : (c == 0 ? true : he);
                const T vw = even ? oxv : part, ve = even ? part : oxv;
                const T xw = pw ? vw : T(0), xe = pe ? ve : T(0);
                const T xs = hs[rr] ? ys[rr][c] : T(0), xn = hn[rr] ? yn[rr][c] : T(0);
                const T xb = hb ? zb[rr][c] : T(0), xt = ht ? zt[rr][c] : T(0);
                const T sx = (xw + xe) * ihx2;
                const T syy = (xs + xn) * ihy2;
                const T szz = (xb + xt) * ihz2;
                sn[rr][c] = sx + syy + szz;
            }
    }
    double dsum = 0.0;
#pragma unroll
    for (int rr = 0; rr < RB_ROWS; ++rr) {
        if (!ok[rr]) continue;
        const int o = pos[rr];
        const long go = (long)color * L.half + o;
        T out[2], res[2];
        double z[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int di = diag_count(2 * (pc + c) + h[rr], L.nx, L.bt[0], L.bt[1]) * 16 + cyz[rr];
            if (RES == 2) {
                T dg;
                if (di == DIAG_INTERIOR) dg = dg_int;
                else dg = T(L.diag[di]);
                res[c] = bo[rr][c] - (dg * ov[rr][c] - sn[rr][c]);
                continue;
            }
            T id;
            if (di == DIAG_INTERIOR) id = id_int;
            else id = T(L.idiag[di]);
            if (ZI) {
                out[c] = (bo[rr][c] + T(0)) * id;
            } else {
                const T xn = (bo[rr][c] + sn[rr][c]) * id;
                out[c] = xn;
                if (RES == 1) {
                    T dg;
                    if (di == DIAG_INTERIOR) dg = dg_int;
                    else dg = T(L.diag[di]);
                    res[c] = bo[rr][c] - (dg * xn - sn[rr][c]);
                }
                z[c] = double(xn);
            }
        }
        if (RES == 2) {
            st2(r + go, res[0], res[1]);
            continue;
        }
        if (ZI) st2(oth + o, T(0), T(0));
        st2(own + o, out[0], out[1]);
        if (RES == 1) st2(r + go, res[0], res[1]);
        if (DOT) {
            if (z64) {
                double r0, r1;
                ld2(r64 + go, r0, r1);
                st2(z64 + go, z[0], z[1]);
                dsum += r0 * z[0];
                dsum += r1 * z[1];
            } else {
                dsum += double(bo[rr][0]) * z[0];
                dsum += double(bo[rr][1]) * z[1];
            }
        }
    }
    if (DOT) {
        __shared__ double red[RB];
        const int nt = kb.dx * kb.dy, t = kb.ty * kb.dx + kb.tx;
        red[t] = dsum;
        __syncthreads();
        for (int st = nt / 2; st > 0; st >>= 1) {
            if (t < st) red[t] += red[t + st];
            __syncthreads();
        }
        if (t == 0) partial[pofs + ((long)kb.bz * kb.gy + kb.by) * kb.gx + kb.bx] = red[0];
    }
}

template <typename T, typename TB, bool DOT = false, int RES = 0, bool ZI = false, bool X2 = false>
__device__ __forceinline__ void rb_half_body(const Blk &kb, const LevelDev &L, T *__restrict__ x, const TB *__restrict__ b,
                                                int color, const double *__restrict__ stop,
                                                double *__restrict__ partial, long pofs, T *__restrict__ r,
                                                const double *__restrict__ r64,
                                                double *__restrict__ z64) {
    if (*stop != 0.0) return;
    if constexpr (X2) {
        static_assert(std::is_same<T, TB>::value, "the two-pair sweep reads b in the level's type");
        rb_half2_body<T, DOT, RES, ZI>(kb, L, x, b, color, partial, pofs, r, r64, z64);
        return;
    }
    const T id_int = T(L.idiag[DIAG_INTERIOR]), dg_int = T(L.diag[DIAG_INTERIOR]);
    const int ip = kb.bx * kb.dx + kb.tx;
    const int k = kb.bz;
    double dsum = 0.0;
    if (L.split) {
        // Two phases over the thread's rows: every load (b, the other
        // colour's neighbours, the own value of the residual pass) is issued
        // before the first store, so the rows' loads are in flight together.
        // Stores go to this colour's half and x is read from the other half
        // (the residual pass reads its own cells but does not write x), so
        // the order changes no value.
        T *own = x + (long)color * L.half;
        T *oth = x + (long)(1 - color) * L.half;
        bool ok[RB_ROWS];
        int pos[RB_ROWS], dix[RB_ROWS];
        TB bo[RB_ROWS];
        T sn[RB_ROWS], ov[RB_ROWS];
        // the thread's two rows are adjacent (j0, j0 + 1); rows and pairs
        // past the level's end load from the last one (valid addresses, no
        // branch) and are dropped in phase two
        const int j0 = (kb.by * kb.dy + kb.ty) * RB_ROWS;
        const int ipc = min(ip, (L.nx >> 1) - 1);
        PairGeo g[RB_ROWS];
        int h[RB_ROWS];
#pragma unroll
        for (int rr = 0; rr < RB_ROWS; ++rr) {
            const int j = j0 + rr;
            ok[rr] = j < L.ny && ip < (L.nx >> 1);
            g[rr] = pair_geo(L, ipc, min(j, L.ny - 1), k);
            h[rr] = g[rr].ca == color ? 0 : 1;   // which cell of the pair has this colour
            pos[rr] = g[rr].pos;
            dix[rr] = (h[rr] ? g[rr].cx1 : g[rr].cx0) * 16 + g[rr].cyz;
            bo[rr] = b[(long)color * L.half + g[rr].pos];
            if (RES == 2) ov[rr] = own[g[rr].pos];
        }
        if (ZI && RES != 2) {
            sn[0] = sn[1] = T(0);
        } else {
            // nbsum_pair of both rows with two loads shared: when both rows
            // exist, row 0's north neighbour is row 1's pair value and row
            // 1's south neighbour is row 0's
            const T *q0 = oth + g[0].pos, *q1 = oth + g[1].pos;
            const T a = q0[0], c = q1[0];
            const T ys0 = q0[g[0].hs ? g[0].ds : 0], yn1 = q1[g[1].hn ? g[1].dn : 0];
            const T xo0 = q0[h[0] == 0 ? (g[0].hw ? g[0].dw : 0) : (g[0].he ? g[0].de : 0)];
            const T xo1 = q1[h[1] == 0 ? (g[1].hw ? g[1].dw : 0) : (g[1].he ? g[1].de : 0)];
            const T zb0 = q0[g[0].hb ? g[0].db : 0], zt0 = q0[g[0].ht ? g[0].dt : 0];
            const T zb1 = q1[g[1].hb ? g[1].db : 0], zt1 = q1[g[1].ht ? g[1].dt : 0];
            T yn0 = c, ys1 = a;
            if (j0 + 1 >= L.ny) yn0 = q0[g[0].hn ? g[0].dn : 0];   // row 1 does not exist
            sn[0] = nbsum_pair_vals(g[0], h[0], a, xo0, ys0, yn0, zb0, zt0, L);
            sn[1] = nbsum_pair_vals(g[1], h[1], c, xo1, ys1, yn1, zb1, zt1, L);
        }
#pragma unroll
        for (int rr = 0; rr < RB_ROWS; ++rr) {
            if (!ok[rr]) continue;
            const int o = pos[rr], di = dix[rr];
            if (RES == 2) {
                T dg;
                if (di == DIAG_INTERIOR) dg = dg_int;
                else dg = T(L.diag[di]);
                r[(long)color * L.half + o] = T(bo[rr]) - (dg * ov[rr] - sn[rr]);
                continue;
            }
            // cells with both neighbours in every dimension (all but the
            // boundary layers) share table entry DIAG_INTERIOR, read once
            T id;
            if (di == DIAG_INTERIOR) id = id_int;
            else id = T(L.idiag[di]);
            if (ZI) {
                oth[o] = T(0);
                own[o] = (T(bo[rr]) + T(0)) * id;
            } else {
                const T xn = (T(bo[rr]) + sn[rr]) * id;
                own[o] = xn;
                if (RES == 1) {
                    T dg;
                    if (di == DIAG_INTERIOR) dg = dg_int;
                    else dg = T(L.diag[di]);
                    r[(long)color * L.half + o] = T(bo[rr]) - (dg * xn - sn[rr]);
                }
                if (DOT) {
                    const long go = (long)color * L.half + o;
                    if (z64) {
                        z64[go] = double(xn);
                        dsum += r64[go] * double(xn);
                    } else {
                        dsum += double(bo[rr]) * double(xn);
                    }
                }
            }
        }
    } else {
#pragma unroll
        for (int rr = 0; rr < RB_ROWS; ++rr) {
            const int j = (kb.by * kb.dy + kb.ty) * RB_ROWS + rr;
            if (j >= L.ny) continue;
            const int i = 2 * ip + ((j + k + color) & 1);
            const int o = lidx(L, i, j, k);
            if (ZI) {
                const int io = 2 * ip + ((j + k + color + 1) & 1);
                if (io < L.nx) x[lidx(L, io, j, k)] = T(0);
                if (i >= L.nx) continue;
                const T id = T(L.idiag[diag_index(L, i, j, k)]);
                x[o] = (T(b[o]) + T(0)) * id;
                continue;
            }
            if (i >= L.nx) continue;
            const int di = diag_index(L, i, j, k);
            const T sn = nbsum(x, L, i, j, k);
            if (RES == 2) {
                r[o] = T(b[o]) - (T(L.diag[di]) * x[o] - sn);
                continue;
            }
            const T xn = (T(b[o]) + sn) * T(L.idiag[di]);
            x[o] = xn;
            if (RES == 1) r[o] = T(b[o]) - (T(L.diag[di]) * xn - sn);
            if (DOT) {
                if (z64) {
                    z64[o] = double(xn);
                    dsum += r64[o] * double(xn);
                } else {
                    dsum += double(b[o]) * double(xn);
                }
            }
        }
    }
    if (DOT) {
        __shared__ double red[RB];
        const int nt = kb.dx * kb.dy, t = kb.ty * kb.dx + kb.tx;
        red[t] = dsum;
        __syncthreads();
        for (int st = nt / 2; st > 0; st >>= 1) {
            if (t < st) red[t] += red[t + st];
            __syncthreads();
        }
        if (t == 0) partial[pofs + ((long)kb.bz * kb.gy + kb.by) * kb.gx + kb.bx] = red[0];
    }
}
template <typename T, typename TB, bool DOT = false, int RES = 0, bool ZI = false, bool X2 = false>
__global__ __launch_bounds__(RB) void k_rb_half(LevelDev L, T *__restrict__ x, const TB *__restrict__ b,
                                                int color, const double *__restrict__ stop,
                                                double *__restrict__ partial, long pofs, T *__restrict__ r,
                                                const double *__restrict__ r64 = nullptr,
                                                double *__restrict__ z64 = nullptr) {
    rb_half_body<T, TB, DOT, RES, ZI, X2>(hw_blk(), L, x, b, color, stop, partial, pofs, r, r64, z64);
}

dim3 half_grid(int nx, int ny, int nz) {
    const LaunchShape rb = shape_for(div_up(nx, 2), ny, nz);
    return dim3(rb.grid.x, div_up(ny, RB_ROWS * (int)rb.block.y), rb.grid.z);
}

// The half sweep's launch on a level: two-pair threads (rb_half2_body) where
// the level is split and its rows hold an even number of pairs, else one
// pair per thread.
struct SweepShape {
    dim3 grid, block;
    bool x2;
};
SweepShape sweep_shape(const LevelDev &d) {
    const bool x2 = d.split && (d.nx >> 1) % 2 == 0;
    const LaunchShape sh = shape_for(x2 ? d.nx / 4 : div_up(d.nx, 2), d.ny, d.nz);
    return {dim3(sh.grid.x, div_up(d.ny, RB_ROWS * (int)sh.block.y), sh.grid.z), sh.block, x2};
}

// cell index of a virtual coordinate one step beyond the level: a wall side
// clamps to the edge (a coarse cell's restriction weighs the edge fine cell
// twice; a fine cell's neighbour parent there is its own parent, see
// parents), a periodic side wraps. Coordinates further out only occur in the
// unused part of a tile and are clamped to some valid cell.
__device__ __forceinline__ int tile_coord(int v, int n, int lo, int hi) {
    if (v < 0) v = lo != IBMDEM_BC_PERIODIC ? 0 : v + n;
    else if (v >= n) v = hi != IBMDEM_BC_PERIODIC ? n - 1 : v - n;
    return min(max(v, 0), n - 1);
}

// b_c = (1/8) sum w r_f over the 4x4x4 fine cells, w = wz*wy*wx from
// (1/4,3/4,3/4,1/4). A block owns RS_CX x RS_CY coarse cells of one coarse
// plane; the fine residual they weigh (2 RS_CX + 2 by 2 RS_CY + 2 by 4 fine
// cells at virtual coordinates, mapped by tile_coord) is staged in LDS with
// coalesced loads, and each thread sums its 64 terms from there in the
// (z, y, x) order. With xc, the coarse level's first smoothing sweep (x = 0
// before it: colour 0 takes (b + 0)/diag, colour 1 stays 0) is done here
// too, each cell's value as that sweep computes it from the b just written,
// and the coarse level's V-cycle starts with its second half sweep.
constexpr int RS_CX = 32, RS_CY = 8;
constexpr int RS_RX = 2 * RS_CX + 2, RS_RY = 2 * RS_CY + 2, RS_TILE = RS_RX * RS_RY * 4;
static_assert(RS_CX * RS_CY == RB, "one coarse cell per thread");
template <typename T>
__device__ __forceinline__ void restrict_body(const Blk &kb, const LevelDev &F, const T *__restrict__ rf, const LevelDev &Cl,
                                                 T *__restrict__ bc, T *__restrict__ xc, const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    __shared__ T tile[RS_TILE];
    const int t = kb.tx;
    const int I0 = kb.bx * RS_CX, J0 = kb.by * RS_CY, K = kb.bz;
    // The tile's rows r = z RS_RY + y: wavefront w loads columns 0..63 of
    // the rows of fine plane z = w; a row's (j, k) and its part of the
    // storage index are the same on every lane (scalars, k once per
    // wavefront), the lane's part (its column's i) is computed once. Columns
    // 64.. of every row are one more load for the first threads. All loads
    // are issued before the first LDS store waits.
    constexpr int NW = RB / 64, NROW = RS_RY * 4, RPW = NROW / NW, NEXTRA = (RS_RX - 64) * NROW;
    static_assert(NW == 4 && RPW == RS_RY && RS_RX >= 64 && NEXTRA <= RB, "row split of the restriction tile");
    const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int half = F.split ? F.half : 0, rl = F.split ? F.nx >> 1 : F.nx;
    const int il = tile_coord(2 * I0 - 1 + lane, F.nx, F.bt[0], F.bt[1]);
    // the lane's offset within a row for row parity 0 and 1
    const int lo0 = F.split ? (il & 1) * half + (il >> 1) : il;
    const int lo1 = F.split ? ((il & 1) ^ 1) * half + (il >> 1) : il;
    T v[RPW], ve = T(0);
    const int k = tile_coord(2 * K - 1 + w, F.nz, F.bt[4], F.bt[5]);
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int j = tile_coord(2 * J0 - 1 + u, F.ny, F.bt[2], F.bt[3]);
        const int rowoff = (k * F.ny + j) * rl;
        v[u] = rf[rowoff + (((j + k) & 1) ? lo1 : lo0)];
    }
    if (t < NEXTRA) {
        const int r = t / (RS_RX - 64), x = 64 + t % (RS_RX - 64), y = r % RS_RY, z = r / RS_RY;
        const int i = tile_coord(2 * I0 - 1 + x, F.nx, F.bt[0], F.bt[1]);
        const int j = tile_coord(2 * J0 - 1 + y, F.ny, F.bt[2], F.bt[3]);
        const int k = tile_coord(2 * K - 1 + z, F.nz, F.bt[4], F.bt[5]);
        ve = rf[lidx(F, i, j, k)];
    }
#pragma unroll
    for (int u = 0; u < RPW; ++u) tile[(w * RS_RY + u) * RS_RX + lane] = v[u];
    if (t < NEXTRA) tile[(t / (RS_RX - 64)) * RS_RX + 64 + t % (RS_RX - 64)] = ve;
    __syncthreads();
    const int I = I0 + t % RS_CX, J = J0 + t / RS_CX;
    if (I >= Cl.nx || J >= Cl.ny) return;
    const T wt[4] = {T(0.25), T(0.75), T(0.75), T(0.25)};
    const T *base = tile + 2 * (t / RS_CX) * RS_RX + 2 * (t % RS_CX);
    T s = T(0);
    for (int c3 = 0; c3 < 4; ++c3)
        for (int c2 = 0; c2 < 4; ++c2)
            for (int c1 = 0; c1 < 4; ++c1) {
                T w = wt[c3] * wt[c2] * wt[c1];
                s = s + w * base[(c3 * RS_RY + c2) * RS_RX + c1];
            }
    const T bv = T(0.125) * s;
    const int o = lidx(Cl, I, J, K);
    bc[o] = bv;
    if (xc) xc[o] = ((I + J + K) & 1) ? T(0) : (bv + T(0)) * T(Cl.idiag[diag_index(Cl, I, J, K)]);
}
template <typename T>
__global__ __launch_bounds__(RB) void k_restrict(LevelDev F, const T *__restrict__ rf, LevelDev Cl,
                                                 T *__restrict__ bc, T *__restrict__ xc,
                                                 const double *__restrict__ stop) {
    restrict_body<T>(hw_blk(), F, rf, Cl, bc, xc, stop);
}

__device__ __forceinline__ void parents(int f, int nc, int lo, int hi, int &p0, int &p1) {
    int cc = f / 2;
    int nb = (f & 1) ? cc + 1 : cc - 1;
    if (nb < 0) nb = lo != IBMDEM_BC_PERIODIC ? cc : nb + nc;
    else if (nb >= nc) nb = hi != IBMDEM_BC_PERIODIC ? cc : nb - nc;
    p0 = cc; p1 = nb;
}

// x_f += sum (wz*wy*wx) x_c over the 8 parents, weights (3/4 own, 1/4
// neighbour). The block's coarse parents (its x-pairs' coarse columns +-1,
// its rows' coarse rows +-1, the plane's two parent planes) are staged in LDS
// once, so each coarse value is read from memory once per block instead of
// by every fine cell it weighs into; the sum is the same 8 products in the
// same order.
constexpr int PRO_TILE = 1024;   // (bx + 2) * (by / 2 + 2) * 2 <= 528 for the shapes of shape_for
template <typename T>
__device__ __forceinline__ void prolong_body(const Blk &kb, const LevelDev &F, T *__restrict__ xf, const LevelDev &Cl,
                                                const T *__restrict__ xc, const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    __shared__ T tile[PRO_TILE];
    const int bx = kb.dx, by = kb.dy;
    const int TX = bx + 2, TY = by / 2 + 2;
    const int I0 = kb.bx * bx - 1, J0 = (kb.by * by) / 2 - 1;
    const int k = kb.bz;
    int pk[2];
    parents(k, Cl.nz, Cl.bt[4], Cl.bt[5], pk[0], pk[1]);
    const int t = kb.ty * bx + kb.tx;
    // the thread's two fine values are loaded first, so their latency
    // overlaps the staging of the coarse tile
    const int ip = kb.bx * bx + kb.tx;
    const int j = kb.by * by + kb.ty;
    T xo[2] = {T(0), T(0)};
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (j < F.ny && 2 * ip + h < F.nx) xo[h] = xf[lidx(F, 2 * ip + h, j, k)];
    if (bx == 64) {
        // one wavefront per kb.ty: wavefront w stages columns 0..63 of
        // tile rows w, w + by, ... (row r = c TY + y); a row's (J, plane) and
        // its part of the storage index are scalars, the lane's column index
        // is computed once; columns 64 and 65 are one more load for the first
        // threads. All loads are issued before the LDS stores wait.
        const int lane = kb.tx, w = kb.ty, NROW = 2 * TY;
        const int half = Cl.split ? Cl.half : 0, rl = Cl.split ? Cl.nx >> 1 : Cl.nx;
        const int Il = tile_coord(I0 + lane, Cl.nx, Cl.bt[0], Cl.bt[1]);
        const int lo0 = Cl.split ? (Il & 1) * half + (Il >> 1) : Il;
        const int lo1 = Cl.split ? ((Il & 1) ^ 1) * half + (Il >> 1) : Il;
        T v[4], ve = T(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = w + u * by;
            v[u] = T(0);
            if (r < NROW) {
                const int y = r % TY, c = r / TY;
                const int J = tile_coord(J0 + y, Cl.ny, Cl.bt[2], Cl.bt[3]), K = pk[c];
                v[u] = xc[(K * Cl.ny + J) * rl + (((J + K) & 1) ? lo1 : lo0)];
            }
        }
        if (t < 2 * NROW) {
            const int r = t >> 1, x = 64 + (t & 1), y = r % TY, c = r / TY;
            const int I = tile_coord(I0 + x, Cl.nx, Cl.bt[0], Cl.bt[1]);
            const int J = tile_coord(J0 + y, Cl.ny, Cl.bt[2], Cl.bt[3]);
            ve = xc[lidx(Cl, I, J, pk[c])];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = w + u * by;
            if (r < NROW) tile[r * TX + lane] = v[u];
        }
        if (t < 2 * NROW) tile[(t >> 1) * TX + 64 + (t & 1)] = ve;
    } else
    // up to four loads of a thread in flight before the LDS stores wait
    for (int q0 = t; q0 < 2 * TY * TX; q0 += 4 * bx * by) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * bx * by;
            if (q < 2 * TY * TX) {
                const int x = q % TX, y = (q / TX) % TY, c = q / (TX * TY);
                const int I = tile_coord(I0 + x, Cl.nx, Cl.bt[0], Cl.bt[1]);
                const int J = tile_coord(J0 + y, Cl.ny, Cl.bt[2], Cl.bt[3]);
                v[u] = xc[lidx(Cl, I, J, pk[c])];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * bx * by;
            if (q < 2 * TY * TX) tile[q] = v[u];
        }
    }
    __syncthreads();
    if (j >= F.ny) return;
    const T wt[2] = {T(0.75), T(0.25)};
    // tile rows of the own and the neighbour coarse row
    const int ty[2] = {j / 2 - J0, j / 2 + ((j & 1) ? 1 : -1) - J0};
    // the thread's x-pair (2p, 2p+1), one cell of each colour
    for (int h = 0; h < 2; ++h) {
        const int i = 2 * ip + h;
        if (i >= F.nx) return;
        const int tx[2] = {i / 2 - I0, i / 2 + ((i & 1) ? 1 : -1) - I0};
        T s = T(0);
        for (int c3 = 0; c3 < 2; ++c3)
            for (int c2 = 0; c2 < 2; ++c2)
                for (int c1 = 0; c1 < 2; ++c1) {
                    T w = wt[c3] * wt[c2] * wt[c1];
                    s = s + w * tile[(c3 * TY + ty[c2]) * TX + tx[c1]];
                }
        xf[lidx(F, i, j, k)] = xo[h] + s;
    }
}
template <typename T>
__global__ __launch_bounds__(RB) void k_prolong(LevelDev F, T *__restrict__ xf, LevelDev Cl,
                                                const T *__restrict__ xc, const double *__restrict__ stop) {
    prolong_body<T>(hw_blk(), F, xf, Cl, xc, stop);
}

// The same prolongation with each thread on two adjacent x-pairs (pc, pc +
// 1) of its row, on split fine levels whose rows hold an even number of
// pairs: the thread's four fine values are two two-value loads (and
// stores) of the colour halves. The block's coarse parents (its pairs'
// coarse columns +-1, rows +-1, the plane's two parent planes) are staged in
// LDS as prolong_body stages them; each fine value gets the same 8 products
// in the same order.
constexpr int PRO2_TILE = 1040;   // 2 (by / 2 + 2) (2 bx + 2) for the shapes of shape_for
template <typename T>
__device__ __forceinline__ void prolong2_body(const Blk &kb, const LevelDev &F, T *__restrict__ xf,
                                              const LevelDev &Cl, const T *__restrict__ xc,
                                              const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    __shared__ T tile[PRO2_TILE];
    const int bx = kb.dx, by = kb.dy;
    const int TX = 2 * bx + 2, TY = by / 2 + 2;
    const int I0 = 2 * kb.bx * bx - 1, J0 = (kb.by * by) / 2 - 1;
    const int k = kb.bz;
    int pk[2];
    parents(k, Cl.nz, Cl.bt[4], Cl.bt[5], pk[0], pk[1]);
    const int t = kb.ty * bx + kb.tx;
    const int rl = F.nx >> 1;
    const int p0 = 2 * (kb.bx * bx + kb.tx), pc = min(p0, rl - 2);
    const int j = kb.by * by + kb.ty, jc = min(j, F.ny - 1);
    const int ca = (jc + k) & 1;
    const long pos = (long)(k * F.ny + jc) * rl + pc;
    // fine values first (even cells of the two pairs in half ca, odd cells in
    // half 1 - ca), so their latency overlaps the staging
    T xo[2][2];
    ld2(xf + (long)ca * F.half + pos, xo[0][0], xo[0][1]);
    ld2(xf + (long)(1 - ca) * F.half + pos, xo[1][0], xo[1][1]);
    // The tile's rows r = c TY + y (coarse row J0 + y of parent plane c).
    // Rows of 64 or more columns: wavefront w stages rows w, w + NW, ...; a
    // row's (J, K) are the same on every lane (scalars), and lane l takes
    // columns l, l + 64, l + 128 (< TX), whose coarse indices it maps once.
    // Column x is stored at (x & 1) * TXH + x / 2 of its row: the summing
    // loop's lanes read columns two apart, which land in consecutive words.
    const int TXH = bx + 1, NROW = 2 * TY, NW = bx * by / 64;
    if (TX >= 64) {
        const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
        int Ic[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) Ic[u] = tile_coord(I0 + lane + 64 * u, Cl.nx, Cl.bt[0], Cl.bt[1]);
        for (int r = w; r < NROW; r += NW) {
            const int y = r % TY, c = r / TY;
            const int J = tile_coord(J0 + y, Cl.ny, Cl.bt[2], Cl.bt[3]), K = pk[c];
            T v[3];
#pragma unroll
            for (int u = 0; u < 3; ++u)
                if (lane + 64 * u < TX) v[u] = xc[lidx(Cl, Ic[u], J, K)];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const int x = lane + 64 * u;
                if (x < TX) tile[r * TX + (x & 1) * TXH + (x >> 1)] = v[u];
            }
        }
    } else {
        // short rows (blocks narrower than 32 pair-pairs): the tile's
        // elements dealt to the threads in order, up to four loads of a
        // thread in flight before the LDS stores wait
        const int NT = NROW * TX;
        for (int q0 = t; q0 < NT; q0 += 4 * bx * by) {
            T v[4];
            int dst[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u * bx * by;
                dst[u] = -1;
                if (q < NT) {
                    const int r = q / TX, x = q - r * TX, y = r % TY, c = r / TY;
                    const int I = tile_coord(I0 + x, Cl.nx, Cl.bt[0], Cl.bt[1]);
                    const int J = tile_coord(J0 + y, Cl.ny, Cl.bt[2], Cl.bt[3]);
                    v[u] = xc[lidx(Cl, I, J, pk[c])];
                    dst[u] = r * TX + (x & 1) * TXH + (x >> 1);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (dst[u] >= 0) tile[dst[u]] = v[u];
        }
    }
    __syncthreads();
    if (j >= F.ny || p0 >= rl) return;
    const T wt[2] = {T(0.75), T(0.25)};
    const int ty[2] = {j / 2 - J0, j / 2 + ((j & 1) ? 1 : -1) - J0};
    T out[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = 2 * (pc + c) + h;
            const int tx[2] = {i / 2 - I0, i / 2 + ((i & 1) ? 1 : -1) - I0};
            const int sx[2] = {(tx[0] & 1) * TXH + (tx[0] >> 1), (tx[1] & 1) * TXH + (tx[1] >> 1)};
            T s = T(0);
            for (int c3 = 0; c3 < 2; ++c3)
                for (int c2 = 0; c2 < 2; ++c2)
                    for (int c1 = 0; c1 < 2; ++c1) {
                        T wgt = wt[c3] * wt[c2] * wt[c1];
                        s = s + wgt * tile[(c3 * TY + ty[c2]) * TX + sx[c1]];
                    }
            out[h][c] = xo[h][c] + s;
        }
    st2(xf + (long)ca * F.half + pos, out[0][0], out[0][1]);
    st2(xf + (long)(1 - ca) * F.half + pos, out[1][0], out[1][1]);
}
template <typename T>
__global__ __launch_bounds__(RB) void k_prolong2(LevelDev F, T *__restrict__ xf, LevelDev Cl,
                                                 const T *__restrict__ xc, const double *__restrict__ stop) {
    prolong2_body<T>(hw_blk(), F, xf, Cl, xc, stop);
}

// The prolongation's launch onto fine level d (one row per thread): two
// pairs per thread where sweep_shape takes two, else one; ok = false when
// the block's coarse tile would not fit the kernel's LDS tile.
struct ProlongShape {
    dim3 grid, block;
    bool x2, ok;
};
ProlongShape prolong_shape(const LevelDev &d) {
    const bool x2 = d.split && (d.nx >> 1) % 2 == 0;
    const LaunchShape sh = shape_for(x2 ? d.nx / 4 : div_up(d.nx, 2), d.ny, d.nz);
    const unsigned bx = sh.block.x, by = sh.block.y;
    const unsigned need = x2 ? (2 * bx + 2) * (by / 2 + 2) * 2 : (bx + 2) * (by / 2 + 2) * 2;
    return {sh.grid, sh.block, x2, by % 2 == 0 && need <= (unsigned)(x2 ? PRO2_TILE : PRO_TILE)};
}

// coarsest level, one work-group. Small levels (<= COARSE_LDS cells, every
// level the grids here coarsen to) live in LDS with a precomputed neighbour
// table whose missing (wall) neighbours point at a zero slot, so the sweeps
// evaluate exactly nbsum's expression without index arithmetic; larger
// coarsest levels run the same sweeps in global memory.
constexpr int COARSE_LDS = 512;

// The symmetric red-black sweeps of a coarsest level of n <= 64 CQ cells on
// one wavefront (lane l owns the cells q = l + 64 t): the lane's neighbour
// indices, 1/diagonal, right-hand side and colour are read from the tables
// once into registers, so a half sweep is one round of LDS loads of x (the
// zero slot x[n] for a missing neighbour). Cells of one colour do not read
// each other, so the lane's order over them does not matter; each cell's
// expression is the table sweep's.
template <int CQ, typename T>
__device__ __forceinline__ void coarse_reg_sweeps(T *x, const T *b, const short (*nb)[COARSE_LDS], const T *id,
                                                  const unsigned char *col, int n, int sweeps, T ihx2, T ihy2,
                                                  T ihz2) {
    const int lane = threadIdx.x & 63;
    int nbr[CQ][6], cr[CQ];
    T br[CQ], idr[CQ];
#pragma unroll
    for (int t = 0; t < CQ; ++t) {
        const int q = lane + 64 * t;
        cr[t] = -1;
        br[t] = T(0);
        idr[t] = T(0);
        for (int d = 0; d < 6; ++d) nbr[t][d] = n;
        if (q < n) {
            for (int d = 0; d < 6; ++d) nbr[t][d] = nb[d][q];
            br[t] = b[q];
            idr[t] = id[q];
            cr[t] = col[q];
        }
    }
    auto half = [&](int color) {
#pragma unroll
        for (int t = 0; t < CQ; ++t)
            if (cr[t] == color) {
                const T sx = (x[nbr[t][0]] + x[nbr[t][1]]) * ihx2;
                const T sy = (x[nbr[t][2]] + x[nbr[t][3]]) * ihy2;
                const T sz = (x[nbr[t][4]] + x[nbr[t][5]]) * ihz2;
                x[lane + 64 * t] = (br[t] + (sx + sy + sz)) * idr[t];
            }
        wave_barrier_lds();
    };
    for (int s = 0; s < sweeps; ++s) { half(0); half(1); }
    for (int s = 0; s < sweeps; ++s) { half(1); half(0); }
}
// the register-held sweeps for levels of up to 256 cells; false: larger
template <typename T>
__device__ __forceinline__ bool coarse_sweeps_small(T *x, const T *b, const short (*nb)[COARSE_LDS], const T *id,
                                                    const unsigned char *col, int n, int sweeps, T ihx2, T ihy2,
                                                    T ihz2) {
    if (n <= 64) coarse_reg_sweeps<1, T>(x, b, nb, id, col, n, sweeps, ihx2, ihy2, ihz2);
    else if (n <= 128) coarse_reg_sweeps<2, T>(x, b, nb, id, col, n, sweeps, ihx2, ihy2, ihz2);
    else if (n <= 256) coarse_reg_sweeps<4, T>(x, b, nb, id, col, n, sweeps, ihx2, ihy2, ihz2);
    else return false;
    return true;
}

template <typename T>
__device__ __forceinline__ void coarse_lds_body(const Blk &kb, const LevelDev &L, T *__restrict__ xg, const T *__restrict__ bg,
                                                   int sweeps, const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    __shared__ T x[COARSE_LDS + 1], b[COARSE_LDS], id[COARSE_LDS], t[COARSE_LDS];
    __shared__ short nb[6][COARSE_LDS];
    __shared__ unsigned char col[COARSE_LDS];
    const int n = L.nx * L.ny * L.nz;
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    for (int q = kb.tx; q < n; q += 64) {
        const int i = q % L.nx, j = (q / L.nx) % L.ny, k = q / (L.nx * L.ny);
        const int a[3] = {i, j, k}, m[3] = {L.nx, L.ny, L.nz};
        for (int d = 0; d < 3; ++d)
            for (int dir = 0; dir < 2; ++dir) {
                int p[3] = {a[0], a[1], a[2]};
                int v = p[d] + (dir ? 1 : -1);
                int idx;
                if (v < 0) idx = L.bt[2 * d] != IBMDEM_BC_PERIODIC ? -1 : v + m[d];
                else if (v >= m[d]) idx = L.bt[2 * d + 1] != IBMDEM_BC_PERIODIC ? -1 : v - m[d];
                else idx = v;
                p[d] = idx;
                nb[2 * d + dir][q] = (short)(idx < 0 ? n : (p[2] * L.ny + p[1]) * L.nx + p[0]);
            }
        id[q] = T(L.idiag[diag_index(L, i, j, k)]);
        col[q] = (unsigned char)((i + j + k) & 1);
        b[q] = bg[lidx(L, i, j, k)];
        x[q] = T(0);
    }
    if (kb.tx == 0) x[n] = T(0);
    __syncthreads();
    auto upd = [&](int q) -> T {
        const T sx = (x[nb[0][q]] + x[nb[1][q]]) * ihx2;
        const T sy = (x[nb[2][q]] + x[nb[3][q]]) * ihy2;
        const T sz = (x[nb[4][q]] + x[nb[5][q]]) * ihz2;
        return (b[q] + (sx + sy + sz)) * id[q];
    };
    auto half = [&](int color) {
        for (int q = kb.tx; q < n; q += 64)
            if (col[q] == color) x[q] = upd(q);
        __syncthreads();
    };
    if (L.rb_valid) {
        if (!coarse_sweeps_small<T>(x, b, nb, id, col, n, sweeps, ihx2, ihy2, ihz2)) {
            for (int s = 0; s < sweeps; ++s) { half(0); half(1); }
            for (int s = 0; s < sweeps; ++s) { half(1); half(0); }
        }
    } else {
        const T om = T(0.8);
        for (int s = 0; s < 2 * sweeps; ++s) {
            for (int q = kb.tx; q < n; q += 64) t[q] = x[q] + om * (upd(q) - x[q]);
            __syncthreads();
            for (int q = kb.tx; q < n; q += 64) x[q] = t[q];
            __syncthreads();
        }
    }
    for (int q = kb.tx; q < n; q += 64) xg[lidx(L, q % L.nx, (q / L.nx) % L.ny, q / (L.nx * L.ny))] = x[q];
}
template <typename T>
__global__ __launch_bounds__(64) void k_coarse_lds(LevelDev L, T *__restrict__ xg, const T *__restrict__ bg,
                                                   int sweeps, const double *__restrict__ stop) {
    coarse_lds_body<T>(hw_blk(), L, xg, bg, sweeps, stop);
}

template <typename T>
__device__ __forceinline__ void coarse_body(const Blk &kb, const LevelDev &L, T *__restrict__ x, const T *__restrict__ b,
                                                     T *__restrict__ tmp, int sweeps, const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    const long n = (long)L.nx * L.ny * L.nz;
    for (long q = kb.tx; q < n; q += COARSE_T) x[q] = T(0);
    __syncthreads();
    auto half = [&](int color) {
        for (long q = kb.tx; q < n; q += COARSE_T) {
            int i, j, k;
            lcell(L, (int)q, i, j, k);
            if (((i + j + k) & 1) != color) continue;
            x[q] = (b[q] + nbsum(x, L, i, j, k)) * T(L.idiag[diag_index(L, i, j, k)]);
        }
        __syncthreads();
    };
    if (L.rb_valid) {
        for (int s = 0; s < sweeps; ++s) { half(0); half(1); }
        for (int s = 0; s < sweeps; ++s) { half(1); half(0); }
    } else {
        const T om = T(0.8);
        for (int s = 0; s < 2 * sweeps; ++s) {
            for (long q = kb.tx; q < n; q += COARSE_T) {
                int i, j, k;
                lcell(L, (int)q, i, j, k);
                T xn = (b[q] + nbsum(x, L, i, j, k)) * T(L.idiag[diag_index(L, i, j, k)]);
                tmp[q] = x[q] + om * (xn - x[q]);
            }
            __syncthreads();
            for (long q = kb.tx; q < n; q += COARSE_T) x[q] = tmp[q];
            __syncthreads();
        }
    }
}
template <typename T>
__global__ __launch_bounds__(COARSE_T) void k_coarse(LevelDev L, T *__restrict__ x, const T *__restrict__ b,
                                                     T *__restrict__ tmp, int sweeps, const double *__restrict__ stop) {
    coarse_body<T>(hw_blk(), L, x, b, tmp, sweeps, stop);
}

// ---------------------------------------------------------------------------
// The small levels of the hierarchy as one work-group: from level `first`
// down to the coarsest, the whole V-cycle (smoothing, residual, restriction,
// coarse sweeps, prolongation) runs in LDS with a barrier between phases, in
// place of a few dozen launches that each cost more than their work. The
// coarsest level sweeps on one wavefront from neighbour tables, as
// k_coarse_lds does. Every cell's arithmetic is the multi-launch path's, so
// the result is the same bit for bit; b of level `first` comes from global
// memory and its x goes back there for the prolongation to the level above.
// ---------------------------------------------------------------------------
constexpr int TAIL_T = 1024;
constexpr int TAIL_BYTES = 112 * 1024;   // x (one extra zero slot on the coarsest level), b, r of the levels
constexpr int TAIL_MAXLEV = 6;           // levels of one tail (their descriptors are copied to LDS)
static_assert(sizeof(LevelDev) % 4 == 0, "level descriptors copied by words");
// (the tail's levels hold at most TAIL_BYTES / 12 cells, so its index
// divisions go through udiv_small)
static_assert(TAIL_BYTES / 12 < (1 << 14), "tail levels small enough for udiv_small");

template <typename T>
__device__ __forceinline__ void tail_half(const LevelDev &L, T *x, const T *b, int color, bool zero) {
    const int n = L.nx * L.ny * L.nz;
    if (L.split) {   // by x-pairs, as k_rb_half
        const int rl = L.nx >> 1;
        T *own = x + color * L.half, *oth = x + (1 - color) * L.half;
        for (int p = threadIdx.x; p < L.half; p += TAIL_T) {
            const int row = udiv_small(p, rl, L.m_row), kr = udiv_small(row, L.ny, L.m_ny);
            const PairGeo g = pair_geo(L, p - row * rl, row - kr * L.ny, kr);
            const int h = g.ca == color ? 0 : 1;
            const T id = T(L.idiag[(h ? g.cx1 : g.cx0) * 16 + g.cyz]);
            const T bo = T(b[color * L.half + p]);
            if (zero) {
                oth[p] = T(0);
                own[p] = (bo + T(0)) * id;
            } else {
                own[p] = (bo + nbsum_pair(oth, g, h, L)) * id;
            }
        }
        __syncthreads();
        return;
    }
    for (int q = threadIdx.x; q < n; q += TAIL_T) {
        int i, j, k;
        lcell_small(L, q, i, j, k);
        if (((i + j + k) & 1) != color) {
            if (zero) x[q] = T(0);
            continue;
        }
        const T id = T(L.idiag[diag_index(L, i, j, k)]);
        x[q] = zero ? (T(b[q]) + T(0)) * id : (T(b[q]) + nbsum(x, L, i, j, k)) * id;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void tail_residual(const LevelDev &L, const T *x, const T *b, T *r) {
    const int n = L.nx * L.ny * L.nz;
    if (L.split) {   // by x-pairs, as the residual of k_rb_half
        const int rl = L.nx >> 1;
        for (int p = threadIdx.x; p < L.half; p += TAIL_T) {
            const int row = udiv_small(p, rl, L.m_row), kr = udiv_small(row, L.ny, L.m_ny);
            const PairGeo g = pair_geo(L, p - row * rl, row - kr * L.ny, kr);
            for (int h = 0; h < 2; ++h) {
                const int own = (h ? 1 - g.ca : g.ca) * L.half, oth = (h ? g.ca : 1 - g.ca) * L.half;
                const T dg = T(L.diag[(h ? g.cx1 : g.cx0) * 16 + g.cyz]);
                r[own + p] = T(b[own + p]) - (dg * x[own + p] - nbsum_pair(x + oth, g, h, L));
            }
        }
        __syncthreads();
        return;
    }
    for (int q = threadIdx.x; q < n; q += TAIL_T) {
        int i, j, k;
        lcell_small(L, q, i, j, k);
        const T dg = T(L.diag[diag_index(L, i, j, k)]);
        r[q] = T(b[q]) - (dg * x[q] - nbsum(x, L, i, j, k));
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void tail_restrict(const LevelDev &F, const T *rf, const LevelDev &Cl, T *bc, T *xc) {
    const T wt[4] = {T(0.25), T(0.75), T(0.75), T(0.25)};
    const int n = Cl.nx * Cl.ny * Cl.nz;
    const int rl = F.split ? F.nx >> 1 : F.nx, hs = F.split ? F.half : 0;
    for (int q = threadIdx.x; q < n; q += TAIL_T) {
        int I, J, K;
        lcell_small(Cl, q, I, J, K);
        // lidx(F, i, j, k) = ((i & 1) ^ ((j + k) & 1)) * half + row(j, k) + (i >> 1)
        // when split, row(j, k) + i otherwise: its parts per tap coordinate
        int xo[4], xp[4], ro[16], rp[16];
        for (int a = 0; a < 4; ++a) {
            const int fi = tile_coord(2 * I - 1 + a, F.nx, F.bt[0], F.bt[1]);
            xo[a] = F.split ? fi >> 1 : fi;
            xp[a] = fi & 1;
        }
        for (int c3 = 0; c3 < 4; ++c3) {
            const int fk = tile_coord(2 * K - 1 + c3, F.nz, F.bt[4], F.bt[5]);
            for (int c2 = 0; c2 < 4; ++c2) {
                const int fj = tile_coord(2 * J - 1 + c2, F.ny, F.bt[2], F.bt[3]);
                ro[4 * c3 + c2] = (fk * F.ny + fj) * rl;
                rp[4 * c3 + c2] = (fj + fk) & 1;
            }
        }
        T s = T(0);
        for (int c3 = 0; c3 < 4; ++c3)
            for (int c2 = 0; c2 < 4; ++c2)
                for (int c1 = 0; c1 < 4; ++c1) {
                    T w = wt[c3] * wt[c2] * wt[c1];
                    const int r = 4 * c3 + c2;
                    s = s + w * rf[(xp[c1] ^ rp[r]) * hs + ro[r] + xo[c1]];
                }
        const T v = T(0.125) * s;
        bc[q] = v;
        // the coarse level's first sweep, as k_restrict does it with xc
        if (xc) xc[q] = ((I + J + K) & 1) ? T(0) : (v + T(0)) * T(Cl.idiag[diag_index(Cl, I, J, K)]);
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void tail_prolong(const LevelDev &F, T *xf, const LevelDev &Cl, const T *xc) {
    const T wt[2] = {T(0.75), T(0.25)};
    const int n = F.nx * F.ny * F.nz;
    const int rl = Cl.split ? Cl.nx >> 1 : Cl.nx, hs = Cl.split ? Cl.half : 0;
    for (int q = threadIdx.x; q < n; q += TAIL_T) {
        int i, j, k;
        lcell_small(F, q, i, j, k);
        int pi[2], pj[2], pk[2];
        parents(i, Cl.nx, Cl.bt[0], Cl.bt[1], pi[0], pi[1]);
        parents(j, Cl.ny, Cl.bt[2], Cl.bt[3], pj[0], pj[1]);
        parents(k, Cl.nz, Cl.bt[4], Cl.bt[5], pk[0], pk[1]);
        // lidx(Cl, .) by parts, as in tail_restrict
        int xo[2], xp[2], ro[4], rp[4];
        for (int a = 0; a < 2; ++a) {
            xo[a] = Cl.split ? pi[a] >> 1 : pi[a];
            xp[a] = pi[a] & 1;
        }
        for (int c3 = 0; c3 < 2; ++c3)
            for (int c2 = 0; c2 < 2; ++c2) {
                ro[2 * c3 + c2] = (pk[c3] * Cl.ny + pj[c2]) * rl;
                rp[2 * c3 + c2] = (pj[c2] + pk[c3]) & 1;
            }
        T s = T(0);
        for (int c3 = 0; c3 < 2; ++c3)
            for (int c2 = 0; c2 < 2; ++c2)
                for (int c1 = 0; c1 < 2; ++c1) {
                    T w = wt[c3] * wt[c2] * wt[c1];
                    const int r = 2 * c3 + c2;
                    s = s + w * xc[(xp[c1] ^ rp[r]) * hs + ro[r] + xo[c1]];
                }
        xf[q] = xf[q] + s;
    }
    __syncthreads();
}

// The coarsest level (n <= COARSE_LDS cells, consistent red-black colouring)
// on wavefront 0: neighbour tables whose missing (wall) neighbours point at
// the zero slot x[n], then the symmetric sweeps of k_coarse_lds.
template <typename T>
__device__ __forceinline__ void tail_coarse(const LevelDev &L, T *x, const T *b, int sweeps,
                                            short (*nb)[COARSE_LDS], T *id, unsigned char *col) {
    const int n = L.nx * L.ny * L.nz;
    const int m[3] = {L.nx, L.ny, L.nz};
    for (int q = threadIdx.x; q < n; q += 64) {
        int a[3];
        lcell_small(L, q, a[0], a[1], a[2]);
        for (int d = 0; d < 3; ++d)
            for (int dir = 0; dir < 2; ++dir) {
                int p[3] = {a[0], a[1], a[2]};
                const int v = p[d] + (dir ? 1 : -1);
                int idx;
                if (v < 0) idx = L.bt[2 * d] != IBMDEM_BC_PERIODIC ? -1 : v + m[d];
                else if (v >= m[d]) idx = L.bt[2 * d + 1] != IBMDEM_BC_PERIODIC ? -1 : v - m[d];
                else idx = v;
                p[d] = idx;
                nb[2 * d + dir][q] = (short)(idx < 0 ? n : lidx(L, p[0], p[1], p[2]));
            }
        id[q] = T(L.idiag[diag_index(L, a[0], a[1], a[2])]);
        col[q] = (unsigned char)((a[0] + a[1] + a[2]) & 1);
        x[q] = T(0);
    }
    if (threadIdx.x == 0) x[n] = T(0);
    wave_barrier_lds();
    const T ihx2 = T(L.ihx2), ihy2 = T(L.ihy2), ihz2 = T(L.ihz2);
    auto half = [&](int color) {
        for (int q = threadIdx.x; q < n; q += 64)
            if (col[q] == color) {
                const T sx = (x[nb[0][q]] + x[nb[1][q]]) * ihx2;
                const T sy = (x[nb[2][q]] + x[nb[3][q]]) * ihy2;
                const T sz = (x[nb[4][q]] + x[nb[5][q]]) * ihz2;
                x[q] = (b[q] + (sx + sy + sz)) * id[q];
            }
        wave_barrier_lds();
    };
    if (coarse_sweeps_small<T>(x, b, nb, id, col, n, sweeps, ihx2, ihy2, ihz2)) return;
    for (int s = 0; s < sweeps; ++s) { half(0); half(1); }
    for (int s = 0; s < sweeps; ++s) { half(1); half(0); }
}

template <typename T>
__device__ __forceinline__ void vcycle_tail_body(const Blk &kb, const LevelDev *__restrict__ lev, int first, int last,
                                                        const T *__restrict__ bg, T *__restrict__ xg, int pre,
                                                        int post, int coarse_sweeps,
                                                        const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    __shared__ T mem[TAIL_BYTES / sizeof(T)];
    __shared__ short s_nb[6][COARSE_LDS];
    __shared__ T s_id[COARSE_LDS];
    __shared__ unsigned char s_col[COARSE_LDS];
    __shared__ int s_off[IBMDEM_MAX_LEVELS + 1];   // level l: x at off, then b, then r
    // the levels' descriptors (sizes, boundary types, diagonal tables) in
    // LDS: every phase reads them, and a level's first phase would otherwise
    // wait on its scalar and table loads from global memory
    __shared__ LevelDev s_lev[TAIL_MAXLEV];
    {
        const int nw = (last - first + 1) * (int)(sizeof(LevelDev) / 4);
        const int *src = (const int *)(lev + first);
        int *dst = (int *)s_lev;
        for (int w = kb.tx; w < nw; w += TAIL_T) dst[w] = src[w];
    }
    if (kb.tx == 0) {
        int off = 0;
        for (int l = first; l <= last; ++l) {
            s_off[l] = off;
            off += 3 * lev[l].nx * lev[l].ny * lev[l].nz + (l == last ? 1 : 0);
        }
    }
    __syncthreads();
    auto LV = [&](int l) -> const LevelDev & { return s_lev[l - first]; };
    auto cells = [&](int l) { return LV(l).nx * LV(l).ny * LV(l).nz; };
    auto X = [&](int l) { return mem + s_off[l]; };
    auto B = [&](int l) { return mem + s_off[l] + cells(l) + (l == last ? 1 : 0); };
    auto R = [&](int l) { return B(l) + cells(l); };
    for (int q = kb.tx; q < cells(first); q += TAIL_T) B(first)[q] = bg[q];
    __syncthreads();
    // below the first level, the restriction into a level also does its
    // first sweep (the coarsest level sets its own x)
    for (int l = first; l < last; ++l) {
        for (int q = 0; q < pre; ++q) {
            if (q > 0 || l == first) tail_half(LV(l), X(l), B(l), 0, q == 0);
            tail_half(LV(l), X(l), B(l), 1, false);
        }
        tail_residual(LV(l), X(l), B(l), R(l));
        tail_restrict(LV(l), R(l), LV(l + 1), B(l + 1), l + 1 < last ? X(l + 1) : (T *)nullptr);
    }
    if (kb.tx < 64) tail_coarse(LV(last), X(last), B(last), coarse_sweeps, s_nb, s_id, s_col);
    __syncthreads();
    for (int l = last - 1; l >= first; --l) {
        tail_prolong(LV(l), X(l), LV(l + 1), X(l + 1));
        for (int q = 0; q < post; ++q) {
            tail_half(LV(l), X(l), B(l), 1, false);
            tail_half(LV(l), X(l), B(l), 0, false);
        }
    }
    for (int q = kb.tx; q < cells(first); q += TAIL_T) xg[q] = X(first)[q];
}
template <typename T>
__global__ __launch_bounds__(TAIL_T) void k_vcycle_tail(const LevelDev *__restrict__ lev, int first, int last,
                                                        const T *__restrict__ bg, T *__restrict__ xg, int pre,
                                                        int post, int coarse_sweeps,
                                                        const double *__restrict__ stop) {
    vcycle_tail_body<T>(hw_blk(), lev, first, last, bg, xg, pre, post, coarse_sweeps, stop);
}

// ---------------------------------------------------------------------------
// CG kernels (FP64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_reduce(double s, double *red) {
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = RED_T / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    return red[0];
}

// Ap = A p and one partial sum of p.Ap per block. A thread owns an x-pair
// (both colours), a block 64 x 4 pairs of one plane: a block's plane
// neighbours are read by the block
// gridDim.x * gridDim.y positions away, a multiple of the XCD count for the
// level-0 shapes, so they are served by the same L2.
__device__ __forceinline__ void matvec_dot_body(const Blk &kb, const LevelDev &L, const double *__restrict__ p,
                                                   double *__restrict__ Ap, double *__restrict__ partial,
                                                   const double *__restrict__ scal) {
    if (scal[SC_CONV] != 0.0) return;
    __shared__ double red[RB];
    // a thread takes the x-pair ip of two adjacent rows j0, j0 + 1 (blocks as
    // the half sweep); when both rows exist, a cell's north neighbour in row
    // j0 is row j0 + 1's value of the same cell of its pair, and a cell's
    // south neighbour in row j0 + 1 is row j0's: those loads are shared
    const int ip = kb.bx * kb.dx + kb.tx;
    const int j0 = (kb.by * kb.dy + kb.ty) * 2, k = kb.bz;
    double s = 0.0;
    if (L.split) {
        const int ipc = min(ip, (L.nx >> 1) - 1);
        const double dg_int = L.diag[DIAG_INTERIOR];
        const bool adj = j0 + 1 < L.ny;
        PairGeo g[2];
        bool ok[2];
        long oo[2][2];
        double pv[2][2], dgv[2][2], a[2][2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            ok[r] = j0 + r < L.ny && ip < (L.nx >> 1);
            g[r] = pair_geo(L, ipc, min(j0 + r, L.ny - 1), k);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                oo[r][h] = (long)(h ? 1 - g[r].ca : g[r].ca) * L.half + g[r].pos;
                const int di = (h ? g[r].cx1 : g[r].cx0) * 16 + g[r].cyz;
                double dg;
                if (di == DIAG_INTERIOR) dg = dg_int;
                else dg = L.diag[di];
                dgv[r][h] = dg;
                pv[r][h] = p[oo[r][h]];
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const PairGeo &G = g[r];
                const double *q = p + (long)(h ? G.ca : 1 - G.ca) * L.half + G.pos;
                const double xo = q[h == 0 ? (G.hw ? G.dw : 0) : (G.he ? G.de : 0)];
                const double zb = q[G.hb ? G.db : 0], zt = q[G.ht ? G.dt : 0];
                double ys, yn;
                if (r == 0) {
                    ys = q[G.hs ? G.ds : 0];
                    if (adj) yn = pv[1][h];
                    else yn = q[G.hn ? G.dn : 0];
                } else {
                    yn = q[G.hn ? G.dn : 0];
                    if (adj) ys = pv[0][h];
                    else ys = q[G.hs ? G.ds : 0];
                }
                a[r][h] = dgv[r][h] * pv[r][h] - nbsum_pair_vals(G, h, pv[r][1 - h], xo, ys, yn, zb, zt, L);
            }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (!ok[r]) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                Ap[oo[r][h]] = a[r][h];
                s += pv[r][h] * a[r][h];
            }
        }
    } else {
        for (int r = 0; r < 2; ++r) {
            const int j = j0 + r;
            if (j >= L.ny) break;
            for (int h = 0; h < 2; ++h) {
                const int i = 2 * ip + h;
                if (i >= L.nx) break;
                const int o = lidx(L, i, j, k);
                const double av = L.diag[diag_index(L, i, j, k)] * p[o] - nbsum(p, L, i, j, k);
                Ap[o] = av;
                s += p[o] * av;
            }
        }
    }
    const int nt = kb.dx * kb.dy, t = kb.ty * kb.dx + kb.tx;
    red[t] = s;
    __syncthreads();
    for (int st = nt / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) partial[((long)kb.bz * kb.gy + kb.by) * kb.gx + kb.bx] = red[0];
}
__global__ __launch_bounds__(RB) void k_matvec_dot(LevelDev L, const double *__restrict__ p,
                                                   double *__restrict__ Ap, double *__restrict__ partial,
                                                   const double *__restrict__ scal) {
    matvec_dot_body(hw_blk(), L, p, Ap, partial, scal);
}

// The same product with each thread on two adjacent x-pairs (pc, pc + 1,
// pc even) of its two rows, on split levels whose rows hold an even number
// of pairs (sweep_shape's condition and launch shape): both cells of the two
// pairs come from two-value loads of the two colour halves, and a cell's
// x neighbours are mostly its pairs' other cells. Each Ap value is
// matvec_dot_body's (the same diagonal product less nbsum_pair_vals' sum);
// the partial sums follow this kernel's blocks.
__device__ __forceinline__ void matvec2_dot_body(const Blk &kb, const LevelDev &L, const double *__restrict__ p,
                                                 double *__restrict__ Ap, double *__restrict__ partial,
                                                 const double *__restrict__ scal) {
    if (scal[SC_CONV] != 0.0) return;
    __shared__ double red[RB];
    const double ihx2 = L.ihx2, ihy2 = L.ihy2, ihz2 = L.ihz2;
    const double dg_int = L.diag[DIAG_INTERIOR];
    const int rl = L.nx >> 1, pl = rl * L.ny;
    const int p0 = 2 * (kb.bx * kb.dx + kb.tx), k = kb.bz;
    const int j0 = (kb.by * kb.dy + kb.ty) * 2;
    const bool px = L.bt[0] == IBMDEM_BC_PERIODIC, py = L.bt[2] == IBMDEM_BC_PERIODIC,
               pz = L.bt[4] == IBMDEM_BC_PERIODIC;
    const int pc = min(p0, rl - 2);
    const bool hw = pc > 0 || px, he = pc + 2 < rl || px;
    const int dw = pc > 0 ? -1 : rl - 1, de = pc + 2 < rl ? 2 : 2 - rl;
    const bool hb = k > 0 || pz, ht = k < L.nz - 1 || pz;
    const int db = k > 0 ? -pl : (L.nz - 1) * pl, dt = k < L.nz - 1 ? pl : -(L.nz - 1) * pl;
    const int cz = diag_count(k, L.nz, L.bt[4], L.bt[5]);
    const bool adj = j0 + 1 < L.ny;
    // per row: v[r][hh][c] = the value of cell h = hh of pair pc + c (hh 0:
    // the even cell, in half ca; 1: the odd cell, in half 1 - ca)
    bool ok[2], hs[2], hn[2];
    int pos[2], ca[2], cyz[2];
    double v[2][2][2], zb[2][2][2], zt[2][2][2], xw0[2], xe1[2], ys[2][2][2], yn[2][2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int j = min(j0 + r, L.ny - 1);
        ok[r] = j0 + r < L.ny && p0 < rl;
        pos[r] = (k * L.ny + j) * rl + pc;
        ca[r] = (j + k) & 1;
        hs[r] = j > 0 || py;
        hn[r] = j < L.ny - 1 || py;
        cyz[r] = diag_count(j, L.ny, L.bt[2], L.bt[3]) * 4 + cz;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const double *q = p + (long)(hh ? 1 - ca[r] : ca[r]) * L.half + pos[r];
            ld2(q, v[r][hh][0], v[r][hh][1]);
            ld2(q + (hb ? db : 0), zb[r][hh][0], zb[r][hh][1]);
            ld2(q + (ht ? dt : 0), zt[r][hh][0], zt[r][hh][1]);
        }
        // the west neighbour of pair pc's even cell (the odd cell of pair
        // pc - 1) and the east neighbour of pair pc + 1's odd cell (the even
        // cell of pair pc + 2)
        xw0[r] = p[(long)(1 - ca[r]) * L.half + pos[r] + (hw ? dw : 0)];
        xe1[r] = p[(long)ca[r] * L.half + pos[r] + (he ? de : 0)];
    }
    // y neighbours: cell (i, j +- 1) is in the other half of its row at the
    // same pair position; row 0's north and row 1's south are the other
    // row's values when both rows exist
    {
        const int jr0 = min(j0, L.ny - 1), jr1 = min(j0 + 1, L.ny - 1);
        const int ds0 = jr0 > 0 ? -rl : (L.ny - 1) * rl, dn1 = jr1 < L.ny - 1 ? rl : -(L.ny - 1) * rl;
        const int dn0 = jr0 < L.ny - 1 ? rl : -(L.ny - 1) * rl;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            // cell h of row 0 has colour ca[0] ^ h; its south neighbour lies in
            // the half of that colour's opposite, i.e. half (ca[0] ^ h) ^ 1 of row jr0 - 1
            const double *q0 = p + (long)((ca[0] ^ hh) ^ 1) * L.half + pos[0];
            const double *q1 = p + (long)((ca[1] ^ hh) ^ 1) * L.half + pos[1];
            ld2(q0 + (hs[0] ? ds0 : 0), ys[0][hh][0], ys[0][hh][1]);
            ld2(q1 + (hn[1] ? dn1 : 0), yn[1][hh][0], yn[1][hh][1]);
            if (adj) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    yn[0][hh][c] = v[1][hh][c];
                    ys[1][hh][c] = v[0][hh][c];
                }
            } else {
                ld2(q0 + (hn[0] ? dn0 : 0), yn[0][hh][0], yn[0][hh][1]);
                ys[1][hh][0] = ys[0][hh][0];
                ys[1][hh][1] = ys[0][hh][1];
            }
        }
    }
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        double a[2][2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                // nbsum_pair_vals of cell hh of pair pc + c: its pair's other
                // cell and its other x neighbour (the adjacent pair's cell)
                const double part = v[r][1 - hh][c];
                double xo;
                bool pw = true, pe = true;
                if (hh == 0) {
                    xo = c == 0 ? xw0[r] : v[r][1][0];
                    pw = c == 0 ? hw : true;
                } else {
                    xo = c == 0 ? v[r][0][1] : xe1[r];
                    pe = c == 0 ? true : he;
                }
                const double vw = hh == 0 ? xo : part, ve = hh == 0 ? part : xo;
                const double xw = pw ? vw : 0.0, xe = pe ? ve : 0.0;
                const double xs = hs[r] ? ys[r][hh][c] : 0.0, xn = hn[r] ? yn[r][hh][c] : 0.0;
                const double xb = hb ? zb[r][1 - hh][c] : 0.0, xt = ht ? zt[r][1 - hh][c] : 0.0;
                const double sx = (xw + xe) * ihx2;
                const double syy = (xs + xn) * ihy2;
                const double szz = (xb + xt) * ihz2;
                const int di = diag_count(2 * (pc + c) + hh, L.nx, L.bt[0], L.bt[1]) * 16 + cyz[r];
                const double dg = di == DIAG_INTERIOR ? dg_int : L.diag[di];
                a[hh][c] = dg * v[r][hh][c] - (sx + syy + szz);
            }
        if (!ok[r]) continue;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            st2(Ap + (long)(hh ? 1 - ca[r] : ca[r]) * L.half + pos[r], a[hh][0], a[hh][1]);
#pragma unroll
            for (int c = 0; c < 2; ++c) s += v[r][hh][c] * a[hh][c];
        }
    }
    const int nt = kb.dx * kb.dy, t = kb.ty * kb.dx + kb.tx;
    red[t] = s;
    __syncthreads();
    for (int st = nt / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) partial[((long)kb.bz * kb.gy + kb.by) * kb.gx + kb.bx] = red[0];
}
__global__ __launch_bounds__(RB) void k_matvec2_dot(LevelDev L, const double *__restrict__ p,
                                                    double *__restrict__ Ap, double *__restrict__ partial,
                                                    const double *__restrict__ scal) {
    matvec2_dot_body(hw_blk(), L, p, Ap, partial, scal);
}

// CG iteration k >= 2, first kernel: p_k = z_k + beta_{k-1} p_{k-1} in place,
// and the deferred x_k = x_{k-1} + alpha_{k-1} p_{k-1}
__device__ __forceinline__ void cg_px_body(const Blk &kb, long n, double *__restrict__ p, const double *__restrict__ z,
                                                 double *__restrict__ x, const double *__restrict__ scal) {
    if (scal[SC_CONV] != 0.0) return;
    const double beta = scal[SC_BETA], alpha = scal[SC_ALPHA];
    for (long q = (long)kb.bx * RED_T + kb.tx; q < n; q += (long)kb.gx * RED_T) {
        const double po = p[q];
        p[q] = z[q] + beta * po;
        x[q] = x[q] + alpha * po;
    }
}
__global__ __launch_bounds__(RED_T) void k_cg_px(long n, double *__restrict__ p, const double *__restrict__ z,
                                                 double *__restrict__ x, const double *__restrict__ scal) {
    cg_px_body(hw_blk(), n, p, z, x, scal);
}

// r -= alpha Ap; partial sums of r.r (x is updated one iteration later, by
// k_cg_px or, after the last iteration, by k_gauge)
// (and, in the mixed form, the FP32 copy of r the V-cycle reads)
__device__ __forceinline__ void cg_update_r_body(const Blk &kb, long n, double *__restrict__ r, const double *__restrict__ Ap,
                                                       const double *__restrict__ scal, double *__restrict__ partial,
                                                       float *__restrict__ r32) {
    if (scal[SC_CONV] != 0.0) return;
    __shared__ double red[RED_T];
    const double alpha = scal[SC_ALPHA];
    const long chunk = (n + kb.gx - 1) / kb.gx;
    const long lo = (long)kb.bx * chunk, hi = min(n, lo + chunk);
    double s = 0.0;
    for (long q = lo + kb.tx; q < hi; q += RED_T) {
        double rq = r[q] - alpha * Ap[q];
        r[q] = rq;
        if (r32) r32[q] = float(rq);
        s += rq * rq;
    }
    double t = block_reduce(s, red);
    if (kb.tx == 0) partial[kb.bx] = t;
}
__global__ __launch_bounds__(RED_T) void k_cg_update_r(long n, double *__restrict__ r, const double *__restrict__ Ap,
                                                       const double *__restrict__ scal, double *__restrict__ partial,
                                                       float *__restrict__ r32) {
    cg_update_r_body(hw_blk(), n, r, Ap, scal, partial, r32);
}

// partial sums of a.b
__device__ __forceinline__ void dot_body(const Blk &kb, long n, const double *__restrict__ a,
                                         const double *__restrict__ b, double *__restrict__ partial,
                                         const double *__restrict__ scal) {
    if (scal[SC_CONV] != 0.0) return;
    __shared__ double red[RED_T];
    const long chunk = (n + kb.gx - 1) / kb.gx;
    const long lo = (long)kb.bx * chunk, hi = min(n, lo + chunk);
    // four independent sums over the thread's strided elements (four
    // rounds of loads in flight), then (s0 + s1) + (s2 + s3): a fixed
    // association set by (n, grid, RED_T)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    long q = lo + kb.tx;
    for (; q + 3 * RED_T < hi; q += 4 * RED_T) {
        s0 += a[q] * b[q];
        s1 += a[q + RED_T] * b[q + RED_T];
        s2 += a[q + 2 * RED_T] * b[q + 2 * RED_T];
        s3 += a[q + 3 * RED_T] * b[q + 3 * RED_T];
    }
    for (; q < hi; q += RED_T) s0 += a[q] * b[q];
    double t = block_reduce((s0 + s1) + (s2 + s3), red);
    if (kb.tx == 0) partial[kb.bx] = t;
}
__global__ __launch_bounds__(RED_T) void k_dot(long n, const double *__restrict__ a, const double *__restrict__ b,
                                               double *__restrict__ partial, const double *__restrict__ scal) {
    dot_body(hw_blk(), n, a, b, partial, scal);
}

// finalize a partial sum into scal[slot]; op: 0 store, 1 alpha = rz/pAp,
// 2 beta = rz_new/rz and rz = rz_new, 3 r.r and the convergence test of
// iteration it, 4 ||b||^2 and tol = rtol ||b||
// Each thread sums its strided partials with four independent accumulators
// (the loads of a round are in flight together), then a tree.
constexpr int FIN_T = 1024;
__device__ __forceinline__ void final_body(const Blk &kb, const double *__restrict__ partial, int np, double *__restrict__ scal,
                                                 int slot, int op, int it, double rtol) {
    if (scal[SC_CONV] != 0.0) return;
    __shared__ double red[FIN_T];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int q = kb.tx;
    for (; q + 3 * FIN_T < np; q += 4 * FIN_T)
        for (int u = 0; u < 4; ++u) a[u] += partial[q + u * FIN_T];
    for (int u = 0; q < np; q += FIN_T, ++u) a[u] += partial[q];
    red[kb.tx] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    for (int st = FIN_T / 2; st > 0; st >>= 1) {
        if (kb.tx < st) red[kb.tx] += red[kb.tx + st];
        __syncthreads();
    }
    const double t = red[0];
    if (kb.tx == 0) {
        scal[slot] = t;
        if (op == 1) scal[SC_ALPHA] = scal[SC_RZ] / t;
        if (op == 2) { scal[SC_BETA] = t / scal[SC_RZ]; scal[SC_RZ] = t; }
        if (op == 3) scal[SC_DONE] = it;
        if (op == 3 && (sqrt(t) <= scal[SC_TOL] || !(t == t))) { scal[SC_CONV] = 1.0; scal[SC_ITER] = it; }
        if (op == 4) {
            scal[SC_TOL] = rtol * sqrt(t);
            if (t == 0.0) { scal[SC_CONV] = 1.0; scal[SC_ITER] = 0; scal[SC_RR] = 0.0; }
        }
    }
}
__global__ __launch_bounds__(FIN_T) void k_final(const double *__restrict__ partial, int np, double *__restrict__ scal,
                                                 int slot, int op, int it, double rtol) {
    final_body(hw_blk(), partial, np, scal, slot, op, it, rtol);
}


// r = rhs with the pinned row, x = 0; r and x in level-0 storage, rhs in
// field order (cell (0,0,0) is position 0 of both). A thread takes the
// x-pair (2p, 2p+1) of row (j, k), blocks of shape_for over the level's x-pairs.
__device__ __forceinline__ void cg_init_body(const Blk &kb, const LevelDev &L, const double *__restrict__ rhs,
                                             double *__restrict__ r, double *__restrict__ x,
                                             double *__restrict__ scal, float *__restrict__ r32) {
    const int ip = kb.bx * kb.dx + kb.tx;
    const int j = kb.by * kb.dy + kb.ty, k = kb.bz;
    const double pin = scal[SC_PIN];
    if (kb.bx == 0 && kb.by == 0 && kb.bz == 0 && kb.tx == 0 && kb.ty == 0) {
        scal[SC_CONV] = 0.0;
        scal[SC_ITER] = 0.0;
        scal[SC_DONE] = 0.0;
    }
    if (j >= L.ny || 2 * ip >= L.nx) return;
    // both cells' right-hand sides are loaded before the stores (the second
    // from the first's place when the row ends at an odd length)
    const bool two = 2 * ip + 1 < L.nx;
    const long f0 = ((long)k * L.ny + j) * L.nx + 2 * ip;
    const double v[2] = {rhs[f0], rhs[f0 + (two ? 1 : 0)]};
    for (int h = 0; h < (two ? 2 : 1); ++h) {
        const int q = lidx(L, 2 * ip + h, j, k);
        const double rq = (q == 0 && !L.dirichlet) ? -pin : v[h];
        r[q] = rq;
        if (r32) r32[q] = float(rq);   // the mixed form's FP32 copy
        x[q] = 0.0;
    }
}
__global__ __launch_bounds__(RB) void k_cg_init(LevelDev L, const double *__restrict__ rhs,
                                                double *__restrict__ r, double *__restrict__ x,
                                                double *__restrict__ scal, float *__restrict__ r32) {
    cg_init_body(hw_blk(), L, rhs, r, x, scal, r32);
}

__device__ __forceinline__ void copy_body(const Blk &kb, long n, double *__restrict__ d, const double *__restrict__ s) {
    for (long q = (long)kb.bx * RED_T + kb.tx; q < n; q += (long)kb.gx * RED_T) d[q] = s[q];
}
__global__ __launch_bounds__(RED_T) void k_copy(long n, double *__restrict__ d, const double *__restrict__ s) {
    copy_body(hw_blk(), n, d, s);
}
template <typename TD, typename TS>
__global__ __launch_bounds__(RED_T) void k_convert(long n, TD *__restrict__ d, const TS *__restrict__ s,
                                                   const double *__restrict__ stop) {
    if (*stop != 0.0) return;
    for (long q = (long)blockIdx.x * RED_T + threadIdx.x; q < n; q += (long)gridDim.x * RED_T) d[q] = TD(s[q]);
}

// gauge P'(1,1,1) = 0: pcorr = x - x(1,1,1), x in level-0 storage, pcorr in
// field order (no shift with an outflow side, where P' = 0 on the face), x
// taking the last completed CG iteration's deferred update x + alpha p
// (thread per x-pair of a row, blocks of shape_for over the x-pairs)
__device__ __forceinline__ void gauge_body(const Blk &kb, const LevelDev &L, const double *__restrict__ x,
                                              const double *__restrict__ p, const double *__restrict__ scal,
                                              double *__restrict__ pc) {
    const int done = (int)scal[SC_DONE];
    const double alpha = done > 0 ? scal[SC_ALPHA] : 0.0;
    auto xv = [&](long q) { return done > 0 ? x[q] + alpha * p[q] : x[q]; };
    const double x0 = L.dirichlet ? 0.0 : xv(0);
    const int ip = kb.bx * kb.dx + kb.tx;
    const int j = kb.by * kb.dy + kb.ty, k = kb.bz;
    if (j >= L.ny || 2 * ip >= L.nx) return;
    // both cells' values are loaded before the stores (the second from the
    // first's place when the row ends at an odd length)
    const bool two = 2 * ip + 1 < L.nx;
    const double v0 = xv(lidx(L, 2 * ip, j, k)), v1 = xv(lidx(L, two ? 2 * ip + 1 : 2 * ip, j, k));
    const long f0 = ((long)k * L.ny + j) * L.nx + 2 * ip;
    pc[f0] = v0 - x0;
    if (two) pc[f0 + 1] = v1 - x0;
}
__global__ __launch_bounds__(RB) void k_gauge(LevelDev L, const double *__restrict__ x,
                                              const double *__restrict__ p, const double *__restrict__ scal,
                                              double *__restrict__ pc) {
    gauge_body(hw_blk(), L, x, p, scal, pc);
}

int stream_blocks(long n) { return (int)std::min<long>(4096, std::max<long>(1, (n + RED_T - 1) / RED_T)); }

LevelDev make_level(int nx, int ny, int nz, const int bt[6], const double h[3]) {
    LevelDev L;
    std::memset(&L, 0, sizeof L);
    L.nx = nx; L.ny = ny; L.nz = nz;
    for (int s = 0; s < 6; ++s) L.bt[s] = bt[s];
    L.ihx2 = 1.0 / (h[0] * h[0]);
    L.ihy2 = 1.0 / (h[1] * h[1]);
    L.ihz2 = 1.0 / (h[2] * h[2]);
    for (int cx = 0; cx < 4; ++cx)
        for (int cy = 0; cy < 4; ++cy)
            for (int cz = 0; cz < 4; ++cz) {
                double dsum0 = cx * L.ihx2, dsum1 = cy * L.ihy2, dsum2 = cz * L.ihz2;
                double d = dsum0 + dsum1 + dsum2;
                L.diag[cx * 16 + cy * 4 + cz] = d;
                L.idiag[cx * 16 + cy * 4 + cz] = d > 0 ? 1.0 / d : 0.0;
            }
    L.dirichlet = 0;
    for (int s = 0; s < 6; ++s)
        if (bt[s] == IBMDEM_BC_OUTFLOW) L.dirichlet = 1;
    const int n[3] = {nx, ny, nz};
    L.rb_valid = 1;
    for (int d = 0; d < 3; ++d)
        if (bt[2 * d] == IBMDEM_BC_PERIODIC && (n[d] % 2)) L.rb_valid = 0;
    L.split = (nx % 2) == 0 ? 1 : 0;
    L.half = L.split ? nx * ny * nz / 2 : 0;
    L.m_row = udiv_magic(L.split ? nx / 2 : nx);
    L.m_ny = udiv_magic(ny);
    L.m_nxy = udiv_magic(nx * ny);
    return L;
}

// one half sweep of a level, two-pair or one-pair threads as sweep_shape says
template <typename T, bool DOT, int RES, bool ZI>
void launch_half(const SweepShape &sw, hipStream_t s, const LevelDev &d, T *x, const T *b, int color,
                 const double *stop, double *partial, long pofs, T *r, const double *r64 = nullptr,
                 double *z64 = nullptr) {
    if (sw.x2)
        hipLaunchKernelGGL((k_rb_half<T, T, DOT, RES, ZI, true>), sw.grid, sw.block, 0, s, d, x, b, color, stop,
                           partial, pofs, r, r64, z64);
    else
        hipLaunchKernelGGL((k_rb_half<T, T, DOT, RES, ZI>), sw.grid, sw.block, 0, s, d, x, b, color, stop, partial,
                           pofs, r, r64, z64);
}

template <typename T>
int vcycle(ibmdem_ctx *c, MGLevel<T> *lev, int l, bool dot = false) {
    MGLevel<T> &L = lev[l];
    const LevelDev &d = L.d;
    hipStream_t s = c->stream;
    if (l >= c->tail_first) {
        hipLaunchKernelGGL((k_vcycle_tail<T>), dim3(1), dim3(TAIL_T), 0, s, c->dlev, l, c->nlev - 1, L.b, L.x,
                           c->cs.mg_pre, c->cs.mg_post, c->cs.mg_coarse_sweeps, c->scal + SC_CONV);
        CHECK_LAUNCH(c);
        return IBMDEM_OK;
    }
    if (l == c->nlev - 1) {
        if (L.n <= COARSE_LDS)
            hipLaunchKernelGGL((k_coarse_lds<T>), dim3(1), dim3(64), 0, s, d, L.x, L.b, c->cs.mg_coarse_sweeps,
                               c->scal + SC_CONV);
        else
            hipLaunchKernelGGL((k_coarse<T>), dim3(1), dim3(COARSE_T), 0, s, d, L.x, L.b, L.r,
                               c->cs.mg_coarse_sweeps, c->scal + SC_CONV);
        CHECK_LAUNCH(c);
        return IBMDEM_OK;
    }
    // level-0 half-sweeps are the profiled (dominant) kernel: each reads the
    // right-hand side and the other colour of half the cells and writes them;
    // the residual variant also writes r of its half, the mixed form's r.z
    // variant also reads the FP64 r and writes the FP64 z of its half
    const double hb = l == 0 ? 1.5 * sizeof(T) * (double)L.n : 0.0;
    const double hb_res = hb + (l == 0 ? 0.5 * sizeof(T) * (double)L.n : 0.0);
    const double hb_dot = hb + (l == 0 && !std::is_same<T, double>::value ? 8.0 * (double)L.n : 0.0);
    const SweepShape sw = sweep_shape(d);
    const long nbh = (long)sw.grid.x * sw.grid.y * sw.grid.z;
    // The pre- and the post-smoothing sweeps at level 0 are each timed as one
    // group of back-to-back launches (one event pair per group).
    double grp_bytes = 0.0;
    int grp_n = 0;
    auto group_begin = [&]() {
        if (l == 0) prof_begin(c, IBMDEM_K_MG_SMOOTH);
        grp_bytes = 0.0;
        grp_n = 0;
    };
    auto group_end = [&]() {
        if (l == 0) prof_end(c, IBMDEM_K_MG_SMOOTH, grp_bytes, grp_n);
    };
    // res: 1 = the sweep also leaves the residual of its cells
    auto half = [&](int color, int zero, int dot_slot = -1, int res = 0) {
        ++grp_n;
        if (dot_slot >= 0) {
            // the mixed form reads the FP64 r and writes the FP64 z as well
            const bool mixed = !std::is_same<T, double>::value;
            launch_half<T, true, 0, false>(sw, s, d, L.x, L.b, color, c->scal + SC_CONV, c->partial,
                                           dot_slot * nbh, (T *)nullptr, mixed ? c->cg_r : nullptr,
                                           mixed ? c->cg_z : nullptr);
            grp_bytes += hb_dot;
            return;
        }
        if (res)
            launch_half<T, false, 1, false>(sw, s, d, L.x, L.b, color, c->scal + SC_CONV, nullptr, 0L, L.r);
        else if (zero)
            launch_half<T, false, 0, true>(sw, s, d, L.x, L.b, color, c->scal + SC_CONV, nullptr, 0L, nullptr);
        else
            launch_half<T, false, 0, false>(sw, s, d, L.x, L.b, color, c->scal + SC_CONV, nullptr, 0L, nullptr);
        grp_bytes += res ? hb_res : hb;
    };
    // pre-smoothing (red, black); the last black sweep also leaves r of the
    // black cells, then r of the red cells
    // below level 0 the restriction that wrote b also did the first sweep
    group_begin();
    for (int q = 0; q < c->cs.mg_pre; ++q) {
        if (q > 0 || l == 0) half(0, q == 0 ? 1 : 0);
        half(1, 0, -1, q == c->cs.mg_pre - 1 ? 1 : 0);
        CHECK_LAUNCH(c);
    }
    group_end();
    launch_half<T, false, 2, false>(sw, s, d, L.x, L.b, 0, c->scal + SC_CONV, nullptr, 0L, L.r);
    CHECK_LAUNCH(c);
    MGLevel<T> &C = lev[l + 1];
    const dim3 cs(div_up(C.d.nx, RS_CX), div_up(C.d.ny, RS_CY), C.d.nz);
    // the coarse level's first sweep in the restriction when that level is
    // swept by launches (not the work-group tail, not the coarsest)
    const bool czero = l + 1 < c->tail_first && l + 1 != c->nlev - 1;
    hipLaunchKernelGGL((k_restrict<T>), cs, dim3(RB), 0, s, d, L.r, C.d, C.b, czero ? C.x : (T *)nullptr,
                       c->scal + SC_CONV);
    CHECK_LAUNCH(c);
    int rc = vcycle<T>(c, lev, l + 1);
    if (rc) return rc;
    const ProlongShape ps = prolong_shape(d);
    if (!ps.ok) {
        return set_error(c, IBMDEM_ERR_ARG, "prolongation tile of " + std::to_string(ps.block.x) + "x" +
                                                std::to_string(ps.block.y) + " threads does not fit");
    }
    if (ps.x2)
        hipLaunchKernelGGL((k_prolong2<T>), ps.grid, ps.block, 0, s, d, L.x, C.d, C.x, c->scal + SC_CONV);
    else
        hipLaunchKernelGGL((k_prolong<T>), ps.grid, ps.block, 0, s, d, L.x, C.d, C.x, c->scal + SC_CONV);
    CHECK_LAUNCH(c);
    group_begin();
    for (int q = 0; q < c->cs.mg_post; ++q) {
        // the last post sweep at level 0 finalizes z: with dot, it also
        // leaves the partial sums of r.z (colour 1, then colour 0)
        const bool last = dot && l == 0 && q == c->cs.mg_post - 1;
        half(1, 0, last ? 0 : -1);
        half(0, 0, last ? 1 : -1);
        CHECK_LAUNCH(c);
    }
    group_end();
    return IBMDEM_OK;
}

int dot_into(ibmdem_ctx *c, const double *a, const double *b, int slot, int op);

// z = M^{-1} r
int precondition(ibmdem_ctx *c) {
    const long n = c->n;
    if (c->cs.precision == IBMDEM_MIXED) {
        hipLaunchKernelGGL((k_convert<float, double>), dim3(stream_blocks(n)), dim3(RED_T), 0, c->stream, n,
                           c->lev32[0].b, c->cg_r, c->scal + SC_CONV);
        CHECK_LAUNCH(c);
        int rc = vcycle<float>(c, c->lev32, 0);
        if (rc) return rc;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_blocks(n)), dim3(RED_T), 0, c->stream, n,
                           c->cg_z, c->lev32[0].x, c->scal + SC_CONV);
        CHECK_LAUNCH(c);
        return IBMDEM_OK;
    }
    c->lev64[0].b = c->cg_r;
    c->lev64[0].x = c->cg_z;
    return vcycle<double>(c, c->lev64, 0);
}

// z = M^{-1} r, then r.z into scal[slot] with k_final's op. In the FP64 form
// with post-smoothing the level-0 sweeps that finalize z also leave the
// partial sums of r.z (one per block and colour), so z is not read again.
int precondition_dot(ibmdem_ctx *c, int slot, int op) {
    if (c->cs.mg_post < 1 || c->nlev < 2) {
        int rc = precondition(c);
        if (rc) return rc;
        return dot_into(c, c->cg_r, c->cg_z, slot, op);
    }
    int rc;
    if (c->cs.precision == IBMDEM_MIXED) {
        // the FP32 copy of r was written with r (k_cg_init, k_cg_update_r);
        // the last level-0 sweeps write the FP64 z
        rc = vcycle<float>(c, c->lev32, 0, true);
    } else {
        c->lev64[0].b = c->cg_r;
        c->lev64[0].x = c->cg_z;
        rc = vcycle<double>(c, c->lev64, 0, true);
    }
    if (rc) return rc;
    const dim3 g = sweep_shape(c->lev64[0].d).grid;   // the r.z sweeps' blocks
    const int np = 2 * (int)(g.x * g.y * g.z);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(FIN_T), 0, c->stream, c->partial, np, c->scal, slot, op, 0,
                       c->cs.poisson_rtol);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int dot_into(ibmdem_ctx *c, const double *a, const double *b, int slot, int op) {
    hipLaunchKernelGGL(k_dot, dim3(RED_BLOCKS), dim3(RED_T), 0, c->stream, c->n, a, b, c->partial, c->scal);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(FIN_T), 0, c->stream, c->partial, RED_BLOCKS, c->scal, slot, op, 0,
                       c->cs.poisson_rtol);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int read_scalars(ibmdem_ctx *c) {
    HIPX(c, hipMemcpyAsync(c->scal_host, c->scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    return IBMDEM_OK;
}

}  // namespace

// partial sums the level-0 matrix-vector product (one per block) and the
// last level-0 sweeps with r.z (one per block and colour) write
long poisson_partials(int nx, int ny, int nz) {
    const LaunchShape mv = shape_for(div_up(nx, 2), ny, nz);
    const dim3 h = half_grid(nx, ny, nz);
    return std::max<long>((long)mv.grid.x * mv.grid.y * mv.grid.z, 2L * h.x * h.y * h.z);
}

int poisson_setup(ibmdem_ctx *c) {
    int bt[6];
    for (int s = 0; s < 6; ++s) bt[s] = c->g.bt[s];
    double h[3] = {c->g.dx, c->g.dy, c->g.dz};
    int nx = c->g.nx, ny = c->g.ny, nz = c->g.nz;
    const int maxl = c->cs.mg_max_levels > 0 ? std::min(c->cs.mg_max_levels, IBMDEM_MAX_LEVELS) : IBMDEM_MAX_LEVELS;
    c->nlev = 0;
    while (true) {
        LevelDev L = make_level(nx, ny, nz, bt, h);
        long n = (long)nx * ny * nz;
        const int l = c->nlev;
        c->lev64[l].d = L; c->lev64[l].n = n;
        c->lev32[l].d = L; c->lev32[l].n = n;
        int rc;
        if (c->cs.precision == IBMDEM_MIXED) {
            if ((rc = dev_alloc(c, &c->lev32[l].x, n))) return rc;
            if ((rc = dev_alloc(c, &c->lev32[l].b, n))) return rc;
            if ((rc = dev_alloc(c, &c->lev32[l].r, n))) return rc;
        } else {
            if (l > 0) {
                if ((rc = dev_alloc(c, &c->lev64[l].x, n))) return rc;
                if ((rc = dev_alloc(c, &c->lev64[l].b, n))) return rc;
            }
            if ((rc = dev_alloc(c, &c->lev64[l].r, n))) return rc;
        }
        c->nlev++;
        if (c->nlev >= maxl) break;
        if ((nx % 2) || (ny % 2) || (nz % 2)) break;
        if (nx < 4 || ny < 4 || nz < 4) break;
        nx /= 2; ny /= 2; nz /= 2;
        h[0] *= 2; h[1] *= 2; h[2] *= 2;
    }
    // the small levels whose x, b and r fit the tail work-group's LDS
    // together run as one launch per V-cycle, when the coarsest level suits
    // the one-wavefront table sweeps (IBMDEM_NO_TAIL=1 keeps one launch per
    // operation, for comparison)
    const size_t elem = c->cs.precision == IBMDEM_MIXED ? sizeof(float) : sizeof(double);
    const LevelDev &lc = c->lev64[c->nlev - 1].d;
    c->tail_first = c->nlev;
    if (c->nlev >= 2 && lc.rb_valid && c->lev64[c->nlev - 1].n <= COARSE_LDS) {
        size_t need = elem;   // the zero slot
        for (int l = c->nlev - 1; l >= 1; --l) {
            need += 3 * (size_t)c->lev64[l].n * elem;
            if (need > (size_t)TAIL_BYTES || c->nlev - l > TAIL_MAXLEV) break;
            c->tail_first = l;
        }
        if (c->tail_first == c->nlev - 1) c->tail_first = c->nlev;   // the coarsest alone: k_coarse_lds
    }
    const char *nt = std::getenv("IBMDEM_NO_TAIL");
    if (nt && *nt && *nt != '0') c->tail_first = c->nlev;
    int rc;
    if ((rc = dev_alloc(c, &c->dlev, (size_t)c->nlev))) return rc;
    LevelDev host[IBMDEM_MAX_LEVELS];
    for (int l = 0; l < c->nlev; ++l) host[l] = c->lev64[l].d;
    HIPX(c, hipMemcpyAsync(c->dlev, host, c->nlev * sizeof(LevelDev), hipMemcpyHostToDevice, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
    return IBMDEM_OK;
}

namespace {

// one CG iteration (it is its 1-based number), queued without any host wait
int cg_iteration(ibmdem_ctx *c, int it) {
    const long n = c->n;
    hipStream_t s = c->stream;
    int rc;
    const LevelDev &L0 = c->lev64[0].d;
    const SweepShape mv = sweep_shape(L0);   // two rows per thread, two pairs where the level allows
    const int mv_blocks = (int)(mv.grid.x * mv.grid.y * mv.grid.z);
    if (it > 1) {
        hipLaunchKernelGGL(k_cg_px, dim3(stream_blocks(n)), dim3(RED_T), 0, s, n, c->cg_p, c->cg_z, c->cg_x,
                           c->scal);
        CHECK_LAUNCH(c);
    }
    prof_begin(c, IBMDEM_K_MATVEC);
    if (mv.x2)
        hipLaunchKernelGGL(k_matvec2_dot, mv.grid, mv.block, 0, s, L0, c->cg_p, c->cg_Ap, c->partial, c->scal);
    else
        hipLaunchKernelGGL(k_matvec_dot, mv.grid, mv.block, 0, s, L0, c->cg_p, c->cg_Ap, c->partial, c->scal);
    prof_end(c, IBMDEM_K_MATVEC, 2.0 * 8.0 * (double)n);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(FIN_T), 0, s, c->partial, mv_blocks, c->scal, SC_PAP, 1, it,
                       c->cs.poisson_rtol);
    CHECK_LAUNCH(c);
    prof_begin(c, IBMDEM_K_CG_UPDATE);
    hipLaunchKernelGGL(k_cg_update_r, dim3(RED_BLOCKS), dim3(RED_T), 0, s, n, c->cg_r, c->cg_Ap, c->scal,
                       c->partial, c->cs.precision == IBMDEM_MIXED ? c->lev32[0].b : (float *)nullptr);
    prof_end(c, IBMDEM_K_CG_UPDATE, 3.0 * 8.0 * (double)n);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(FIN_T), 0, s, c->partial, RED_BLOCKS, c->scal, SC_RR, 3, it,
                       c->cs.poisson_rtol);
    CHECK_LAUNCH(c);
    return precondition_dot(c, SC_RZ_NEW, 2);
}

}  // namespace

// Start of the solve, queued without any host wait, in two parts: first
// the pinned-row right-hand side, r = b, x = 0, ||b|| and the tolerance (a
// zero right-hand side converges at once); then z = M^-1 r, p = z, r.z
// (which a batch runs for all its cases at once).
int poisson_begin_rhs(ibmdem_ctx *c) {
    const long n = c->n;
    hipStream_t s = c->stream;
    int rc;
    // pinned row: b[0] = -sum_{q>0} b[q] (none with an outflow side)
    if (!c->lev64[0].d.dirichlet && (rc = reduce_sum(c, c->rhs, nullptr, n, SC_PIN, 1))) return rc;
    const LevelDev &L0 = c->lev64[0].d;
    const LaunchShape fs = shape_for(div_up(L0.nx, 2), L0.ny, L0.nz);
    hipLaunchKernelGGL(k_cg_init, fs.grid, fs.block, 0, s, L0, c->rhs, c->cg_r, c->cg_x, c->scal,
                       c->cs.precision == IBMDEM_MIXED ? c->lev32[0].b : (float *)nullptr);
    CHECK_LAUNCH(c);
    return dot_into(c, c->cg_r, c->cg_r, SC_BNORM2, 4);
}

int poisson_begin_precondition(ibmdem_ctx *c) {
    int rc;
    if ((rc = precondition_dot(c, SC_RZ, 0))) return rc;
    hipLaunchKernelGGL(k_copy, dim3(stream_blocks(c->n)), dim3(RED_T), 0, c->stream, c->n, c->cg_p, c->cg_z);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int poisson_begin(ibmdem_ctx *c) {
    int rc;
    if ((rc = poisson_begin_rhs(c))) return rc;
    return poisson_begin_precondition(c);
}

// Iterations first..last, queued without any host wait. Once the device sets
// SC_CONV every later CG kernel returns at once, so x is the iterate of the
// converging iteration however many iterations were queued.
int poisson_iterations(ibmdem_ctx *c, int first, int last) {
    for (int it = first; it <= last; ++it) {
        int rc = cg_iteration(c, it);
        if (rc) return rc;
    }
    return IBMDEM_OK;
}

// Iterations from `first` until convergence or the cap: the host keeps at
// most two iterations in flight and reads the convergence flag of the older
// one, so the stream never drains inside the solve.
int poisson_iterate_to_convergence(ibmdem_ctx *c, int first) {
    hipStream_t s = c->stream;
    const int maxit = c->cs.poisson_max_iter;
    for (int q = 0; q <= maxit + 1; ++q) c->conv_host[q] = 0.0;
    HIPX(c, hipMemcpyAsync(c->conv_host + first - 1, c->scal + SC_CONV, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPX(c, hipEventRecord(c->iter_ev[(first - 1) & 1], s));
    for (int it = first; it <= maxit; ++it) {
        int rc = cg_iteration(c, it);
        if (rc) return rc;
        HIPX(c, hipMemcpyAsync(c->conv_host + it, c->scal + SC_CONV, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPX(c, hipEventRecord(c->iter_ev[it & 1], s));
        // wait for the previous iteration, then look at its flag
        HIPX(c, hipEventSynchronize(c->iter_ev[(it - 1) & 1]));
        if (c->conv_host[it - 1] != 0.0) break;
    }
    return IBMDEM_OK;
}

// Wait for the solve, report it, fail on a non-converged or non-finite one.
int poisson_check(ibmdem_ctx *c, ibmdem_solve_info *info) {
    int rc;
    if ((rc = read_scalars(c))) return rc;
    return poisson_check_scalars(c, nullptr, info);
}

// The same report from solver scalars already on the host (scal, or the
// context's own copy when null).
int poisson_check_scalars(ibmdem_ctx *c, const double *scal, ibmdem_solve_info *info) {
    if (scal) std::memcpy(c->scal_host, scal, SC_COUNT * sizeof(double));
    const double bnorm = std::sqrt(c->scal_host[SC_BNORM2]);
    const int conv = c->scal_host[SC_CONV] != 0.0;
    const double rr = bnorm > 0 ? c->scal_host[SC_RR] : 0.0;
    const int it = conv ? (int)c->scal_host[SC_ITER] : c->cs.poisson_max_iter;
    c->last_cg_iterations = it;
    if (info) {
        info->iterations = it;
        info->converged = conv;
        info->rel_residual = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
        info->rhs_norm = bnorm;
    }
    // a non-finite right-hand side (||b|| is then NaN or inf) or residual
    if (!(bnorm - bnorm == 0.0) || !(rr == rr))
        return set_error(c, IBMDEM_ERR_NAN, "non-finite right-hand side or residual in the Poisson solve");
    if (!conv) return set_error(c, IBMDEM_ERR_NOT_CONVERGED, "Poisson solve did not converge");
    return IBMDEM_OK;
}

// P' = x - x(1,1,1): the reference's gauge (poisson.F90:78,97-100)
int poisson_gauge(ibmdem_ctx *c) {
    const LevelDev &L0 = c->lev64[0].d;
    const LaunchShape fs = shape_for(div_up(L0.nx, 2), L0.ny, L0.nz);
    hipLaunchKernelGGL(k_gauge, fs.grid, fs.block, 0, c->stream, L0, c->cg_x, c->cg_p, c->scal, c->pcorr);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int poisson_solve(ibmdem_ctx *c, ibmdem_solve_info *info) {
    int rc;
    if ((rc = poisson_begin(c))) return rc;
    if ((rc = poisson_iterate_to_convergence(c, 1))) return rc;
    // an unconverged solve still delivers its last iterate, as the oracle does
    const int chk = poisson_check(c, info);
    if (chk && chk != IBMDEM_ERR_NOT_CONVERGED) return chk;
    if ((rc = poisson_gauge(c))) return rc;
    return chk;
}

// ---------------------------------------------------------------------------
// The batched solve: one launch per solver operation over a group of cases.
//
// A group's cases share the precision, the smoothing counts and the
// iteration cap; grids, level counts, where the one-work-group tail starts
// and the coarsest-level sweeps are per case. A launch runs each member's
// blocks with the body, block decomposition and arguments of the member's
// own launch, and a member sees its operations in the order its own solve
// queues them, so its arithmetic is bitwise that of solving it alone. A
// block finds its case in a table of first blocks (each a multiple of 8, so
// a case's blocks keep the XCD placement of its own launch); a member that
// has converged returns at once, as it does alone.
// ---------------------------------------------------------------------------
namespace {

template <typename T>
struct RbArg {
    BGrid g;
    const LevelDev *L;
    T *x;
    const T *b;
    const double *stop;
    double *partial;
    long pofs;
    T *r;
    const double *r64;
    double *z64;
};
template <typename T, bool DOT, int RES, bool ZI, bool X2>
__global__ __launch_bounds__(RB) void kb_rb_half(BTable<RbArg<T>> tab, int color) {
    batch_for(tab, [&](const RbArg<T> &a, const Blk &kb) {
        rb_half_body<T, T, DOT, RES, ZI, X2>(kb, *cst(a.L), glb(a.x), glb(a.b), color, glb(a.stop), glb(a.partial),
                                             a.pofs, glb(a.r), glb(a.r64), glb(a.z64));
    });
}

template <typename T>
struct RsArg {
    BGrid g;
    const LevelDev *F, *C;
    const T *rf;
    T *bc, *xc;   // xc: the coarse level's first sweep too (null: not)
    const double *stop;
};
template <typename T>
__global__ __launch_bounds__(RB) void kb_restrict(BTable<RsArg<T>> tab) {
    batch_for(tab, [&](const RsArg<T> &a, const Blk &kb) {
        restrict_body<T>(kb, *cst(a.F), glb(a.rf), *cst(a.C), glb(a.bc), a.xc ? glb(a.xc) : (T *)nullptr,
                         glb(a.stop));
    });
}

template <typename T>
struct PrArg {
    BGrid g;
    const LevelDev *F, *C;
    T *xf;
    const T *xc;
    const double *stop;
};
template <typename T>
__global__ __launch_bounds__(RB) void kb_prolong(BTable<PrArg<T>> tab) {
    batch_for(tab, [&](const PrArg<T> &a, const Blk &kb) {
        prolong_body<T>(kb, *cst(a.F), glb(a.xf), *cst(a.C), glb(a.xc), glb(a.stop));
    });
}
template <typename T>
__global__ __launch_bounds__(RB) void kb_prolong2(BTable<PrArg<T>> tab) {
    batch_for(tab, [&](const PrArg<T> &a, const Blk &kb) {
        prolong2_body<T>(kb, *cst(a.F), glb(a.xf), *cst(a.C), glb(a.xc), glb(a.stop));
    });
}

// the bottom of a case's V-cycle: the one-work-group tail (first < last),
// or the coarsest level alone from LDS tables or in global memory
template <typename T>
struct BtArg {
    BGrid g;
    const LevelDev *lev;
    int first, last, pre, post, sweeps;
    T *x;
    const T *b;
    T *tmp;
    const double *stop;
};
template <typename T>
__global__ __launch_bounds__(TAIL_T) void kb_tail(BTable<BtArg<T>> tab) {
    batch_for(tab, [&](const BtArg<T> &a, const Blk &kb) {
        vcycle_tail_body<T>(kb, cst(a.lev), a.first, a.last, glb(a.b), glb(a.x), a.pre, a.post, a.sweeps,
                            glb(a.stop));
    });
}
template <typename T>
__global__ __launch_bounds__(64) void kb_coarse_lds(BTable<BtArg<T>> tab) {
    batch_for(tab, [&](const BtArg<T> &a, const Blk &kb) {
        coarse_lds_body<T>(kb, *cst(a.lev + a.last), glb(a.x), glb(a.b), a.sweeps, glb(a.stop));
    });
}
template <typename T>
__global__ __launch_bounds__(COARSE_T) void kb_coarse(BTable<BtArg<T>> tab) {
    batch_for(tab, [&](const BtArg<T> &a, const Blk &kb) {
        coarse_body<T>(kb, *cst(a.lev + a.last), glb(a.x), glb(a.b), glb(a.tmp), a.sweeps, glb(a.stop));
    });
}

// the CG vector operations and the scalar step
struct CgArg {
    BGrid g;
    const LevelDev *L;
    long n;
    double *p, *z, *x, *r, *Ap;
    double *partial;
    double *scal;
    float *r32;
    int np;           // partial sums k_final adds
    double rtol;
    const double *rhs;   // the solve's right-hand side (field order)
};
__global__ __launch_bounds__(RB) void kb_matvec_dot(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        matvec_dot_body(kb, *cst(a.L), glb(a.p), glb(a.Ap), glb(a.partial), glb(a.scal));
    });
}
__global__ __launch_bounds__(RB) void kb_matvec2_dot(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        matvec2_dot_body(kb, *cst(a.L), glb(a.p), glb(a.Ap), glb(a.partial), glb(a.scal));
    });
}
__global__ __launch_bounds__(RED_T) void kb_cg_px(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) { cg_px_body(kb, a.n, glb(a.p), glb(a.z), glb(a.x), glb(a.scal)); });
}
__global__ __launch_bounds__(RED_T) void kb_cg_update_r(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        cg_update_r_body(kb, a.n, glb(a.r), glb(a.Ap), glb(a.scal), glb(a.partial), glb(a.r32));
    });
}
__global__ __launch_bounds__(RED_T) void kb_copy(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) { copy_body(kb, a.n, glb(a.p), glb(a.z)); });
}
__global__ __launch_bounds__(FIN_T) void kb_final(BTable<CgArg> tab, int slot, int op, int it) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        final_body(kb, glb(a.partial), a.np, glb(a.scal), slot, op, it, a.rtol);
    });
}

// every case's solver scalars into one array (case q at q * SC_COUNT)
__global__ __launch_bounds__(64) void kb_gather_scalars(const double *const *__restrict__ scal, int ncase,
                                                        double *__restrict__ out) {
    const int q = blockIdx.x;
    if (q < ncase && threadIdx.x < SC_COUNT) out[(long)q * SC_COUNT + threadIdx.x] = glb(scal[q])[threadIdx.x];
}

// Profiling of the batched solve: while armed, run_ops records an event
// before the first launch and after every launch, so the elapsed time
// between two events is one launch's run time once the stream runs behind
// the host (the launches are queued far faster than they run).
constexpr int BK_COUNT = 9;
struct BProf {
    bool on = false, armed = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    std::vector<double> bytes;
    int n = 0;                          // launches recorded
};
template <typename T, bool DOT, int RES, bool ZI, bool X2>
void launch_rb(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL((kb_rb_half<T, DOT, RES, ZI, X2>), dim3(o.blocks), dim3(RB), 0, s, table<RbArg<T>>(o),
                       o.color);
}
template <typename T>
void launch_restrict(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_restrict<T>, dim3(o.blocks), dim3(RB), 0, s, table<RsArg<T>>(o));
}
template <typename T>
void launch_prolong(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_prolong<T>, dim3(o.blocks), dim3(RB), 0, s, table<PrArg<T>>(o));
}
template <typename T>
void launch_prolong2(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_prolong2<T>, dim3(o.blocks), dim3(RB), 0, s, table<PrArg<T>>(o));
}
template <typename T>
void launch_tail(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_tail<T>, dim3(o.blocks), dim3(TAIL_T), 0, s, table<BtArg<T>>(o));
}
template <typename T>
void launch_coarse_lds(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_coarse_lds<T>, dim3(o.blocks), dim3(64), 0, s, table<BtArg<T>>(o));
}
template <typename T>
void launch_coarse(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_coarse<T>, dim3(o.blocks), dim3(COARSE_T), 0, s, table<BtArg<T>>(o));
}
void launch_matvec(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_matvec_dot, dim3(o.blocks), dim3(RB), 0, s, table<CgArg>(o));
}
void launch_matvec2(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_matvec2_dot, dim3(o.blocks), dim3(RB), 0, s, table<CgArg>(o));
}
void launch_px(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_cg_px, dim3(o.blocks), dim3(RED_T), 0, s, table<CgArg>(o));
}
void launch_update_r(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_cg_update_r, dim3(o.blocks), dim3(RED_T), 0, s, table<CgArg>(o));
}
void launch_copy(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_copy, dim3(o.blocks), dim3(RED_T), 0, s, table<CgArg>(o));
}
void launch_final(const BOp &o, hipStream_t s, int it) {
    hipLaunchKernelGGL(kb_final, dim3(o.blocks), dim3(FIN_T), 0, s, table<CgArg>(o), o.slot, o.op,
                       o.with_it ? it : 0);
}

template <typename T>
MGLevel<T> *levels(ibmdem_ctx *c);
template <>
MGLevel<double> *levels<double>(ibmdem_ctx *c) { return c->lev64; }
template <>
MGLevel<float> *levels<float>(ibmdem_ctx *c) { return c->lev32; }

// the level where a case's V-cycle turns: its tail's first level, else its coarsest
int bottom_level(const ibmdem_ctx *c) { return std::min(c->tail_first, c->nlev - 1); }

// z = M^-1 r for every case, with r.z into scal[slot] (k_final's op): the
// batched form of vcycle<T>(c, lev, 0, true) and precondition_dot
template <typename T>
int plan_precondition_dot(Arena &ar, std::vector<BOp> &ops, const std::vector<ibmdem_ctx *> &cs, int slot,
                          int op, std::string &err) {
    const bool mixed = !std::is_same<T, double>::value;
    const int pre = cs[0]->cs.mg_pre, post = cs[0]->cs.mg_post;
    int depth = 0;
    for (ibmdem_ctx *c : cs) depth = std::max(depth, bottom_level(c));
    int rc;
    // a half sweep of level l of every case whose V-cycle goes below l: one
    // launch for the cases that sweep it with two-pair threads, one for the
    // others (sweep_shape)
    auto half = [&](int l, int color, int zero, int dot_slot, int res) {
        std::vector<RbArg<T>> args[2];
        BOp o[2];
        // per cell of level 0: the half's b and the other colour read, the
        // half's x written (the first sweep reads no x); the residual
        // variants also read the half's own x or write its r; the mixed
        // form's r.z variant also reads the FP64 r and writes the FP64 z
        const double per = dot_slot >= 0 ? 1.5 * sizeof(T) + (mixed ? 8.0 : 0.0)
                           : res          ? 2.0 * sizeof(T)
                           : zero         ? 1.0 * sizeof(T)
                                          : 1.5 * sizeof(T);
        for (ibmdem_ctx *c : cs) {
            if (l >= bottom_level(c)) continue;
            MGLevel<T> &L = levels<T>(c)[l];
            const SweepShape sw = sweep_shape(L.d);
            RbArg<T> a{};
            a.g = bgrid(sw.grid, sw.block);
            a.L = c->dlev + l;
            a.x = L.x;
            a.b = L.b;
            a.stop = c->scal + SC_CONV;
            a.partial = c->partial;
            a.pofs = dot_slot >= 0 ? dot_slot * (long)sw.grid.x * sw.grid.y * sw.grid.z : 0;
            a.r = res ? L.r : nullptr;
            a.r64 = dot_slot >= 0 && mixed ? c->cg_r : nullptr;
            a.z64 = dot_slot >= 0 && mixed ? c->cg_z : nullptr;
            args[sw.x2].push_back(a);
            if (l == 0) o[sw.x2].bytes += per * (double)c->n;
        }
        for (int v = 0; v < 2; ++v) {
            o[v].color = color;
            o[v].kind = l == 0 ? IBMDEM_BK_SMOOTH0 : IBMDEM_BK_SMOOTH;
        }
        const BLaunch fn[2] = {
            dot_slot >= 0 ? launch_rb<T, true, 0, false, false>
            : res == 2    ? launch_rb<T, false, 2, false, false>
            : res == 1    ? launch_rb<T, false, 1, false, false>
            : zero        ? launch_rb<T, false, 0, true, false>
                          : launch_rb<T, false, 0, false, false>,
            dot_slot >= 0 ? launch_rb<T, true, 0, false, true>
            : res == 2    ? launch_rb<T, false, 2, false, true>
            : res == 1    ? launch_rb<T, false, 1, false, true>
            : zero        ? launch_rb<T, false, 0, true, true>
                          : launch_rb<T, false, 0, false, true>};
        int r2 = add_op(ar, ops, fn[1], RB, args[1], o[1], err);
        return r2 ? r2 : add_op(ar, ops, fn[0], RB, args[0], o[0], err);
    };
    // the bottoms at level l
    auto bottoms = [&](int l) {
        std::vector<BtArg<T>> tail, lds, glob;
        for (ibmdem_ctx *c : cs) {
            if (bottom_level(c) != l) continue;
            MGLevel<T> &L = levels<T>(c)[l];
            BtArg<T> a{};
            a.lev = c->dlev;
            a.first = l;
            a.last = c->nlev - 1;
            a.pre = c->cs.mg_pre;
            a.post = c->cs.mg_post;
            a.sweeps = c->cs.mg_coarse_sweeps;
            a.x = L.x;
            a.b = L.b;
            a.tmp = L.r;
            a.stop = c->scal + SC_CONV;
            if (l >= c->tail_first) {
                a.g = bgrid(dim3(1), dim3(TAIL_T));
                tail.push_back(a);
            } else if (L.n <= COARSE_LDS) {
                a.g = bgrid(dim3(1), dim3(64));
                lds.push_back(a);
            } else {
                a.g = bgrid(dim3(1), dim3(COARSE_T));
                glob.push_back(a);
            }
        }
        int r;
        BOp o;
        o.kind = IBMDEM_BK_BOTTOM;
        if ((r = add_op(ar, ops, launch_tail<T>, TAIL_T, tail, o, err))) return r;
        if ((r = add_op(ar, ops, launch_coarse_lds<T>, 64, lds, o, err))) return r;
        return add_op(ar, ops, launch_coarse<T>, COARSE_T, glob, o, err);
    };
    for (int l = 0; l < depth; ++l) {
        // below level 0 every case that sweeps level l had its first sweep
        // done by the restriction into l
        for (int q = 0; q < pre; ++q) {
            if ((q > 0 || l == 0) && (rc = half(l, 0, q == 0, -1, 0))) return rc;
            if ((rc = half(l, 1, 0, -1, q == pre - 1 ? 1 : 0))) return rc;
        }
        if ((rc = half(l, 0, 0, -1, 2))) return rc;
        std::vector<RsArg<T>> rs;
        for (ibmdem_ctx *c : cs) {
            if (l >= bottom_level(c)) continue;
            MGLevel<T> &F = levels<T>(c)[l], &C = levels<T>(c)[l + 1];
            RsArg<T> a{};
            a.g = bgrid(dim3(div_up(C.d.nx, RS_CX), div_up(C.d.ny, RS_CY), C.d.nz), dim3(RB));
            a.F = c->dlev + l;
            a.C = c->dlev + l + 1;
            a.rf = F.r;
            a.bc = C.b;
            a.xc = l + 1 < bottom_level(c) ? C.x : nullptr;
            a.stop = c->scal + SC_CONV;
            rs.push_back(a);
        }
        BOp ors;
        ors.kind = IBMDEM_BK_RESTRICT;
        // the fine residual read, the coarse right-hand side (and where the
        // first coarse sweep is fused, the coarse x) written
        for (ibmdem_ctx *c : cs)
            if (l < bottom_level(c))
                ors.bytes += sizeof(T) * (double)(levels<T>(c)[l].n + levels<T>(c)[l + 1].n *
                                                  (l + 1 < bottom_level(c) ? 2 : 1));
        if ((rc = add_op(ar, ops, launch_restrict<T>, RB, rs, ors, err))) return rc;
        if ((rc = bottoms(l + 1))) return rc;
    }
    for (int l = depth - 1; l >= 0; --l) {
        std::vector<PrArg<T>> pr[2];
        BOp opr[2];
        for (ibmdem_ctx *c : cs) {
            if (l >= bottom_level(c)) continue;
            MGLevel<T> &F = levels<T>(c)[l], &C = levels<T>(c)[l + 1];
            const ProlongShape sh = prolong_shape(F.d);
            if (!sh.ok) {
                err = "prolongation tile does not fit";
                return IBMDEM_ERR_ARG;
            }
            PrArg<T> a{};
            a.g = bgrid(sh.grid, sh.block);
            a.F = c->dlev + l;
            a.C = c->dlev + l + 1;
            a.xf = F.x;
            a.xc = C.x;
            a.stop = c->scal + SC_CONV;
            pr[sh.x2].push_back(a);
            // the coarse solution read, the fine one read and written
            opr[sh.x2].bytes += sizeof(T) * (double)(2 * F.n + C.n);
        }
        opr[0].kind = opr[1].kind = IBMDEM_BK_PROLONG;
        if ((rc = add_op(ar, ops, launch_prolong2<T>, RB, pr[1], opr[1], err))) return rc;
        if ((rc = add_op(ar, ops, launch_prolong<T>, RB, pr[0], opr[0], err))) return rc;
        for (int q = 0; q < post; ++q) {
            const bool last = l == 0 && q == post - 1;
            if ((rc = half(l, 1, 0, last ? 0 : -1, 0))) return rc;
            if ((rc = half(l, 0, 0, last ? 1 : -1, 0))) return rc;
        }
    }
    std::vector<CgArg> fin;
    for (ibmdem_ctx *c : cs) {
        const dim3 g = sweep_shape(c->lev64[0].d).grid;   // the r.z sweeps' blocks
        CgArg a{};
        a.g = bgrid(dim3(1), dim3(FIN_T));
        a.partial = c->partial;
        a.scal = c->scal;
        a.np = 2 * (int)(g.x * g.y * g.z);
        a.rtol = c->cs.poisson_rtol;
        fin.push_back(a);
    }
    BOp o;
    o.slot = slot;
    o.op = op;
    o.kind = IBMDEM_BK_SCALAR;
    return add_op(ar, ops, launch_final, FIN_T, fin, o, err);
}

// the CG vectors of every case (the mixed form's FP32 copy of r with them)
CgArg cg_arg(ibmdem_ctx *c, dim3 grid, dim3 block, int np) {
    CgArg a{};
    a.g = bgrid(grid, block);
    a.L = c->dlev;
    a.n = c->n;
    a.p = c->cg_p; a.z = c->cg_z; a.x = c->cg_x; a.r = c->cg_r; a.Ap = c->cg_Ap;
    a.partial = c->partial;
    a.scal = c->scal;
    a.r32 = c->cs.precision == IBMDEM_MIXED ? c->lev32[0].b : nullptr;
    a.np = np;
    a.rtol = c->cs.poisson_rtol;
    return a;
}

template <typename T>
int plan_all(Arena &ar, std::vector<BOp> &begin, std::vector<BOp> &iter, const std::vector<ibmdem_ctx *> &cs,
             std::string &err) {
    int rc;
    // poisson_begin_precondition: z = M^-1 r with r.z, p = z
    if ((rc = plan_precondition_dot<T>(ar, begin, cs, SC_RZ, 0, err))) return rc;
    std::vector<CgArg> cp, px, mv[2], fpap, up, frr;
    double mv_cells[2] = {0.0, 0.0};
    for (ibmdem_ctx *c : cs) {
        const SweepShape sh = sweep_shape(c->lev64[0].d);
        const int mvb = (int)(sh.grid.x * sh.grid.y * sh.grid.z);
        cp.push_back(cg_arg(c, dim3(stream_blocks(c->n)), dim3(RED_T), 0));
        px.push_back(cg_arg(c, dim3(stream_blocks(c->n)), dim3(RED_T), 0));
        mv[sh.x2].push_back(cg_arg(c, sh.grid, sh.block, 0));
        mv_cells[sh.x2] += (double)c->n;
        fpap.push_back(cg_arg(c, dim3(1), dim3(FIN_T), mvb));
        up.push_back(cg_arg(c, dim3(RED_BLOCKS), dim3(RED_T), 0));
        frr.push_back(cg_arg(c, dim3(1), dim3(FIN_T), RED_BLOCKS));
    }
    double cells = 0.0;
    for (ibmdem_ctx *c : cs) cells += (double)c->n;
    const bool mixed = !std::is_same<T, double>::value;
    BOp o;
    o.kind = IBMDEM_BK_VECTOR;
    o.bytes = 16.0 * cells;   // z read, p written
    if ((rc = add_op(ar, begin, launch_copy, RED_T, cp, o, err))) return rc;
    // cg_iteration
    o.skip_first = true;
    o.bytes = 40.0 * cells;   // p, x read and written, z read
    if ((rc = add_op(ar, iter, launch_px, RED_T, px, o, err))) return rc;
    o = BOp();
    o.kind = IBMDEM_BK_MATVEC;
    o.bytes = 16.0 * mv_cells[1];   // p read, Ap written
    if ((rc = add_op(ar, iter, launch_matvec2, RB, mv[1], o, err))) return rc;
    o.bytes = 16.0 * mv_cells[0];
    if ((rc = add_op(ar, iter, launch_matvec, RB, mv[0], o, err))) return rc;
    o = BOp();
    o.kind = IBMDEM_BK_SCALAR;
    o.slot = SC_PAP; o.op = 1; o.with_it = true;
    if ((rc = add_op(ar, iter, launch_final, FIN_T, fpap, o, err))) return rc;
    BOp ou;
    ou.kind = IBMDEM_BK_VECTOR;
    ou.bytes = (24.0 + (mixed ? 4.0 : 0.0)) * cells;   // r read and written, Ap read (FP32 r written)
    if ((rc = add_op(ar, iter, launch_update_r, RED_T, up, ou, err))) return rc;
    o.slot = SC_RR; o.op = 3;
    if ((rc = add_op(ar, iter, launch_final, FIN_T, frr, o, err))) return rc;
    return plan_precondition_dot<T>(ar, iter, cs, SC_RZ_NEW, 2, err);
}

int prof_event(BProf &pr, int i, hipStream_t s, std::string &err) {
    while ((int)pr.ev.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            err = "profiling event creation failed";
            return IBMDEM_ERR_HIP;
        }
        pr.ev.push_back(e);
    }
    if (hipEventRecord(pr.ev[i], s) != hipSuccess) {
        err = "profiling event record failed";
        return IBMDEM_ERR_HIP;
    }
    return IBMDEM_OK;
}

int run_ops(const std::vector<BOp> &ops, hipStream_t s, int it, BProf &pr, std::string &err) {
    const bool rec = pr.on && pr.armed;
    int rc;
    if (rec && pr.n == 0 && (rc = prof_event(pr, 0, s, err))) return rc;
    for (const BOp &o : ops) {
        if (o.skip_first && it == 1) continue;
        o.fn(o, s, it);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            err = std::string("batched solver launch: ") + hipGetErrorString(e);
            return IBMDEM_ERR_HIP;
        }
        if (rec) {
            if ((rc = prof_event(pr, pr.n + 1, s, err))) return rc;
            pr.kind.resize(pr.n + 1);
            pr.bytes.resize(pr.n + 1);
            pr.kind[pr.n] = o.kind;
            pr.bytes[pr.n] = o.bytes;
            pr.n++;
        }
    }
    return IBMDEM_OK;
}

}  // namespace

struct PoissonBatch {
    std::vector<ibmdem_ctx *> cs;
    std::vector<BOp> begin, iter;
    BProf prof;
    char *arena = nullptr;             // device copy of the tables
    const double **scal_ptrs = nullptr;   // in the arena
    double *gather_dev = nullptr;
    double *gather_host = nullptr;     // pinned
};

bool poisson_batchable(const ibmdem_ctx *c) {
    // the V-cycle must leave r.z with its last sweeps (precondition_dot)
    return c->nlev >= 2 && c->cs.mg_post >= 1;
}

void poisson_batch_destroy(PoissonBatch *pb) {
    if (!pb) return;
    for (hipEvent_t e : pb->prof.ev) (void)hipEventDestroy(e);
    if (pb->arena) (void)hipFree(pb->arena);
    if (pb->gather_dev) (void)hipFree(pb->gather_dev);
    if (pb->gather_host) (void)hipHostFree(pb->gather_host);
    delete pb;
}

int poisson_batch_create(const std::vector<ibmdem_ctx *> &cs, PoissonBatch **out, std::string &err) {
    *out = nullptr;
    if (cs.empty()) {
        err = "empty batch group";
        return IBMDEM_ERR_ARG;
    }
    const ibmdem_ctx *c0 = cs[0];
    for (const ibmdem_ctx *c : cs) {
        if (!poisson_batchable(c) || c->cs.precision != c0->cs.precision || c->cs.mg_pre != c0->cs.mg_pre ||
            c->cs.mg_post != c0->cs.mg_post || c->cs.poisson_max_iter != c0->cs.poisson_max_iter) {
            err = "cases of a batched solve must share precision, smoothing counts and iteration cap";
            return IBMDEM_ERR_ARG;
        }
    }
    PoissonBatch *pb = new (std::nothrow) PoissonBatch();
    if (!pb) return IBMDEM_ERR_ALLOC;
    pb->cs = cs;
    // the FP64 V-cycle's level 0 is the CG vectors (as precondition sets it)
    for (ibmdem_ctx *c : cs) {
        c->lev64[0].b = c->cg_r;
        c->lev64[0].x = c->cg_z;
    }
    Arena ar;
    int rc = c0->cs.precision == IBMDEM_MIXED ? plan_all<float>(ar, pb->begin, pb->iter, cs, err)
                                              : plan_all<double>(ar, pb->begin, pb->iter, cs, err);
    std::vector<const double *> sp;
    for (ibmdem_ctx *c : cs) sp.push_back(c->scal);
    const size_t sp_off = ar.put(sp.data(), sp.size() * sizeof(double *));
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMalloc((void **)&pb->arena, ar.bytes.size());
        if (e == hipSuccess) e = hipMemcpy(pb->arena, ar.bytes.data(), ar.bytes.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void **)&pb->gather_dev, cs.size() * SC_COUNT * sizeof(double));
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&pb->gather_host, cs.size() * SC_COUNT * sizeof(double), hipHostMallocDefault);
        if (e != hipSuccess) {
            err = std::string("batched solver set-up: ") + hipGetErrorString(e);
            rc = IBMDEM_ERR_HIP;
        }
    }
    if (rc) {
        poisson_batch_destroy(pb);
        return rc;
    }
    for (auto *ops : {&pb->begin, &pb->iter})
        for (BOp &o : *ops) {
            o.first = (const int *)(pb->arena + o.first_off);
            o.arg = pb->arena + o.arg_off;
        }
    pb->scal_ptrs = (const double **)(pb->arena + sp_off);
    *out = pb;
    return IBMDEM_OK;
}

int poisson_batch_begin(PoissonBatch *pb, hipStream_t s, std::string &err) {
    return run_ops(pb->begin, s, 0, pb->prof, err);
}

int poisson_batch_iteration(PoissonBatch *pb, hipStream_t s, int it, std::string &err) {
    return run_ops(pb->iter, s, it, pb->prof, err);
}

void poisson_batch_set_profile(PoissonBatch *pb, bool on) { pb->prof.on = on; }
bool poisson_batch_profiling(const PoissonBatch *pb) { return pb->prof.on; }

void poisson_batch_profile_arm(PoissonBatch *pb, bool armed) {
    pb->prof.armed = armed;
    if (armed) pb->prof.n = 0;
}

int poisson_batch_profile_collect(PoissonBatch *pb, double *ms, int64_t *launches, double *bytes) {
    const BProf &pr = pb->prof;
    if (!pr.on) return IBMDEM_OK;
    for (int i = 0; i < pr.n; ++i) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, pr.ev[i], pr.ev[i + 1]) != hipSuccess) return IBMDEM_ERR_HIP;
        const int k = pr.kind[i];
        if (k < 0 || k >= BK_COUNT) continue;
        ms[k] += t;
        launches[k] += 1;
        bytes[k] += pr.bytes[i];
    }
    return IBMDEM_OK;
}

int poisson_batch_gather(PoissonBatch *pb, hipStream_t s, std::string &err) {
    const int n = (int)pb->cs.size();
    hipLaunchKernelGGL(kb_gather_scalars, dim3(n), dim3(64), 0, s, pb->scal_ptrs, n, pb->gather_dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(pb->gather_host, pb->gather_dev, (size_t)n * SC_COUNT * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) {
        err = std::string("batched solver scalars: ") + hipGetErrorString(e);
        return IBMDEM_ERR_HIP;
    }
    return IBMDEM_OK;
}

const double *poisson_batch_scalars(const PoissonBatch *pb, int q) { return pb->gather_host + (size_t)q * SC_COUNT; }

int poisson_batch_launches(const PoissonBatch *pb, int k) {
    int n = (int)pb->begin.size();
    for (int it = 1; it <= k; ++it)
        for (const BOp &o : pb->iter) n += !(o.skip_first && it == 1);
    return n;
}

// poisson_gauge of every context of a group as one launch
namespace {
struct GaugeArg {
    BGrid g;
    const LevelDev *L;
    const double *x, *p, *scal;
    double *pc;
};
__global__ __launch_bounds__(RB) void kb_gauge(BTable<GaugeArg> tab) {
    batch_for(tab, [&](const GaugeArg &a, const Blk &kb) {
        gauge_body(kb, *cst(a.L), glb(a.x), glb(a.p), glb(a.scal), glb(a.pc));
    });
}
void launch_gauge(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_gauge, dim3(o.blocks), dim3(RB), 0, s, table<GaugeArg>(o));
}
}  // namespace

int poisson_plan_gauge(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<GaugeArg> args;
    for (ibmdem_ctx *c : cs) {
        const LevelDev &L0 = c->lev64[0].d;
        const LaunchShape fs = shape_for(div_up(L0.nx, 2), L0.ny, L0.nz);
        args.push_back(GaugeArg{bgrid(fs.grid, fs.block), c->dlev, c->cg_x, c->cg_p, c->scal, c->pcorr});
    }
    return add_op(P.ar, P.ops, launch_gauge, RB, args, BOp(), err);
}

// poisson_begin_rhs of every context of a group: the pinned row's sum (the
// contexts without a Dirichlet side), r = b and x = 0, then ||b||
namespace {
__global__ __launch_bounds__(RB) void kb_cg_init(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        cg_init_body(kb, *cst(a.L), glb(a.rhs), glb(a.r), glb(a.x), glb(a.scal), glb(a.r32));
    });
}
__global__ __launch_bounds__(RED_T) void kb_dot_rr(BTable<CgArg> tab) {
    batch_for(tab, [&](const CgArg &a, const Blk &kb) {
        dot_body(kb, a.n, glb((const double *)a.r), glb((const double *)a.r), glb(a.partial), glb(a.scal));
    });
}
void launch_cg_init(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_cg_init, dim3(o.blocks), dim3(RB), 0, s, table<CgArg>(o));
}
void launch_dot_rr(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_dot_rr, dim3(o.blocks), dim3(RED_T), 0, s, table<CgArg>(o));
}
}  // namespace

int poisson_plan_rhs(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<ReduceJob> pin;
    std::vector<CgArg> init, dot, fin;
    for (ibmdem_ctx *c : cs) {
        const LevelDev &L0 = c->lev64[0].d;
        if (!L0.dirichlet) pin.push_back(ReduceJob{c->rhs, nullptr, c->n, 1, c->partial, c->scal + SC_PIN});
        const LaunchShape fs = shape_for(div_up(L0.nx, 2), L0.ny, L0.nz);
        CgArg a = cg_arg(c, fs.grid, fs.block, 0);
        a.rhs = c->rhs;
        init.push_back(a);
        dot.push_back(cg_arg(c, dim3(RED_BLOCKS), dim3(RED_T), 0));
        fin.push_back(cg_arg(c, dim3(1), dim3(FIN_T), RED_BLOCKS));
    }
    int rc;
    if ((rc = plan_reduce_sum(P, pin, err))) return rc;
    if ((rc = add_op(P.ar, P.ops, launch_cg_init, RB, init, BOp(), err))) return rc;
    if ((rc = add_op(P.ar, P.ops, launch_dot_rr, RED_T, dot, BOp(), err))) return rc;
    BOp o;
    o.slot = SC_BNORM2;
    o.op = 4;
    return add_op(P.ar, P.ops, launch_final, FIN_T, fin, o, err);
}
