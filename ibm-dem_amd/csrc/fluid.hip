// fluid.hip — staggered-grid predictor, bulk forcing, divergence and
// projection (fluid.F90:2-250) as HBM-bound stencil kernels.
//
// Layout: interior-only FP64 arrays, x fastest. A thread owns one cell and
// all three velocity components of it. Blocks are 64 x 4 cells in x-y and
// one z plane; the block index is remapped so that each of the 8 XCDs walks
// a contiguous slab of z planes (neighbouring planes share its L2).
#include "common.hpp"
#include "batched.hpp"

#include <map>
#include <type_traits>

#include <cmath>

namespace {

using namespace bk;

constexpr int BX = 64, BY = 4;
constexpr int RED_BLOCKS = 1024;   // fixed: reductions never depend on timing
constexpr int RED_THREADS = 256;

// XCD-aware block order: hardware dispatches block b to XCD b % 8; give XCD
// x the planes [x*nz/8, (x+1)*nz/8) so its L2 sees neighbouring planes.
__device__ __forceinline__ void xcd_block(int &bx, int &by, int &bz) {
    const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const long nb = (long)gx * gy * gz;
    long b = (long)blockIdx.x + (long)gx * (blockIdx.y + (long)gy * blockIdx.z);
    long per = nb / 8;
    if (per > 0 && nb % 8 == 0) b = (b % 8) * per + b / 8;
    bx = (int)(b % gx);
    by = (int)((b / gx) % gy);
    bz = (int)(b / ((long)gx * gy));
}
// the hardware block with xcd_block's coordinates
__device__ __forceinline__ Blk xcd_blk() {
    Blk kb = hw_blk();
    xcd_block(kb.bx, kb.by, kb.bz);
    return kb;
}

using Map3 = AxisMap<-1, 3>;
template <int C, bool OPEN>
struct Tap {
    const double *__restrict__ p;
    const GridDev *g;
    const Map3 *mx, *my, *mz;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return vel_tap<C, OPEN>(p, *g, *mx, *my, *mz, di + 1, dj + 1, dk + 1);
    }
};
template <bool OPEN>
struct TapP {
    const double *__restrict__ p;
    const GridDev *g;
    const Map3 *mx, *my, *mz;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return scal_tap<OPEN>(p, *g, *mx, *my, *mz, di + 1, dj + 1, dk + 1);
    }
};

struct PredParams {
    double dt, visc, rhof, a, c;
    double cp[3];
    double idx, idy, idz, idx2, idy2, idz2, irhof;   // exact reciprocals (used when P2)
};

// ConvectionTensor + ViscousTensor + PressureGradient + VelocityGuess for one
// cell (fluid.F90:15-44, 58-66, 80-82, 118-126), the reference's operand
// order, in compute type CT. Returns the three increments
// a*RHS_noF - c*gradP (without the IBM force, already inside u~).
template <bool P2, typename CT>
__device__ __forceinline__ CT pdiv(CT x, double d, double inv) {
    return P2 ? x * CT(inv) : x / CT(d);
}

template <typename CT, bool P2, class LU, class LV, class LW, class LP>
__device__ __forceinline__ void predictor_cell(const LU &u, const LV &v, const LW &w, const LP &P,
                                               const GridDev &g, const PredParams &pp, CT R[3], CT Pg[3]) {
    const CT h = CT(0.5);
    const CT visc = CT(pp.visc);
    const CT u0 = CT(u(0, 0, 0)), uE = CT(u(1, 0, 0)), uW = CT(u(-1, 0, 0)), uN = CT(u(0, 1, 0)),
             uS = CT(u(0, -1, 0)), uT = CT(u(0, 0, 1)), uB = CT(u(0, 0, -1));
    const CT v0 = CT(v(0, 0, 0)), vE = CT(v(1, 0, 0)), vW = CT(v(-1, 0, 0)), vN = CT(v(0, 1, 0)),
             vS = CT(v(0, -1, 0)), vT = CT(v(0, 0, 1)), vB = CT(v(0, 0, -1));
    const CT w0 = CT(w(0, 0, 0)), wE = CT(w(1, 0, 0)), wW = CT(w(-1, 0, 0)), wN = CT(w(0, 1, 0)),
             wS = CT(w(0, -1, 0)), wT = CT(w(0, 0, 1)), wB = CT(w(0, 0, -1));
    // the diagonal taps of the flux averages
    const CT v_wn = CT(v(-1, 1, 0)), w_wt = CT(w(-1, 0, 1));   // for u
    const CT u_es = CT(u(1, -1, 0)), w_st = CT(w(0, -1, 1));   // for v
    const CT u_eb = CT(u(1, 0, -1)), v_nb = CT(v(0, 1, -1));   // for w
    const CT p0 = CT(P(0, 0, 0)), pW = CT(P(-1, 0, 0)), pS = CT(P(0, -1, 0)), pB = CT(P(0, 0, -1));

    CT uuE = h * (uE + u0) * h * (uE + u0);
    CT uuW = h * (uW + u0) * h * (uW + u0);
    CT uvB = h * (uN + u0) * h * (v_wn + vN);
    CT uvF = h * (uS + u0) * h * (vW + v0);
    CT uwN = h * (uT + u0) * h * (w_wt + wT);
    CT uwS = h * (uB + u0) * h * (wW + w0);

    CT vuE = h * (vE + v0) * h * (u_es + uE);
    CT vuW = h * (vW + v0) * h * (uS + u0);
    CT vvB = h * (vN + v0) * h * (vN + v0);
    CT vvF = h * (vS + v0) * h * (vS + v0);
    CT vwN = h * (vT + v0) * h * (w_st + wT);
    CT vwS = h * (vB + v0) * h * (wS + w0);

    CT wuE = h * (wE + w0) * h * (u_eb + uE);
    CT wuW = h * (wW + w0) * h * (uB + u0);
    CT wvB = h * (wN + w0) * h * (v_nb + vN);
    CT wvF = h * (wS + w0) * h * (vB + v0);
    CT wwN = h * (wT + w0) * h * (wT + w0);
    CT wwS = h * (wB + w0) * h * (wB + w0);

    const double gdx = g.dx, gdy = g.dy, gdz = g.dz;
    CT TIux = pdiv<P2>(uuE - uuW, gdx, pp.idx), TIuy = pdiv<P2>(uvB - uvF, gdy, pp.idy), TIuz = pdiv<P2>(uwN - uwS, gdz, pp.idz);
    CT TIvx = pdiv<P2>(vuE - vuW, gdx, pp.idx), TIvy = pdiv<P2>(vvB - vvF, gdy, pp.idy), TIvz = pdiv<P2>(vwN - vwS, gdz, pp.idz);
    CT TIwx = pdiv<P2>(wuE - wuW, gdx, pp.idx), TIwy = pdiv<P2>(wvB - wvF, gdy, pp.idy), TIwz = pdiv<P2>(wwN - wwS, gdz, pp.idz);

    const double dx2 = gdx * gdx, dy2 = gdy * gdy, dz2 = gdz * gdz;
    CT TVux = pdiv<P2>(uE + uW - CT(2) * u0, dx2, pp.idx2) * visc;
    CT TVuy = pdiv<P2>(uN + uS - CT(2) * u0, dy2, pp.idy2) * visc;
    CT TVuz = pdiv<P2>(uT + uB - CT(2) * u0, dz2, pp.idz2) * visc;
    CT TVvx = pdiv<P2>(vE + vW - CT(2) * v0, dx2, pp.idx2) * visc;
    CT TVvy = pdiv<P2>(vN + vS - CT(2) * v0, dy2, pp.idy2) * visc;
    CT TVvz = pdiv<P2>(vT + vB - CT(2) * v0, dz2, pp.idz2) * visc;
    CT TVwx = pdiv<P2>(wE + wW - CT(2) * w0, dx2, pp.idx2) * visc;
    CT TVwy = pdiv<P2>(wN + wS - CT(2) * w0, dy2, pp.idy2) * visc;
    CT TVwz = pdiv<P2>(wT + wB - CT(2) * w0, dz2, pp.idz2) * visc;

    CT Px = pdiv<P2>(pdiv<P2>(p0 - pW, gdx, pp.idx), pp.rhof, pp.irhof);
    CT Py = pdiv<P2>(pdiv<P2>(p0 - pS, gdy, pp.idy), pp.rhof, pp.irhof);
    CT Pz = pdiv<P2>(pdiv<P2>(p0 - pB, gdz, pp.idz), pp.rhof, pp.irhof);

    R[0] = -(TIux + TIuy + TIuz) + (TVux + TVuy + TVuz) + CT(pp.cp[0]);
    R[1] = -(TIvx + TIvy + TIvz) + (TVvx + TVvy + TVvz) + CT(pp.cp[1]);
    R[2] = -(TIwx + TIwy + TIwz) + (TVwx + TVwy + TVwz) + CT(pp.cp[2]);
    Pg[0] = Px; Pg[1] = Py; Pg[2] = Pz;
}

// Bulk-forcing partial sums with a fixed association, whatever the
// predictor schedule: each 64-cell x row is summed by a butterfly over its
// wavefront (lane 0's value), and each group of four rows (j = 4g+1..4g+4 at
// one plane) by ((r0 + r1) + (r2 + r3)); partial[3 q + c], q = ((k-1) *
// ceil(ny/4) + g) * ceil(nx/64) + i-tile.
__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// rows of this thread's wave into red[row][c] (row = kb.ty)
__device__ __forceinline__ void row_sums(const Blk &kb, const double s[3], double (*red)[3]) {
    double r[3];
    for (int q = 0; q < 3; ++q) r[q] = wave_sum64(s[q]);
    if (kb.tx == 0)
        for (int q = 0; q < 3; ++q) red[kb.ty][q] = r[q];
}
__device__ __forceinline__ void group_partial(double (*red)[3], int g0, double *part) {
    for (int q = 0; q < 3; ++q) part[q] = (red[g0][q] + red[g0 + 1][q]) + (red[g0 + 2][q] + red[g0 + 3][q]);
}
__device__ __forceinline__ long group_index(const GridDev &g, int bx, int jg, int k) {
    return ((long)(k - 1) * ((g.ny + 3) / 4) + jg) * ((g.nx + BX - 1) / BX) + bx;
}

// VelocityGuess fused form: u* = u~ + dt*(a*RHS_noF(u~) - c*gradP_old).
// Because u~ = uold + dt*f after the multi-direct-forcing loop, this equals
// the reference's uold + dt*(a*RHS(u~) + 0*RHSl - c*gradP) for a = 1.
template <typename CT, bool BULK, bool P2, bool OPEN>
__device__ __forceinline__ void predictor_body(const Blk &kb, const GridDev &g, const double *__restrict__ u,
                                               const double *__restrict__ v, const double *__restrict__ w,
                                               const double *__restrict__ P, double *__restrict__ un,
                                               double *__restrict__ vn, double *__restrict__ wn,
                                               const PredParams &pp, double *__restrict__ partial) {
    const int bx = kb.bx, by = kb.by, bz = kb.bz;
    const int i = bx * BX + kb.tx + 1, j = by * BY + kb.ty + 1, k = bz + 1;
    double s[3] = {0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const long o = cidx(g, i, j, k);
        CT R[3], Pg[3], inc[3];
        Map3 mx, my, mz;
        mx.init(i, g.nx, g.bt[0], g.bt[1]);
        my.init(j, g.ny, g.bt[2], g.bt[3]);
        mz.init(k, g.nz, g.bt[4], g.bt[5]);
        predictor_cell<CT, P2>(Tap<0, OPEN>{u, &g, &mx, &my, &mz}, Tap<1, OPEN>{v, &g, &mx, &my, &mz},
                               Tap<2, OPEN>{w, &g, &mx, &my, &mz}, TapP<OPEN>{P, &g, &mx, &my, &mz}, g, pp, R, Pg);
        const CT a = CT(pp.a), c = CT(pp.c);
        for (int q = 0; q < 3; ++q) inc[q] = a * R[q] - c * Pg[q];
        const double dt = pp.dt;
        double nu_ = wall_face<0>(g, i, j, k) ? face_value<0>(g) : u[o] + dt * (double)inc[0];
        double nv_ = wall_face<1>(g, i, j, k) ? face_value<1>(g) : v[o] + dt * (double)inc[1];
        double nw_ = wall_face<2>(g, i, j, k) ? face_value<2>(g) : w[o] + dt * (double)inc[2];
        un[o] = nu_;
        vn[o] = nv_;
        wn[o] = nw_;
        s[0] = nu_; s[1] = nv_; s[2] = nw_;
    }
    if (BULK) {
        __shared__ double red[BY][3];
        row_sums(kb, s, red);
        __syncthreads();
        if (kb.tx == 0 && kb.ty == 0) group_partial(red, 0, partial + 3 * group_index(g, bx, by, bz + 1));
    }
}
template <typename CT, bool BULK, bool P2, bool OPEN>
__global__ __launch_bounds__(BX *BY) void k_predictor(GridDev g, const double *__restrict__ u,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ P, double *__restrict__ un,
                                                      double *__restrict__ vn, double *__restrict__ wn,
                                                      PredParams pp, double *__restrict__ partial) {
    predictor_body<CT, BULK, P2, OPEN>(xcd_blk(), g, u, v, w, P, un, vn, wn, pp, partial);
}

// VelocityGuess(a,b,c) in the reference's own form (fluid.F90:116-131):
// RHS = -sum TI + sum TV + CP + f; u = uold + dt*(a*RHS + b*RHSl - c*gradP);
// RHSl <- RHS. Used by contexts created for the low-storage RK3 scheme, which
// keep uold, RHSl and the full-field IBM force.
struct GeneralArgs {
    const double *uold[3];
    const double *f[3];
    double *rhsl[3];
    double b;
};
template <typename CT, bool BULK, bool P2, bool OPEN>
__global__ __launch_bounds__(BX *BY) void k_predictor_general(GridDev g, const double *__restrict__ u,
                                                              const double *__restrict__ v,
                                                              const double *__restrict__ w,
                                                              const double *__restrict__ P, double *__restrict__ un,
                                                              double *__restrict__ vn, double *__restrict__ wn,
                                                              PredParams pp, GeneralArgs ga,
                                                              double *__restrict__ partial) {
    const Blk kb = xcd_blk();
    const int bx = kb.bx, by = kb.by, bz = kb.bz;
    const int i = bx * BX + kb.tx + 1, j = by * BY + kb.ty + 1, k = bz + 1;
    double s[3] = {0.0, 0.0, 0.0};
    if (i <= g.nx && j <= g.ny) {
        const long o = cidx(g, i, j, k);
        CT R[3], Pg[3];
        Map3 mx, my, mz;
        mx.init(i, g.nx, g.bt[0], g.bt[1]);
        my.init(j, g.ny, g.bt[2], g.bt[3]);
        mz.init(k, g.nz, g.bt[4], g.bt[5]);
        predictor_cell<CT, P2>(Tap<0, OPEN>{u, &g, &mx, &my, &mz}, Tap<1, OPEN>{v, &g, &mx, &my, &mz},
                               Tap<2, OPEN>{w, &g, &mx, &my, &mz}, TapP<OPEN>{P, &g, &mx, &my, &mz}, g, pp, R, Pg);
        const CT a = CT(pp.a), b = CT(ga.b), c = CT(pp.c);
        const bool face[3] = {wall_face<0>(g, i, j, k), wall_face<1>(g, i, j, k), wall_face<2>(g, i, j, k)};
        const double fval[3] = {face_value<0>(g), face_value<1>(g), face_value<2>(g)};
        double *out[3] = {un, vn, wn};
        for (int q = 0; q < 3; ++q) {
            const CT rhs = R[q] + CT(ga.f[q][o]);
            const CT val = a * rhs + b * CT(ga.rhsl[q][o]) - c * Pg[q];
            const double nv = face[q] ? fval[q] : ga.uold[q][o] + pp.dt * (double)val;
            out[q][o] = nv;
            ga.rhsl[q][o] = (double)rhs;
            s[q] = nv;
        }
    }
    if (BULK) {
        __shared__ double red[BY][3];
        row_sums(kb, s, red);
        __syncthreads();
        if (kb.tx == 0 && kb.ty == 0) group_partial(red, 0, partial + 3 * group_index(g, bx, by, bz + 1));
    }
}

// ---------------------------------------------------------------------------
// LDS z-marching predictor. A block of 64 x TY threads owns a 64 x TY cell
// column tile and marches it through KZ planes. The three planes k-1, k,
// k+1 of u, v, w (with a one-cell x/y halo) and plane k of P sit in LDS in
// rotating slots, each element loaded once per block through the ghost
// rules (vel_at / scal_at: the same values the direct kernel's taps read),
// so every stencil tap is an LDS read and each field element crosses HBM
// about (66 (TY+2) / (64 TY)) (KZ+2)/KZ times instead of the direct
// kernel's plane re-reads. The cell arithmetic is predictor_cell's, so the
// result is the direct kernel's bit for bit.
// ---------------------------------------------------------------------------
constexpr int LX = BX + 2;   // tile row with halo

template <typename LT, int TY>
struct LdsPlanes {
    LT v[3][3][TY + 2][LX];   // [component][slot][y][x]
    LT p[TY + 2][LX];         // P, plane k
};

template <typename LT, int TY, int C>
struct LTap {
    const LdsPlanes<LT, TY> *L;
    int sm, s0, sp, tx, ty;   // slots of planes k-1, k, k+1
    __device__ __forceinline__ LT operator()(int di, int dj, int dk) const {
        const int sl = dk < 0 ? sm : (dk > 0 ? sp : s0);
        return L->v[C][sl][ty + 1 + dj][tx + 1 + di];
    }
};
template <typename LT, int TY>
struct LTapP {
    const LdsPlanes<LT, TY> *L;
    int tx, ty;
    LT below;                  // P(i, j, k-1)
    __device__ __forceinline__ LT operator()(int di, int dj, int dk) const {
        return dk < 0 ? below : L->p[ty + 1 + dj][tx + 1 + di];
    }
};

// plane k of component C (or of P when C == 3) into an LDS slot; elements
// beyond the halo of the valid cells are not needed and not loaded
template <typename LT, int TY, int C, bool OPEN>
__device__ __forceinline__ void load_plane(LT (*dst)[LX], const double *__restrict__ f, const GridDev &g, int i0,
                                           int j0, int k, int tid) {
    for (int e = tid; e < (TY + 2) * LX; e += BX * TY) {
        const int ly = e / LX, lx = e - ly * LX;
        const int i = i0 - 1 + lx, j = j0 - 1 + ly;
        LT val = LT(0);
        if (i <= g.nx + 1 && j <= g.ny + 1) {
            if (C == 3) val = LT(scal_at(f, g, i, j, k));
            else val = LT(vel_at<C == 3 ? 0 : C, OPEN>(f, g, i, j, k));
        }
        dst[ly][lx] = val;
    }
}

// The same plane loads on grids without open sides, where every ghost rule
// is a product of per-axis maps (map_index, map_scal: mapped index, sign,
// forced zero). A thread's plane elements (its slots) are the same in every
// plane, so their x and y maps are evaluated once per thread; a plane adds
// only its z map, the same for the whole block.
template <int TY>
struct PlaneSlots {
    static constexpr int NE = (TY + 2) * LX, NT = BX * TY, NS = (NE + NT - 1) / NT;
    int off[4][NS];    // x/y part of the linear index per field (u, v, w, P)
    unsigned neg;      // bit 4 q + f: the x/y maps negate
    unsigned zero;     // bit 4 q + f: forced zero (a wall-normal ghost or an element not loaded)
    bool has[NS];      // the slot is one of the plane's elements
};
template <int TY>
__device__ __forceinline__ void plane_slots(PlaneSlots<TY> &S, const GridDev &g, int i0, int j0, int tid) {
    S.neg = 0u;
    S.zero = 0u;
#pragma unroll
    for (int q = 0; q < PlaneSlots<TY>::NS; ++q) {
        const int e = tid + q * PlaneSlots<TY>::NT;
        S.has[q] = e < PlaneSlots<TY>::NE;
        const int ly = e / LX, lx = e - ly * LX;
        const int i = i0 - 1 + lx, j = j0 - 1 + ly;
        const bool inside = S.has[q] && i <= g.nx + 1 && j <= g.ny + 1;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            int ii = i, jj = j;
            double sg = 1.0;
            bool ok = true;
            if (f == 3) {
                map_scal(ii, g.nx, g.bt[0], g.bt[1], sg);
                map_scal(jj, g.ny, g.bt[2], g.bt[3], sg);
            } else {
                ok = (f == 0 ? map_index<true>(ii, g.nx, g.bt[0], g.bt[1], sg)
                             : map_index<false>(ii, g.nx, g.bt[0], g.bt[1], sg)) &&
                     (f == 1 ? map_index<true>(jj, g.ny, g.bt[2], g.bt[3], sg)
                             : map_index<false>(jj, g.ny, g.bt[2], g.bt[3], sg));
            }
            ok = ok && inside;
            S.off[f][q] = ok ? (jj - 1) * g.nx + (ii - 1) : 0;
            if (!ok) S.zero |= 1u << (4 * q + f);
            if (sg < 0) S.neg |= 1u << (4 * q + f);
        }
    }
}
// plane k of field F (0..2 velocity components, 3 P) through the slots
template <typename LT, int TY, int F>
__device__ __forceinline__ void load_plane_slots(LT (*dst)[LX], const double *__restrict__ f, const GridDev &g,
                                                 const PlaneSlots<TY> &S, int k, int tid) {
    int kk = k;
    double sg = 1.0;
    bool ok = true;
    if (F == 3) map_scal(kk, g.nz, g.bt[4], g.bt[5], sg);
    else if (F == 2) ok = map_index<true>(kk, g.nz, g.bt[4], g.bt[5], sg);
    else ok = map_index<false>(kk, g.nz, g.bt[4], g.bt[5], sg);
    const long zoff = ok ? (long)(kk - 1) * g.nx * g.ny : 0;
    const unsigned zneg = sg < 0 ? 1u : 0u;
    double v[PlaneSlots<TY>::NS];
#pragma unroll
    for (int q = 0; q < PlaneSlots<TY>::NS; ++q) {
        const bool z = !ok || ((S.zero >> (4 * q + F)) & 1u);
        v[q] = f[z ? 0 : zoff + S.off[F][q]];
        const unsigned neg = ((S.neg >> (4 * q + F)) & 1u) ^ zneg;
        v[q] = neg ? -v[q] : v[q];
        v[q] = z ? 0.0 : v[q];
    }
#pragma unroll
    for (int q = 0; q < PlaneSlots<TY>::NS; ++q) {
        const int e = tid + q * PlaneSlots<TY>::NT;
        if (S.has[q]) (&dst[0][0])[e] = LT(v[q]);
    }
}

// The same split in two: the loads of a plane into registers (issued one
// plane ahead, so they are in flight during the current plane's arithmetic)
// and their ghost rule and LDS store.
template <int TY>
struct PlaneFetch {
    double v[PlaneSlots<TY>::NS];
    bool zok;        // the plane's z map is not a forced zero
    unsigned zneg;   // the plane's z map negates
};
template <int TY, int F>
__device__ __forceinline__ void plane_fetch(PlaneFetch<TY> &P, const double *__restrict__ f, const GridDev &g,
                                            const PlaneSlots<TY> &S, int k) {
    int kk = k;
    double sg = 1.0;
    bool ok = true;
    if (F == 3) map_scal(kk, g.nz, g.bt[4], g.bt[5], sg);
    else if (F == 2) ok = map_index<true>(kk, g.nz, g.bt[4], g.bt[5], sg);
    else ok = map_index<false>(kk, g.nz, g.bt[4], g.bt[5], sg);
    const long zoff = ok ? (long)(kk - 1) * g.nx * g.ny : 0;
    P.zok = ok;
    P.zneg = sg < 0 ? 1u : 0u;
#pragma unroll
    for (int q = 0; q < PlaneSlots<TY>::NS; ++q) {
        const bool z = !ok || ((S.zero >> (4 * q + F)) & 1u);
        P.v[q] = f[z ? 0 : zoff + S.off[F][q]];
    }
}
template <typename LT, int TY, int F>
__device__ __forceinline__ void plane_commit(LT (*dst)[LX], const PlaneFetch<TY> &P, const PlaneSlots<TY> &S,
                                             int tid) {
#pragma unroll
    for (int q = 0; q < PlaneSlots<TY>::NS; ++q) {
        const bool z = !P.zok || ((S.zero >> (4 * q + F)) & 1u);
        const unsigned neg = ((S.neg >> (4 * q + F)) & 1u) ^ P.zneg;
        double v = neg ? -P.v[q] : P.v[q];
        v = z ? 0.0 : v;
        if (S.has[q]) (&dst[0][0])[tid + q * PlaneSlots<TY>::NT] = LT(v);
    }
}

template <typename CT, bool BULK, bool P2, bool OPEN, int TY, bool GEN>
__device__ __forceinline__ void predictor_lds_body(const Blk &kb, const GridDev &g, const double *__restrict__ u,
                                                   const double *__restrict__ v, const double *__restrict__ w,
                                                   const double *__restrict__ P, double *__restrict__ un,
                                                   double *__restrict__ vn, double *__restrict__ wn,
                                                   const PredParams &pp, const GeneralArgs &ga, int kz,
                                                   double *__restrict__ partial) {
    __shared__ LdsPlanes<CT, TY> L;
    __shared__ double red[TY][3];
    const int bx = kb.bx, by = kb.by, bz = kb.bz;
    const int tx = kb.tx, ty = kb.ty, tid = tx + BX * ty;
    const int i0 = bx * BX + 1, j0 = by * TY + 1;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = bz * kz + 1, k1 = min(g.nz, k0 + kz - 1);
    const bool mine = i <= g.nx && j <= g.ny;
    PlaneSlots<TY> S;
    if (!OPEN) plane_slots<TY>(S, g, i0, j0, tid);
    // plane kk of field F into dst
    auto plane = [&](auto fld, CT (*dst)[LX], const double *__restrict__ f, int kk) {
        constexpr int F = decltype(fld)::value;
        if (OPEN) load_plane<CT, TY, F, OPEN>(dst, f, g, i0, j0, kk, tid);
        else load_plane_slots<CT, TY, F>(dst, f, g, S, kk, tid);
    };
    using F0 = std::integral_constant<int, 0>;
    using F1 = std::integral_constant<int, 1>;
    using F2 = std::integral_constant<int, 2>;
    using F3 = std::integral_constant<int, 3>;
    // planes k0-1 and k0 of u, v, w
    plane(F0{}, L.v[0][(k0 - 1) % 3], u, k0 - 1);
    plane(F1{}, L.v[1][(k0 - 1) % 3], v, k0 - 1);
    plane(F2{}, L.v[2][(k0 - 1) % 3], w, k0 - 1);
    plane(F0{}, L.v[0][k0 % 3], u, k0);
    plane(F1{}, L.v[1][k0 % 3], v, k0);
    plane(F2{}, L.v[2][k0 % 3], w, k0);
    CT below = mine ? CT(scal_at(P, g, i, j, k0 - 1)) : CT(0);
    // grids without open sides: planes k+1 of u, v, w and k of P are in
    // registers when iteration k starts (the first from here)
    PlaneFetch<TY> fu, fv, fw, fp;
    if (!OPEN) {
        plane_fetch<TY, 0>(fu, u, g, S, k0 + 1);
        plane_fetch<TY, 1>(fv, v, g, S, k0 + 1);
        plane_fetch<TY, 2>(fw, w, g, S, k0 + 1);
        plane_fetch<TY, 3>(fp, P, g, S, k0);
    }
    for (int k = k0; k <= k1; ++k) {
        const int sm = (k - 1) % 3, s0 = k % 3, sp = (k + 1) % 3;
        if (OPEN) {
            plane(F0{}, L.v[0][sp], u, k + 1);
            plane(F1{}, L.v[1][sp], v, k + 1);
            plane(F2{}, L.v[2][sp], w, k + 1);
            plane(F3{}, L.p, P, k);
        } else {
            plane_commit<CT, TY, 0>(L.v[0][sp], fu, S, tid);
            plane_commit<CT, TY, 1>(L.v[1][sp], fv, S, tid);
            plane_commit<CT, TY, 2>(L.v[2][sp], fw, S, tid);
            plane_commit<CT, TY, 3>(L.p, fp, S, tid);
            if (k < k1) {
                plane_fetch<TY, 0>(fu, u, g, S, k + 2);
                plane_fetch<TY, 1>(fv, v, g, S, k + 2);
                plane_fetch<TY, 2>(fw, w, g, S, k + 2);
                plane_fetch<TY, 3>(fp, P, g, S, k + 1);
            }
        }
        __syncthreads();
        double s[3] = {0.0, 0.0, 0.0};
        if (mine) {
            const long o = cidx(g, i, j, k);
            CT R[3], Pg[3];
            LTapP<CT, TY> tp{&L, tx, ty, below};
            predictor_cell<CT, P2>(LTap<CT, TY, 0>{&L, sm, s0, sp, tx, ty}, LTap<CT, TY, 1>{&L, sm, s0, sp, tx, ty},
                                   LTap<CT, TY, 2>{&L, sm, s0, sp, tx, ty}, tp, g, pp, R, Pg);
            below = L.p[ty + 1][tx + 1];
            const bool face[3] = {wall_face<0>(g, i, j, k), wall_face<1>(g, i, j, k), wall_face<2>(g, i, j, k)};
            const double fval[3] = {face_value<0>(g), face_value<1>(g), face_value<2>(g)};
            double *out[3] = {un, vn, wn};
            const double *cur[3] = {u, v, w};
            const CT a = CT(pp.a), c = CT(pp.c);
            for (int q = 0; q < 3; ++q) {
                double nv;
                if (GEN) {
                    const CT rhs = R[q] + CT(ga.f[q][o]);
                    const CT val = a * rhs + CT(ga.b) * CT(ga.rhsl[q][o]) - c * Pg[q];
                    nv = face[q] ? fval[q] : ga.uold[q][o] + pp.dt * (double)val;
                    ga.rhsl[q][o] = (double)rhs;
                } else {
                    const CT inc = a * R[q] - c * Pg[q];
                    nv = face[q] ? fval[q] : cur[q][o] + pp.dt * (double)inc;
                }
                out[q][o] = nv;
                s[q] = nv;
            }
        }
        if (BULK) row_sums(kb, s, red);
        __syncthreads();   // slot sm is reloaded next iteration; red is complete
        // a group of rows beyond ny (64x8 tiles over the last rows) has no partial
        if (BULK && tx == 0 && (ty & 3) == 0 && j0 + ty <= g.ny)
            group_partial(red, ty, partial + 3 * group_index(g, bx, (j0 - 1) / 4 + ty / 4, k));
    }
}
template <typename CT, bool BULK, bool P2, bool OPEN, int TY, bool GEN>
__global__ __launch_bounds__(BX *TY) void k_predictor_lds(GridDev g, const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ P, double *__restrict__ un,
                                                          double *__restrict__ vn, double *__restrict__ wn,
                                                          PredParams pp, GeneralArgs ga, int kz,
                                                          double *__restrict__ partial) {
    predictor_lds_body<CT, BULK, P2, OPEN, TY, GEN>(xcd_blk(), g, u, v, w, P, un, vn, wn, pp, ga, kz, partial);
}

__global__ __launch_bounds__(256) void k_copy3(long n, const double *__restrict__ a, const double *__restrict__ b,
                                               const double *__restrict__ c, double *__restrict__ x,
                                               double *__restrict__ y, double *__restrict__ z) {
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        x[q] = a[q];
        y[q] = b[q];
        z[q] = c[q];
    }
}

// First stage of the bulk sums: block b adds the row-group partials of its
// contiguous chunk (thread t takes every 256th, then a tree) into out[3b + c].
constexpr int BULK_BLOCKS = 256;
__device__ __forceinline__ void bulk_reduce_body(const Blk &kb, const double *__restrict__ partial, long nparts,
                                                 double *__restrict__ out) {
    __shared__ double red[3][256];
    const int t = kb.tx;
    const long chunk = (nparts + kb.gx - 1) / kb.gx;
    const long lo = (long)kb.bx * chunk, hi = min(nparts, lo + chunk);
    double s[3] = {0, 0, 0};
    for (long q = lo + t; q < hi; q += 256)
        for (int c = 0; c < 3; ++c) s[c] += partial[3 * q + c];
    for (int c = 0; c < 3; ++c) red[c][t] = s[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int c = 0; c < 3; ++c) red[c][t] += red[c][t + st];
        __syncthreads();
    }
    if (t < 3) out[3 * kb.bx + t] = red[t][0];
}
__global__ __launch_bounds__(256) void k_bulk_reduce(const double *__restrict__ partial, long nparts,
                                                     double *__restrict__ out) {
    bulk_reduce_body(hw_blk(), partial, nparts, out);
}

// ModifyBulkVelocity finalize (fluid.F90:177-183): one block sums the block
// partials in a fixed order and stores the shifts dt*modifyP.
__device__ __forceinline__ void bulk_finalize_body(const Blk &kb, const double *__restrict__ partial, long nparts,
                                                   double *__restrict__ scal, int sw0, int sw1, int sw2, double t0,
                                                   double t1, double t2, double ntot, double dt) {
    __shared__ double red[3][256];
    const int t = kb.tx;
    double s[3] = {0, 0, 0};
    for (long q = t; q < nparts; q += 256)
        for (int c = 0; c < 3; ++c) s[c] += partial[3 * q + c];
    for (int c = 0; c < 3; ++c) red[c][t] = s[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int c = 0; c < 3; ++c) red[c][t] += red[c][t + st];
        __syncthreads();
    }
    if (t == 0) {
        const int sw[3] = {sw0, sw1, sw2};
        const double tg[3] = {t0, t1, t2};
        for (int c = 0; c < 3; ++c) {
            double sum = red[c][0];
            scal[SC_BULKSUM + c] = sum;
            double mod = 0.0;
            if (sw[c] == 1) mod = -(sum / ntot - tg[c]) / dt;
            scal[SC_MODP + c] = dt * mod;
        }
    }
}
__global__ __launch_bounds__(256) void k_bulk_finalize(const double *__restrict__ partial, long nparts,
                                                       double *__restrict__ scal, int sw0, int sw1, int sw2,
                                                       double t0, double t1, double t2, double ntot,
                                                       double dt) {
    bulk_finalize_body(hw_blk(), partial, nparts, scal, sw0, sw1, sw2, t0, t1, t2, ntot, dt);
}

__global__ void k_bulk_from_sums(double *scal, int sw0, int sw1, int sw2, double t0, double t1, double t2,
                                 double ntot, double dt) {
    const int sw[3] = {sw0, sw1, sw2};
    const double tg[3] = {t0, t1, t2};
    for (int c = 0; c < 3; ++c) {
        double mod = 0.0;
        if (sw[c] == 1) mod = -(scal[SC_BULKSUM + c] / ntot - tg[c]) / dt;
        scal[SC_MODP + c] = dt * mod;
    }
}

// apply the bulk shift in place (standalone ModifyBulkVelocity)
__global__ __launch_bounds__(BX *BY) void k_bulk_apply(GridDev g, double *u, double *v, double *w,
                                                       const double *__restrict__ scal) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int i = bx * BX + threadIdx.x + 1, j = by * BY + threadIdx.y + 1, k = bz + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = cidx(g, i, j, k);
    if (!wall_face<0>(g, i, j, k)) u[o] = u[o] + scal[SC_MODP + 0];
    if (!wall_face<1>(g, i, j, k)) v[o] = v[o] + scal[SC_MODP + 1];
    if (!wall_face<2>(g, i, j, k)) w[o] = w[o] + scal[SC_MODP + 2];
}

// VelocityDivergence(c) (fluid.F90:97-100) on the (pending-shifted) velocity.
// OPEN: the outflow planes are first set to the predicted velocity of the
// last interior face (zero normal gradient), by the thread of the last cell,
// which reads that face as its own.
// A divisor with its reciprocal: when it is a power of two, x / d is
// exactly x * (1/d) (a scaling by a power of two rounds once, to the same
// value), and the kernel multiplies instead of running the division sequence.
struct Divisor {
    double d, inv;
    int p2;
};
inline Divisor divisor(double d) { return Divisor{d, 1.0 / d, is_pow2(d) ? 1 : 0}; }
__device__ __forceinline__ double dv(double x, const Divisor &D) {
    if (D.p2) return x * D.inv;
    return x / D.d;
}

template <bool OPEN>
__device__ __forceinline__ void divergence_body(const Blk &kb, const GridDev &g, const double *__restrict__ u,
                                                const double *__restrict__ v, const double *__restrict__ w,
                                                double *__restrict__ rhs, const double *__restrict__ scal,
                                                int shifted, const Divisor &ddx, const Divisor &ddy,
                                                const Divisor &ddz, const Divisor &ddt, const Divisor &dcrk,
                                                double rhof) {
    const int i = kb.bx * BX + kb.tx + 1, j = kb.by * BY + kb.ty + 1, k = kb.bz + 1;
    if (i > g.nx || j > g.ny) return;
    double su = 0.0, sv = 0.0, sw = 0.0;
    if (shifted) { su = scal[SC_MODP]; sv = scal[SC_MODP + 1]; sw = scal[SC_MODP + 2]; }
    double u1, u0, v1, v0, w1, w0;
    if (OPEN) {
        u0 = vel_at<0, true>(u, g, i, j, k);
        v0 = vel_at<1, true>(v, g, i, j, k);
        w0 = vel_at<2, true>(w, g, i, j, k);
        const long o = cidx(g, i, j, k);
        if (i == g.nx && g.bt[1] == IBMDEM_BC_OUTFLOW) { u1 = u[o]; g.plane[0][plane_off<0>(g, i, j, k)] = u1; }
        else u1 = vel_at<0, true>(u, g, i + 1, j, k);
        if (j == g.ny && g.bt[3] == IBMDEM_BC_OUTFLOW) { v1 = v[o]; g.plane[1][plane_off<1>(g, i, j, k)] = v1; }
        else v1 = vel_at<1, true>(v, g, i, j + 1, k);
        if (k == g.nz && g.bt[5] == IBMDEM_BC_OUTFLOW) { w1 = w[o]; g.plane[2][plane_off<2>(g, i, j, k)] = w1; }
        else w1 = vel_at<2, true>(w, g, i, j, k + 1);
    } else {
        // the six taps' references first, then their loads together
        const VelRef a1 = vel_ref<0>(g, i + 1, j, k), a0 = vel_ref<0>(g, i, j, k);
        const VelRef b1 = vel_ref<1>(g, i, j + 1, k), b0 = vel_ref<1>(g, i, j, k);
        const VelRef c1 = vel_ref<2>(g, i, j, k + 1), c0 = vel_ref<2>(g, i, j, k);
        const double la1 = u[a1.idx], la0 = u[a0.idx], lb1 = v[b1.idx], lb0 = v[b0.idx];
        const double lc1 = w[c1.idx], lc0 = w[c0.idx];
        if (shifted) {
            u1 = vel_val_shift<0>(g, a1, la1, su); u0 = vel_val_shift<0>(g, a0, la0, su);
            v1 = vel_val_shift<1>(g, b1, lb1, sv); v0 = vel_val_shift<1>(g, b0, lb0, sv);
            w1 = vel_val_shift<2>(g, c1, lc1, sw); w0 = vel_val_shift<2>(g, c0, lc0, sw);
        } else {
            u1 = vel_val(a1, la1); u0 = vel_val(a0, la0);
            v1 = vel_val(b1, lb1); v0 = vel_val(b0, lb0);
            w1 = vel_val(c1, lc1); w0 = vel_val(c0, lc0);
        }
    }
    double d = dv(u1 - u0, ddx) + dv(v1 - v0, ddy) + dv(w1 - w0, ddz);
    rhs[cidx(g, i, j, k)] = dv(dv(-d, ddt), dcrk) * rhof;
}
template <bool OPEN>
__global__ __launch_bounds__(BX *BY) void k_divergence(GridDev g, const double *__restrict__ u,
                                                       const double *__restrict__ v,
                                                       const double *__restrict__ w, double *__restrict__ rhs,
                                                       const double *__restrict__ scal, int shifted, Divisor ddx,
                                                       Divisor ddy, Divisor ddz, Divisor ddt, Divisor dcrk,
                                                       double rhof) {
    divergence_body<OPEN>(xcd_blk(), g, u, v, w, rhs, scal, shifted, ddx, ddy, ddz, ddt, dcrk, rhof);
}

// PressureCorrection + VelocityCorrection(c) (fluid.F90:206, 221-223):
// P_new = P + P' and u = u* + c*Px_old*dt - c*(P_new(i)-P_new(i-1))/dx*dt/rhof,
// with Px_old = (P(i)-P(i-1))/dx/rhof from the pressure before correction.
// OPEN: the thread of the last cell before an outflow side corrects the
// outflow plane by the same formula, with the ghost pressures -P and -P_new.
template <bool OPEN>
__device__ __forceinline__ void project_body(const Blk &kb, GridDev g, double *u, double *v, double *w,
                                                    const double *__restrict__ P, const double *__restrict__ Pc,
                                                    double *__restrict__ Pn, const double *__restrict__ scal,
                                                    int shifted, double dt, double cRK, Divisor ddx, Divisor ddy,
                                                    Divisor ddz, Divisor drho) {
    const int i = kb.bx * BX + kb.tx + 1, j = kb.by * BY + kb.ty + 1, k = kb.bz + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = cidx(g, i, j, k);
    double *vel[3] = {u, v, w};
    // every load first (the cell's own u, v, w are read and written by this
    // thread only): the lower neighbours' P and P' through the scalar ghost
    // rule, the cell's P, P' and velocities
    long nb[3];
    bool ng[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nb[c] = scal_ref(g, c == 0 ? i - 1 : i, c == 1 ? j - 1 : j, c == 2 ? k - 1 : k, ng[c]);
    const double p0 = P[o], q0 = Pc[o];
    double pml[3], qml[3], usl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pml[c] = P[nb[c]];
        qml[c] = Pc[nb[c]];
        usl[c] = vel[c][o];
    }
    const double pn0 = p0 + q0;
    Pn[o] = pn0;
    const Divisor d[3] = {ddx, ddy, ddz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double pm = ng[c] ? -pml[c] : pml[c], qm = ng[c] ? -qml[c] : qml[c];
        const double Pold_grad = dv(dv(p0 - pm, d[c]), drho);
        const double pnm = pm + qm;
        const bool face = (c == 0 ? i : c == 1 ? j : k) == 1 && g.bt[2 * c] != IBMDEM_BC_PERIODIC;
        if (face) {
            vel[c][o] = g.bt[2 * c] == IBMDEM_BC_INFLOW ? g.uin[c] : 0.0;
            continue;
        }
        double us = usl[c];
        if (shifted) us = us + scal[SC_MODP + c];
        vel[c][o] = us + cRK * Pold_grad * dt - dv(dv(cRK * (pn0 - pnm), d[c]) * dt, drho);
    }
    if (OPEN) {
        const int ijk[3] = {i, j, k}, n3[3] = {g.nx, g.ny, g.nz};
        for (int c = 0; c < 3; ++c) {
            if (ijk[c] != n3[c] || g.bt[2 * c + 1] != IBMDEM_BC_OUTFLOW) continue;
            const double pg = -p0, png = -pn0;
            const double Pold_grad = dv(dv(pg - p0, d[c]), drho);
            const long po = c == 0 ? plane_off<0>(g, i, j, k) : (c == 1 ? plane_off<1>(g, i, j, k) : plane_off<2>(g, i, j, k));
            const double us = g.plane[c][po];
            g.plane[c][po] = us + cRK * Pold_grad * dt - dv(dv(cRK * (png - pn0), d[c]) * dt, drho);
        }
    }
}
template <bool OPEN>
__global__ __launch_bounds__(BX *BY) void k_project(GridDev g, double *u, double *v, double *w,
                                                    const double *__restrict__ P, const double *__restrict__ Pc,
                                                    double *__restrict__ Pn, const double *__restrict__ scal,
                                                    int shifted, double dt, double cRK, Divisor ddx, Divisor ddy,
                                                    Divisor ddz, Divisor drho) {
    project_body<OPEN>(xcd_blk(), g, u, v, w, P, Pc, Pn, scal, shifted, dt, cRK, ddx, ddy, ddz, drho);
}

// upload-time Boundary: the low boundary faces take their fixed values (zero
// on walls, the inflow velocity on inflow sides); the outflow plane of each
// uploaded normal component (bit c of `reset`) starts from the last interior
// face (zero normal gradient)
__global__ __launch_bounds__(BX *BY) void k_wall_faces(GridDev g, double *u, double *v, double *w, int reset) {
    const int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = cidx(g, i, j, k);
    if (wall_face<0>(g, i, j, k)) u[o] = face_value<0>(g);
    if (wall_face<1>(g, i, j, k)) v[o] = face_value<1>(g);
    if (wall_face<2>(g, i, j, k)) w[o] = face_value<2>(g);
    if ((reset & 1) && i == g.nx && g.bt[1] == IBMDEM_BC_OUTFLOW) g.plane[0][plane_off<0>(g, i, j, k)] = u[o];
    if ((reset & 2) && j == g.ny && g.bt[3] == IBMDEM_BC_OUTFLOW) g.plane[1][plane_off<1>(g, i, j, k)] = v[o];
    if ((reset & 4) && k == g.nz && g.bt[5] == IBMDEM_BC_OUTFLOW) g.plane[2][plane_off<2>(g, i, j, k)] = w[o];
}

// ---------------------------------------------------------------------------
// deterministic reduction: RED_BLOCKS blocks each sum a contiguous chunk,
// thread-strided, then an LDS tree; the finalize block sums the partials the
// same way. The association is fixed by (n, RED_BLOCKS, RED_THREADS).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void reduce_partial_body(const Blk &kb, const double *__restrict__ a,
                                                    const double *__restrict__ b, long n, long skip,
                                                    double *__restrict__ partial) {
    __shared__ double red[RED_THREADS];
    const int t = kb.tx;
    const long chunk = (n + kb.gx - 1) / kb.gx;
    const long lo = (long)kb.bx * chunk, hi = min(n, lo + chunk);
    // four independent sums over the thread's strided elements (four rounds
    // of loads in flight), then (s0 + s1) + (s2 + s3)
    auto term = [&](long q) { return b ? a[q] * b[q] : a[q]; };
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    long q = lo + t;
    for (; q + 3 * RED_THREADS < hi; q += 4 * RED_THREADS) {
        const double v0 = term(q), v1 = term(q + RED_THREADS), v2 = term(q + 2 * RED_THREADS),
                     v3 = term(q + 3 * RED_THREADS);
        if (q >= skip) s0 += v0;
        if (q + RED_THREADS >= skip) s1 += v1;
        if (q + 2 * RED_THREADS >= skip) s2 += v2;
        if (q + 3 * RED_THREADS >= skip) s3 += v3;
    }
    for (; q < hi; q += RED_THREADS)
        if (q >= skip) s0 += term(q);
    red[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) partial[kb.bx] = red[0];
}
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partial(const double *__restrict__ a,
                                                                const double *__restrict__ b, long n,
                                                                long skip, double *__restrict__ partial) {
    reduce_partial_body(hw_blk(), a, b, n, skip, partial);
}

__device__ __forceinline__ void reduce_final_body(const Blk &kb, const double *__restrict__ partial, int np,
                                                  double *__restrict__ out) {
    __shared__ double red[RED_THREADS];
    const int t = kb.tx;
    double s = 0.0;
    for (int q = t; q < np; q += RED_THREADS) s += partial[q];
    red[t] = s;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) *out = red[0];
}
__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *__restrict__ partial, int np,
                                                              double *__restrict__ out) {
    reduce_final_body(hw_blk(), partial, np, out);
}

dim3 field_grid(const GridDev &g) { return dim3(div_up(g.nx, BX), div_up(g.ny, BY), g.nz); }

}  // namespace

int reduce_sum(ibmdem_ctx *c, const double *a, const double *b, long n, int slot, long skip_first) {
    int nb = (int)std::min<long>(RED_BLOCKS, std::max<long>(1, n));
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(RED_THREADS), 0, c->stream, a, b, n, skip_first,
                       c->partial);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, c->stream, c->partial, nb,
                       c->scal + slot);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

// uold <- u at the start of a step (FirstOrderForwardEuler, fluid.F90:258);
// only contexts for the general VelocityGuess keep uold
int fluid_save_uold(ibmdem_ctx *c) {
    if (!c->uold[0]) return IBMDEM_OK;
    const int cur = c->vcur;
    hipLaunchKernelGGL(k_copy3, dim3(4096), dim3(256), 0, c->stream, c->n, c->u[cur], c->v[cur], c->w[cur],
                       c->uold[0], c->uold[1], c->uold[2]);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

namespace {

// Predictor launch shape from case.stencil_tile: 0 (default) LDS z-marching
// 64x4 tiles over 32 planes, 1 the same over 16 planes, 2 LDS 64x8 tiles
// over 32 planes, 3 the direct kernel (one thread per cell, taps through
// L1/L2). All give the same result bit for bit.
struct PredShape {
    dim3 grid, blk;
    int kz, ty;   // ty == 0: direct kernel
};
PredShape pred_shape(const ibmdem_ctx *c) {
    const GridDev &g = c->g;
    int ty = 4, kz = 32;
    switch (c->cs.stencil_tile) {
        case 1: kz = 16; break;
        case 2: ty = 8; break;
        case 3: return {field_grid(g), dim3(BX, BY), 0, 0};
        default: break;
    }
    return {dim3(div_up(g.nx, BX), div_up(g.ny, ty), div_up(g.nz, kz)), dim3(BX, ty), kz, ty};
}

template <typename CT, bool BK, bool PT, bool OP, bool GEN>
void launch_pred(ibmdem_ctx *c, const PredShape &S, const PredParams &pp, const GeneralArgs &ga) {
    const GridDev &g = c->g;
    const int cur = c->vcur, nxt = 1 - c->vcur;
    const double *u = c->u[cur], *v = c->v[cur], *w = c->w[cur], *P = c->P[c->pcur];
    double *un = c->u[nxt], *vn = c->v[nxt], *wn = c->w[nxt];
    if (S.ty == 0) {
        if (GEN)
            hipLaunchKernelGGL((k_predictor_general<CT, BK, PT, OP>), S.grid, S.blk, 0, c->stream, g, u, v, w, P, un,
                               vn, wn, pp, ga, c->partial);
        else
            hipLaunchKernelGGL((k_predictor<CT, BK, PT, OP>), S.grid, S.blk, 0, c->stream, g, u, v, w, P, un, vn, wn,
                               pp, c->partial);
    } else if (S.ty == 8) {
        hipLaunchKernelGGL((k_predictor_lds<CT, BK, PT, OP, 8, GEN>), S.grid, S.blk, 0, c->stream, g, u, v, w, P, un,
                           vn, wn, pp, ga, S.kz, c->partial);
    } else {
        hipLaunchKernelGGL((k_predictor_lds<CT, BK, PT, OP, 4, GEN>), S.grid, S.blk, 0, c->stream, g, u, v, w, P, un,
                           vn, wn, pp, ga, S.kz, c->partial);
    }
}

template <bool GEN>
void dispatch_pred(ibmdem_ctx *c, const PredShape &S, const PredParams &pp, const GeneralArgs &ga, bool bulk) {
    const GridDev &g = c->g;
    const bool p2 = is_pow2(g.dx) && is_pow2(g.dy) && is_pow2(g.dz) && is_pow2(c->cs.rhof);
    const bool mixed = c->cs.precision == IBMDEM_MIXED;
    // open sides exclude bulk forcing (rejected at creation)
    if (grid_open(g)) {
        if (mixed) { if (p2) launch_pred<float, false, true, true, GEN>(c, S, pp, ga); else launch_pred<float, false, false, true, GEN>(c, S, pp, ga); }
        else { if (p2) launch_pred<double, false, true, true, GEN>(c, S, pp, ga); else launch_pred<double, false, false, true, GEN>(c, S, pp, ga); }
    } else if (mixed) {
        if (bulk) { if (p2) launch_pred<float, true, true, false, GEN>(c, S, pp, ga); else launch_pred<float, true, false, false, GEN>(c, S, pp, ga); }
        else { if (p2) launch_pred<float, false, true, false, GEN>(c, S, pp, ga); else launch_pred<float, false, false, false, GEN>(c, S, pp, ga); }
    } else {
        if (bulk) { if (p2) launch_pred<double, true, true, false, GEN>(c, S, pp, ga); else launch_pred<double, true, false, false, GEN>(c, S, pp, ga); }
        else { if (p2) launch_pred<double, false, true, false, GEN>(c, S, pp, ga); else launch_pred<double, false, false, false, GEN>(c, S, pp, ga); }
    }
}

PredParams pred_params(const ibmdem_ctx *c, double a, double cc) {
    const GridDev &g = c->g;
    PredParams pp{c->cs.dt, c->cs.visc, c->cs.rhof, a, cc, {c->cs.cp[0], c->cs.cp[1], c->cs.cp[2]}};
    pp.idx = 1.0 / g.dx; pp.idy = 1.0 / g.dy; pp.idz = 1.0 / g.dz;
    pp.idx2 = 1.0 / (g.dx * g.dx); pp.idy2 = 1.0 / (g.dy * g.dy); pp.idz2 = 1.0 / (g.dz * g.dz);
    pp.irhof = 1.0 / c->cs.rhof;
    return pp;
}

// the row-group partials of the predictor feed the bulk finalize
int bulk_after_predictor(ibmdem_ctx *c) {
    const GridDev &g = c->g;
    const long nparts = (long)div_up(g.nx, BX) * div_up(g.ny, 4) * g.nz;
    // two stages: BULK_BLOCKS chunk sums after the row-group partials, then one block
    double *second = c->partial + 3 * nparts;
    hipLaunchKernelGGL(k_bulk_reduce, dim3(BULK_BLOCKS), dim3(256), 0, c->stream, c->partial, nparts, second);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_bulk_finalize, dim3(1), dim3(256), 0, c->stream, second, (long)BULK_BLOCKS, c->scal,
                       c->cs.bulk_switch[0], c->cs.bulk_switch[1], c->cs.bulk_switch[2], c->cs.bulk_target[0],
                       c->cs.bulk_target[1], c->cs.bulk_target[2], (double)((long)g.nx * g.ny * g.nz), c->cs.dt);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

}  // namespace

int fluid_predictor(ibmdem_ctx *c, double a, double b, double cc) {
    if (c->uold[0]) return fluid_predictor_general(c, a, b, cc);
    if (b != 0.0 || a != 1.0)
        return set_error(c, IBMDEM_ERR_UNSUPPORTED,
                         "general VelocityGuess coefficients need a context created with time_scheme = "
                         "IBMDEM_SCHEME_RK3 (it keeps uold and RHSl)");
    const PredParams pp = pred_params(c, a, cc);
    const PredShape S = pred_shape(c);
    const bool bulk = c->cs.bulk_switch[0] == 1 || c->cs.bulk_switch[1] == 1 || c->cs.bulk_switch[2] == 1;
    GeneralArgs ga{};
    prof_begin(c, IBMDEM_K_PREDICTOR);
    dispatch_pred<false>(c, S, pp, ga, bulk);
    CHECK_LAUNCH(c);
    // reads u, v, w, P; writes u*, v*, w*
    prof_end(c, IBMDEM_K_PREDICTOR, 7.0 * 8.0 * (double)c->n);
    c->vcur = 1 - c->vcur;
    c->bulk_pending = false;
    return bulk ? bulk_after_predictor(c) : IBMDEM_OK;
}

int fluid_predictor_general(ibmdem_ctx *c, double a, double b, double cc) {
    const PredParams pp = pred_params(c, a, cc);
    int rc = ibm_force_field(c, c->ffull[0], c->ffull[1], c->ffull[2]);
    if (rc) return rc;
    GeneralArgs ga;
    for (int q = 0; q < 3; ++q) {
        ga.uold[q] = c->uold[q];
        ga.f[q] = c->ffull[q];
        ga.rhsl[q] = c->rhsl[q];
    }
    ga.b = b;
    const PredShape S = pred_shape(c);
    const bool bulk = c->cs.bulk_switch[0] == 1 || c->cs.bulk_switch[1] == 1 || c->cs.bulk_switch[2] == 1;
    prof_begin(c, IBMDEM_K_PREDICTOR);
    dispatch_pred<true>(c, S, pp, ga, bulk);
    CHECK_LAUNCH(c);
    // reads u, v, w, P, uold (3), f (3), RHSl (3); writes u*, v*, w*, RHSl (3)
    prof_end(c, IBMDEM_K_PREDICTOR, 19.0 * 8.0 * (double)c->n);
    c->vcur = 1 - c->vcur;
    c->bulk_pending = false;
    return bulk ? bulk_after_predictor(c) : IBMDEM_OK;
}

int fluid_bulk(ibmdem_ctx *c, bool apply_now) {
    const bool bulk = c->cs.bulk_switch[0] == 1 || c->cs.bulk_switch[1] == 1 || c->cs.bulk_switch[2] == 1;
    if (!bulk) return IBMDEM_OK;
    const GridDev &g = c->g;
    if (apply_now) {
        // standalone ModifyBulkVelocity: sum u,v,w now (fluid.F90:163-165),
        // then shift in place
        for (int q = 0; q < 3; ++q) {
            const double *f = q == 0 ? c->u[c->vcur] : q == 1 ? c->v[c->vcur] : c->w[c->vcur];
            int rc = reduce_sum(c, f, nullptr, c->n, SC_BULKSUM + q, 0);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_bulk_from_sums, dim3(1), dim3(1), 0, c->stream, c->scal, c->cs.bulk_switch[0],
                           c->cs.bulk_switch[1], c->cs.bulk_switch[2], c->cs.bulk_target[0],
                           c->cs.bulk_target[1], c->cs.bulk_target[2], (double)((long)g.nx * g.ny * g.nz),
                           c->cs.dt);
        CHECK_LAUNCH(c);
        hipLaunchKernelGGL(k_bulk_apply, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                           c->v[c->vcur], c->w[c->vcur], c->scal);
        CHECK_LAUNCH(c);
        c->bulk_pending = false;
    } else {
        // the sums were reduced inside the predictor; divergence and
        // projection apply the shift on the fly
        c->bulk_pending = true;
    }
    return IBMDEM_OK;
}

int fluid_divergence(ibmdem_ctx *c, double cc) {
    const GridDev &g = c->g;
    prof_begin(c, IBMDEM_K_DIVERGENCE);
    if (grid_open(g))
        hipLaunchKernelGGL(k_divergence<true>, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                           c->v[c->vcur], c->w[c->vcur], c->rhs, c->scal, c->bulk_pending ? 1 : 0, divisor(g.dx),
                           divisor(g.dy), divisor(g.dz), divisor(c->cs.dt), divisor(cc), c->cs.rhof);
    else
        hipLaunchKernelGGL(k_divergence<false>, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                           c->v[c->vcur], c->w[c->vcur], c->rhs, c->scal, c->bulk_pending ? 1 : 0, divisor(g.dx),
                           divisor(g.dy), divisor(g.dz), divisor(c->cs.dt), divisor(cc), c->cs.rhof);
    CHECK_LAUNCH(c);
    prof_end(c, IBMDEM_K_DIVERGENCE, 4.0 * 8.0 * (double)c->n);
    return IBMDEM_OK;
}

int fluid_project(ibmdem_ctx *c, double cc) {
    const GridDev &g = c->g;
    const int pn = 1 - c->pcur;
    prof_begin(c, IBMDEM_K_PROJECT);
    if (grid_open(g))
        hipLaunchKernelGGL(k_project<true>, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                           c->v[c->vcur], c->w[c->vcur], c->P[c->pcur], c->pcorr, c->P[pn], c->scal,
                           c->bulk_pending ? 1 : 0, c->cs.dt, cc, divisor(g.dx), divisor(g.dy), divisor(g.dz),
                           divisor(c->cs.rhof));
    else
        hipLaunchKernelGGL(k_project<false>, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                           c->v[c->vcur], c->w[c->vcur], c->P[c->pcur], c->pcorr, c->P[pn], c->scal,
                           c->bulk_pending ? 1 : 0, c->cs.dt, cc, divisor(g.dx), divisor(g.dy), divisor(g.dz),
                           divisor(c->cs.rhof));
    CHECK_LAUNCH(c);
    prof_end(c, IBMDEM_K_PROJECT, 9.0 * 8.0 * (double)c->n);
    c->pcur = pn;
    c->bulk_pending = false;
    return IBMDEM_OK;
}

int fluid_upload_bc(ibmdem_ctx *c, int reset_planes) {
    const GridDev &g = c->g;
    hipLaunchKernelGGL(k_wall_faces, field_grid(g), dim3(BX, BY), 0, c->stream, g, c->u[c->vcur],
                       c->v[c->vcur], c->w[c->vcur], reset_planes);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

// ---------------------------------------------------------------------------
// fluid_project(c, 1) of every context of a group as one launch per kind
// (closed, open boundaries); each case's pointers follow its own parity.
// ---------------------------------------------------------------------------
namespace {
struct ProjArg {
    BGrid g;
    GridDev grid;
    double *u, *v, *w, *Pn;
    const double *P, *Pc, *scal;
    int shifted;
    double dt;
    Divisor ddx, ddy, ddz, drho;
};
template <bool OPEN>
__global__ __launch_bounds__(BX *BY) void kb_project(BTable<ProjArg> tab) {
    batch_for(tab, [&](const ProjArg &a, const Blk &kb) {
        GridDev g = a.grid;
        for (int d = 0; d < 3; ++d) g.plane[d] = glb(g.plane[d]);
        project_body<OPEN>(kb, g, glb(a.u), glb(a.v), glb(a.w), glb(a.P), glb(a.Pc), glb(a.Pn), glb(a.scal),
                           a.shifted, a.dt, 1.0, a.ddx, a.ddy, a.ddz, a.drho);
    });
}
template <bool OPEN>
void launch_project(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_project<OPEN>, dim3(o.blocks), dim3(BX * BY), 0, s, table<ProjArg>(o));
}
}  // namespace

int fluid_plan_project(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<ProjArg> open, closed;
    for (ibmdem_ctx *c : cs) {
        const GridDev &g = c->g;
        ProjArg a{};
        a.g = bgrid(field_grid(g), dim3(BX, BY));
        a.grid = g;
        a.u = c->u[c->vcur]; a.v = c->v[c->vcur]; a.w = c->w[c->vcur];
        a.P = c->P[c->pcur];
        a.Pc = c->pcorr;
        a.Pn = c->P[1 - c->pcur];
        a.scal = c->scal;
        a.shifted = c->bulk_pending ? 1 : 0;
        a.dt = c->cs.dt;
        a.ddx = divisor(g.dx); a.ddy = divisor(g.dy); a.ddz = divisor(g.dz);
        a.drho = divisor(c->cs.rhof);
        (grid_open(g) ? open : closed).push_back(a);
    }
    int rc;
    if ((rc = add_op(P.ar, P.ops, launch_project<false>, BX * BY, closed, BOp(), err))) return rc;
    return add_op(P.ar, P.ops, launch_project<true>, BX * BY, open, BOp(), err);
}

// ---------------------------------------------------------------------------
// The predictor (with its bulk sums) and the divergence of a batched step,
// and reduce_sum's launches for several sums: every context of a group in
// one launch per kernel variant (precision, bulk forcing, power-of-two
// spacing, open sides, tile schedule). Each case's pointers follow its own
// parity, as fluid_predictor and fluid_divergence read it.
// ---------------------------------------------------------------------------
namespace {
struct PredArg {
    BGrid g;
    GridDev grid;
    const double *u, *v, *w, *P;
    double *un, *vn, *wn;
    PredParams pp;
    int kz;
    double *partial;
};
__device__ __forceinline__ GridDev glb_grid(const GridDev &src) {
    GridDev g = src;
    for (int d = 0; d < 3; ++d) g.plane[d] = glb(g.plane[d]);
    return g;
}
template <typename CT, bool BULK, bool P2, bool OPEN>
__global__ __launch_bounds__(BX *BY) void kb_predictor(BTable<PredArg> tab) {
    batch_for(tab, [&](const PredArg &a, const Blk &kb) {
        const GridDev g = glb_grid(a.grid);
        predictor_body<CT, BULK, P2, OPEN>(kb, g, glb(a.u), glb(a.v), glb(a.w), glb(a.P), glb(a.un), glb(a.vn),
                                           glb(a.wn), a.pp, glb(a.partial));
    });
}
template <typename CT, bool BULK, bool P2, bool OPEN, int TY>
__global__ __launch_bounds__(BX *TY) void kb_predictor_lds(BTable<PredArg> tab) {
    batch_for(tab, [&](const PredArg &a, const Blk &kb) {
        const GridDev g = glb_grid(a.grid);
        predictor_lds_body<CT, BULK, P2, OPEN, TY, false>(kb, g, glb(a.u), glb(a.v), glb(a.w), glb(a.P), glb(a.un),
                                                          glb(a.vn), glb(a.wn), a.pp, GeneralArgs{}, a.kz,
                                                          glb(a.partial));
    });
}
template <typename CT, bool BULK, bool P2, bool OPEN>
void launch_predictor(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL((kb_predictor<CT, BULK, P2, OPEN>), dim3(o.blocks), dim3(BX * BY), 0, s, table<PredArg>(o));
}
template <typename CT, bool BULK, bool P2, bool OPEN, int TY>
void launch_predictor_lds(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL((kb_predictor_lds<CT, BULK, P2, OPEN, TY>), dim3(o.blocks), dim3(BX * TY), 0, s,
                       table<PredArg>(o));
}
// the launcher of a tile schedule (PredShape::ty: 0 direct, 4 or 8)
template <typename CT, bool BULK, bool P2, bool OPEN>
BLaunch pred_launcher(int ty) {
    return ty == 0 ? launch_predictor<CT, BULK, P2, OPEN>
         : ty == 8 ? launch_predictor_lds<CT, BULK, P2, OPEN, 8>
                   : launch_predictor_lds<CT, BULK, P2, OPEN, 4>;
}
// dispatch_pred's variant choice (open sides exclude bulk forcing)
template <typename CT>
BLaunch pred_launcher(bool bulk, bool p2, bool open, int ty) {
    if (open) return p2 ? pred_launcher<CT, false, true, true>(ty) : pred_launcher<CT, false, false, true>(ty);
    if (bulk) return p2 ? pred_launcher<CT, true, true, false>(ty) : pred_launcher<CT, true, false, false>(ty);
    return p2 ? pred_launcher<CT, false, true, false>(ty) : pred_launcher<CT, false, false, false>(ty);
}

struct BulkArg {
    BGrid g;
    const double *partial;
    long nparts;
    double *second, *scal;
    int sw[3];
    double tg[3];
    double ntot, dt;
};
__global__ __launch_bounds__(256) void kb_bulk_reduce(BTable<BulkArg> tab) {
    batch_for(tab, [&](const BulkArg &a, const Blk &kb) {
        bulk_reduce_body(kb, glb(a.partial), a.nparts, glb(a.second));
    });
}
__global__ __launch_bounds__(256) void kb_bulk_finalize(BTable<BulkArg> tab) {
    batch_for(tab, [&](const BulkArg &a, const Blk &kb) {
        bulk_finalize_body(kb, glb((const double *)a.second), BULK_BLOCKS, glb(a.scal), a.sw[0], a.sw[1], a.sw[2],
                           a.tg[0], a.tg[1], a.tg[2], a.ntot, a.dt);
    });
}
void launch_bulk_reduce(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_bulk_reduce, dim3(o.blocks), dim3(256), 0, s, table<BulkArg>(o));
}
void launch_bulk_finalize(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_bulk_finalize, dim3(o.blocks), dim3(256), 0, s, table<BulkArg>(o));
}

struct DivArg {
    BGrid g;
    GridDev grid;
    const double *u, *v, *w, *scal;
    double *rhs;
    int shifted;
    Divisor ddx, ddy, ddz, ddt, dcrk;
    double rhof;
};
template <bool OPEN>
__global__ __launch_bounds__(BX *BY) void kb_divergence(BTable<DivArg> tab) {
    batch_for(tab, [&](const DivArg &a, const Blk &kb) {
        const GridDev g = glb_grid(a.grid);
        divergence_body<OPEN>(kb, g, glb(a.u), glb(a.v), glb(a.w), glb(a.rhs), glb(a.scal), a.shifted, a.ddx, a.ddy,
                              a.ddz, a.ddt, a.dcrk, a.rhof);
    });
}
template <bool OPEN>
void launch_divergence(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_divergence<OPEN>, dim3(o.blocks), dim3(BX * BY), 0, s, table<DivArg>(o));
}

struct RedArg {
    BGrid g;
    const double *a, *b;
    long n, skip;
    double *partial, *out;
    int np;
};
__global__ __launch_bounds__(RED_THREADS) void kb_reduce_partial(BTable<RedArg> tab) {
    batch_for(tab, [&](const RedArg &a, const Blk &kb) {
        reduce_partial_body(kb, glb(a.a), a.b ? glb(a.b) : nullptr, a.n, a.skip, glb(a.partial));
    });
}
__global__ __launch_bounds__(RED_THREADS) void kb_reduce_final(BTable<RedArg> tab) {
    batch_for(tab, [&](const RedArg &a, const Blk &kb) {
        reduce_final_body(kb, glb((const double *)a.partial), a.np, glb(a.out));
    });
}
void launch_reduce_partial(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_reduce_partial, dim3(o.blocks), dim3(RED_THREADS), 0, s, table<RedArg>(o));
}
void launch_reduce_final(const BOp &o, hipStream_t s, int) {
    hipLaunchKernelGGL(kb_reduce_final, dim3(o.blocks), dim3(RED_THREADS), 0, s, table<RedArg>(o));
}

bool bulk_on(const ibmdem_ctx *c) {
    return c->cs.bulk_switch[0] == 1 || c->cs.bulk_switch[1] == 1 || c->cs.bulk_switch[2] == 1;
}
}  // namespace

int fluid_plan_predictor(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    // the cases of each launcher, with its block size
    std::map<BLaunch, std::pair<int, std::vector<PredArg>>> by;
    std::vector<BulkArg> bulk_red, bulk_fin;
    for (ibmdem_ctx *c : cs) {
        if (c->uold[0]) {
            err = "batched predictor: the general VelocityGuess (RK3 contexts) is not batched";
            return IBMDEM_ERR_UNSUPPORTED;
        }
        const GridDev &g = c->g;
        const PredShape S = pred_shape(c);
        const bool p2 = is_pow2(g.dx) && is_pow2(g.dy) && is_pow2(g.dz) && is_pow2(c->cs.rhof);
        const bool bulk = bulk_on(c), open = grid_open(g);
        const BLaunch fn = c->cs.precision == IBMDEM_MIXED ? pred_launcher<float>(bulk, p2, open, S.ty)
                                                           : pred_launcher<double>(bulk, p2, open, S.ty);
        const int cur = c->vcur, nxt = 1 - c->vcur;
        PredArg a{};
        a.g = bgrid(S.grid, S.blk);
        a.grid = g;
        a.u = c->u[cur]; a.v = c->v[cur]; a.w = c->w[cur];
        a.P = c->P[c->pcur];
        a.un = c->u[nxt]; a.vn = c->v[nxt]; a.wn = c->w[nxt];
        a.pp = pred_params(c, 1.0, 1.0);
        a.kz = S.kz;
        a.partial = c->partial;
        auto &slot = by[fn];
        slot.first = (int)(S.blk.x * S.blk.y);
        slot.second.push_back(a);
        if (bulk) {
            BulkArg b{};
            b.partial = c->partial;
            b.nparts = (long)div_up(g.nx, BX) * div_up(g.ny, 4) * g.nz;
            b.second = c->partial + 3 * b.nparts;
            b.scal = c->scal;
            for (int q = 0; q < 3; ++q) {
                b.sw[q] = c->cs.bulk_switch[q];
                b.tg[q] = c->cs.bulk_target[q];
            }
            b.ntot = (double)((long)g.nx * g.ny * g.nz);
            b.dt = c->cs.dt;
            b.g = bgrid(dim3(BULK_BLOCKS), dim3(256));
            bulk_red.push_back(b);
            b.g = bgrid(dim3(1), dim3(256));
            bulk_fin.push_back(b);
        }
    }
    int rc;
    for (auto &kv : by)
        if ((rc = add_op(P.ar, P.ops, kv.first, kv.second.first, kv.second.second, BOp(), err))) return rc;
    if ((rc = add_op(P.ar, P.ops, launch_bulk_reduce, 256, bulk_red, BOp(), err))) return rc;
    return add_op(P.ar, P.ops, launch_bulk_finalize, 256, bulk_fin, BOp(), err);
}

int fluid_plan_divergence(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    std::vector<DivArg> open, closed;
    for (ibmdem_ctx *c : cs) {
        const GridDev &g = c->g;
        DivArg a{};
        a.g = bgrid(field_grid(g), dim3(BX, BY));
        a.grid = g;
        a.u = c->u[c->vcur]; a.v = c->v[c->vcur]; a.w = c->w[c->vcur];
        a.scal = c->scal;
        a.rhs = c->rhs;
        a.shifted = c->bulk_pending ? 1 : 0;
        a.ddx = divisor(g.dx); a.ddy = divisor(g.dy); a.ddz = divisor(g.dz);
        a.ddt = divisor(c->cs.dt);
        a.dcrk = divisor(1.0);
        a.rhof = c->cs.rhof;
        (grid_open(g) ? open : closed).push_back(a);
    }
    int rc;
    if ((rc = add_op(P.ar, P.ops, launch_divergence<false>, BX * BY, closed, BOp(), err))) return rc;
    return add_op(P.ar, P.ops, launch_divergence<true>, BX * BY, open, BOp(), err);
}

int plan_reduce_sum(bk::BatchPlan &P, const std::vector<ReduceJob> &jobs, std::string &err) {
    std::vector<RedArg> part, fin;
    for (const ReduceJob &j : jobs) {
        const int nb = (int)std::min<long>(RED_BLOCKS, std::max<long>(1, j.n));
        RedArg a{};
        a.a = j.a; a.b = j.b;
        a.n = j.n; a.skip = j.skip;
        a.partial = j.partial; a.out = j.out;
        a.np = nb;
        a.g = bgrid(dim3(nb), dim3(RED_THREADS));
        part.push_back(a);
        a.g = bgrid(dim3(1), dim3(RED_THREADS));
        fin.push_back(a);
    }
    int rc;
    if ((rc = add_op(P.ar, P.ops, launch_reduce_partial, RED_THREADS, part, BOp(), err))) return rc;
    return add_op(P.ar, P.ops, launch_reduce_final, RED_THREADS, fin, BOp(), err);
}
