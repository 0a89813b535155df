// common.hpp — shared device helpers and the context layout of libibmdem.
//
// Fields live in HBM as interior-only arrays (nx*ny*nz, x fastest). The
// reference keeps two ghost layers per side and refills them after every
// operator (FillGhostRegionForFluid + Boundary, boundary.F90:122-247,
// fluid.F90:240-250); here every kernel evaluates the ghost value it needs
// from the interior with the same rules:
//   periodic side          ghost = wrapped interior value;
//   wall side, tangential  ghost = mirrored interior value, negated
//                          (ghost 0 <- 1, -1 <- 2; n+1 <- n, n+2 <- n-1);
//   wall side, normal      ghost = 0, and the wall face itself (index 1 on
//                          the low side) is kept 0 by every writer;
//   pressure               mirrored without negation (Neumann).
// Corner ghosts compose the per-dimension rules, which is what the
// reference's x-then-y-then-z halo passes produce.
// Open sides (no counterpart in the reference; the same rules are in the
// oracle):
//   inflow (low sides)     normal component = U_n at the face and on every
//                          ghost layer beyond it, whatever the transverse
//                          index; tangential ghosts mirrored and negated
//                          (zero tangential velocity); pressure Neumann;
//   outflow (high sides)   normal component on the boundary face (index
//                          n+1) and beyond = the stored outflow plane at the
//                          mapped transverse index, times the transverse
//                          mirror signs; tangential ghosts mirrored without
//                          negation (zero gradient); pressure and its
//                          correction mirrored and negated (zero at the
//                          face).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/ibmdem.h"

#define IBMDEM_MAX_LEVELS 16

struct GridDev {
    int nx, ny, nz;
    int bt[6];
    double dx, dy, dz;
    double L[3];
    double uin[3];      // normal velocity of an inflow side of dimension d
    double *plane[3];   // outflow plane of dimension d (normal component on
                        // the face n_d+1), first transverse index fastest
};

__host__ __device__ inline bool grid_open(const GridDev &g) {
    for (int s = 0; s < 6; ++s)
        if (g.bt[s] == IBMDEM_BC_INFLOW || g.bt[s] == IBMDEM_BC_OUTFLOW) return true;
    return false;
}

// ---------------------------------------------------------------------------
// boundary rules
// ---------------------------------------------------------------------------
__host__ __device__ inline long cidx(const GridDev &g, int i, int j, int k) {
    // 1-based interior (i,j,k) -> linear index
    return ((long)(k - 1) * g.ny + (j - 1)) * (long)g.nx + (i - 1);
}

// Map one 1-based index of dimension d (size n) onto the interior.
// Returns 0 if the value is a wall-normal ghost (forced zero), flips *sgn on a
// tangential mirror (wall or inflow side; an outflow side mirrors without
// negation).
template <bool NORMAL>
__device__ __forceinline__ bool map_index(int &i, int n, int lo, int hi, double &sgn) {
    if (i < 1) {
        if (lo == IBMDEM_BC_PERIODIC) { i += n; return true; }
        if (NORMAL) return false;
        i = 1 - i; sgn = -sgn; return true;
    }
    if (i > n) {
        if (hi == IBMDEM_BC_PERIODIC) { i -= n; return true; }
        if (NORMAL) return false;
        i = 2 * n + 1 - i;
        if (hi != IBMDEM_BC_OUTFLOW) sgn = -sgn;
        return true;
    }
    return true;
}

// the value every writer keeps on the low boundary face of component C (a
// wall face zero, an inflow face the inflow velocity)
template <int C>
__device__ __forceinline__ double face_value(const GridDev &g) {
    return g.bt[2 * C] == IBMDEM_BC_INFLOW ? g.uin[C] : 0.0;
}

// offset of the mapped transverse cell (a, b) (1-based, a the lower
// dimension) in the outflow plane of dimension C
template <int C>
__device__ __forceinline__ long plane_off(const GridDev &g, int i, int j, int k) {
    return C == 0 ? (long)(j - 1) + (long)g.ny * (k - 1)
                  : (C == 1 ? (long)(i - 1) + (long)g.nx * (k - 1) : (long)(i - 1) + (long)g.nx * (j - 1));
}

// The open-side rules of velocity component C at (i,j,k): returns true and
// sets *val when dimension C's index lies on the inflow face or beyond it, or
// beyond an outflow face.
template <int C, typename T>
__device__ __forceinline__ bool open_value(const GridDev &g, int i, int j, int k, T &val) {
    const int ic = C == 0 ? i : (C == 1 ? j : k);
    const int nc = C == 0 ? g.nx : (C == 1 ? g.ny : g.nz);
    if (ic <= 1 && g.bt[2 * C] == IBMDEM_BC_INFLOW) { val = T(g.uin[C]); return true; }
    if (ic > nc && g.bt[2 * C + 1] == IBMDEM_BC_OUTFLOW) {
        double s = 1.0;
        if (C != 0) map_index<false>(i, g.nx, g.bt[0], g.bt[1], s);
        if (C != 1) map_index<false>(j, g.ny, g.bt[2], g.bt[3], s);
        if (C != 2) map_index<false>(k, g.nz, g.bt[4], g.bt[5], s);
        const T p = T(g.plane[C][plane_off<C>(g, i, j, k)]);
        val = s < 0 ? -p : p;
        return true;
    }
    return false;
}

// velocity component C (0=u,1=v,2=w) at logical 1-based (i,j,k), each index
// in [-1, n+2], with the reference's post-Boundary ghost semantics. OPEN
// kernels (grids with an inflow or outflow side) apply the open-side rules.
template <int C, bool OPEN = false, typename T>
__device__ __forceinline__ T vel_at(const T *__restrict__ f, const GridDev &g, int i, int j, int k) {
    if (OPEN) {
        T ov;
        if (open_value<C>(g, i, j, k, ov)) return ov;
    }
    double s = 1.0;
    if (!map_index<C == 0>(i, g.nx, g.bt[0], g.bt[1], s)) return T(0);
    if (!map_index<C == 1>(j, g.ny, g.bt[2], g.bt[3], s)) return T(0);
    if (!map_index<C == 2>(k, g.nz, g.bt[4], g.bt[5], s)) return T(0);
    T val = f[cidx(g, i, j, k)];
    return s < 0 ? -val : val;
}

// Same, for a field with a pending uniform shift sh (ModifyBulkVelocity's
// u + dt*modifyP, fluid.F90:188-190): the shift is applied to the interior
// value the ghost is copied from; forced zeros (wall-normal ghosts and the
// low wall face) stay zero, as Boundary re-zeroes them after the shift.
template <int C>
__device__ __forceinline__ double vel_at_shift(const double *__restrict__ f, const GridDev &g,
                                               int i, int j, int k, double sh) {
    double s = 1.0;
    if (!map_index<C == 0>(i, g.nx, g.bt[0], g.bt[1], s)) return 0.0;
    if (!map_index<C == 1>(j, g.ny, g.bt[2], g.bt[3], s)) return 0.0;
    if (!map_index<C == 2>(k, g.nz, g.bt[4], g.bt[5], s)) return 0.0;
    int idx = C == 0 ? i : (C == 1 ? j : k);
    if (idx == 1 && g.bt[2 * C] != IBMDEM_BC_PERIODIC) return face_value<C>(g);
    double val = f[cidx(g, i, j, k)] + sh;
    return s < 0 ? -val : val;
}

// A tap of vel_at / vel_at_shift split into its reference (computed
// without touching memory) and its value, so that a kernel can compute all
// its taps' references, issue all their loads together and then form the
// values: the same numbers as vel_at and vel_at_shift.
struct VelRef {
    long idx;    // linear index to load (0 when the tap is a forced zero)
    bool zero;   // wall-normal ghost: the value is 0
    bool neg;    // tangential mirror: the value is negated
    bool face;   // the mapped cell is the low face of a non-periodic side
};
template <int C>
__device__ __forceinline__ VelRef vel_ref(const GridDev &g, int i, int j, int k) {
    double s = 1.0;
    const bool ok = map_index<C == 0>(i, g.nx, g.bt[0], g.bt[1], s) &&
                    map_index<C == 1>(j, g.ny, g.bt[2], g.bt[3], s) &&
                    map_index<C == 2>(k, g.nz, g.bt[4], g.bt[5], s);
    const int idx = C == 0 ? i : (C == 1 ? j : k);
    return VelRef{ok ? cidx(g, i, j, k) : 0L, !ok, s < 0, idx == 1 && g.bt[2 * C] != IBMDEM_BC_PERIODIC};
}
// vel_at's value from the tap's loaded element
__device__ __forceinline__ double vel_val(const VelRef &r, double loaded) {
    const double v = r.neg ? -loaded : loaded;
    return r.zero ? 0.0 : v;
}
// vel_at_shift's value from the tap's loaded element
template <int C>
__device__ __forceinline__ double vel_val_shift(const GridDev &g, const VelRef &r, double loaded, double sh) {
    if (r.zero) return 0.0;
    if (r.face) return face_value<C>(g);
    const double val = loaded + sh;
    return r.neg ? -val : val;
}

// one index of a cell-centred scalar: wrap, or mirror (negated beyond an
// outflow side, where the scalar is zero at the face)
__device__ __forceinline__ void map_scal(int &i, int n, int lo, int hi, double &sgn) {
    if (i < 1) { i = lo == IBMDEM_BC_PERIODIC ? i + n : 1 - i; return; }
    if (i > n) {
        if (hi == IBMDEM_BC_PERIODIC) { i -= n; return; }
        i = 2 * n + 1 - i;
        if (hi == IBMDEM_BC_OUTFLOW) sgn = -sgn;
    }
}

// cell-centred scalar (P, P'): mirror without negation on walls and inflow
// sides, negated beyond an outflow side
template <typename T>
__device__ __forceinline__ T scal_at(const T *__restrict__ f, const GridDev &g, int i, int j, int k) {
    double s = 1.0;
    map_scal(i, g.nx, g.bt[0], g.bt[1], s);
    map_scal(j, g.ny, g.bt[2], g.bt[3], s);
    map_scal(k, g.nz, g.bt[4], g.bt[5], s);
    const T v = f[cidx(g, i, j, k)];
    return s < 0 ? -v : v;
}

// scal_at split into reference and value, as VelRef
__device__ __forceinline__ long scal_ref(const GridDev &g, int i, int j, int k, bool &neg) {
    double s = 1.0;
    map_scal(i, g.nx, g.bt[0], g.bt[1], s);
    map_scal(j, g.ny, g.bt[2], g.bt[3], s);
    map_scal(k, g.nz, g.bt[4], g.bt[5], s);
    neg = s < 0;
    return cidx(g, i, j, k);
}

// is (i,j,k) the low boundary face of velocity component C (a wall face
// kept zero, or an inflow face kept at the inflow velocity)?
template <int C>
__device__ __forceinline__ bool wall_face(const GridDev &g, int i, int j, int k) {
    int idx = C == 0 ? i : (C == 1 ? j : k);
    return idx == 1 && g.bt[2 * C] != IBMDEM_BC_PERIODIC;
}

// Per-thread axis map: for offsets O0..O0+NO-1 around a 1-based index, the
// 0-based interior index the ghost value comes from, whether it is a
// tangential mirror that negates (wall or inflow side), whether it lies
// beyond a wall (forced zero for the wall-normal component), on or beyond an
// inflow face (the inflow velocity for the normal component) or beyond an
// outflow face (the outflow plane for the normal component; the scalars
// negated). Evaluated once per thread so every tap is one load, with no
// divergent boundary path.
template <int O0, int NO>
struct AxisMap {
    int idx[NO];
    bool mir[NO];
    bool wall[NO];
    bool cst[NO];
    bool pln[NO];
    __device__ __forceinline__ void init(int i, int n, int lo, int hi) {
#pragma unroll
        for (int q = 0; q < NO; ++q) {
            int j = i + O0 + q;
            bool m = false, w = false, p = false;
            const bool c = j <= 1 && lo == IBMDEM_BC_INFLOW;
            if (j < 1) {
                if (lo == IBMDEM_BC_PERIODIC) j += n;
                else { j = 1 - j; m = true; w = lo == IBMDEM_BC_WALL; }
            } else if (j > n) {
                if (hi == IBMDEM_BC_PERIODIC) j -= n;
                else {
                    j = 2 * n + 1 - j;
                    p = hi == IBMDEM_BC_OUTFLOW;
                    m = !p;
                    w = hi == IBMDEM_BC_WALL;
                }
            }
            idx[q] = j - 1;
            mir[q] = m;
            wall[q] = w;
            cst[q] = c;
            pln[q] = p;
        }
    }
};

// velocity component C at mapped offsets (a,b,c) of the three axis maps
template <int C, bool OPEN = false, class MX, class MY, class MZ, typename T>
__device__ __forceinline__ T vel_tap(const T *__restrict__ f, const GridDev &g, const MX &mx, const MY &my,
                                     const MZ &mz, int a, int b, int c) {
    const bool zero = (C == 0 && mx.wall[a]) || (C == 1 && my.wall[b]) || (C == 2 && mz.wall[c]);
    const bool neg = ((C != 0 && mx.mir[a]) ^ (C != 1 && my.mir[b])) ^ (C != 2 && mz.mir[c]);
    const long fo = (long)mx.idx[a] + (long)g.nx * ((long)my.idx[b] + (long)g.ny * mz.idx[c]);
    if (OPEN) {
        const bool cst = (C == 0 && mx.cst[a]) || (C == 1 && my.cst[b]) || (C == 2 && mz.cst[c]);
        const bool pln = (C == 0 && mx.pln[a]) || (C == 1 && my.pln[b]) || (C == 2 && mz.pln[c]);
        const long po = C == 0 ? (long)my.idx[b] + (long)g.ny * mz.idx[c]
                               : (C == 1 ? (long)mx.idx[a] + (long)g.nx * mz.idx[c]
                                         : (long)mx.idx[a] + (long)g.nx * my.idx[b]);
        const T v = pln ? T(g.plane[C][po]) : f[fo];
        return cst ? T(g.uin[C]) : (zero ? T(0) : (neg ? -v : v));
    }
    const T v = f[fo];
    return zero ? T(0) : (neg ? -v : v);
}
// cell-centred scalar at mapped offsets (mirror without negation; negated
// beyond an outflow side in OPEN kernels)
template <bool OPEN = false, class MX, class MY, class MZ, typename T>
__device__ __forceinline__ T scal_tap(const T *__restrict__ f, const GridDev &g, const MX &mx, const MY &my,
                                      const MZ &mz, int a, int b, int c) {
    const T v = f[(long)mx.idx[a] + (long)g.nx * ((long)my.idx[b] + (long)g.ny * mz.idx[c])];
    if (OPEN) {
        const bool neg = (mx.pln[a] ^ my.pln[b]) ^ mz.pln[c];
        return neg ? -v : v;
    }
    return v;
}

// x / d. When d is a power of two (P2), 1/d is exact and x * (1/d) rounds to
// the same double as x / d, so kernels take the multiply without changing a
// bit; otherwise they divide.
template <bool P2>
__device__ __forceinline__ double qdiv(double x, double d, double inv) {
    return P2 ? x * inv : x / d;
}
inline bool is_pow2(double d) {
    if (!(d > 0.0)) return false;
    int e;
    return frexp(d, &e) == 0.5;
}

// ---------------------------------------------------------------------------
// Roma-Peskin-Berger 3-point kernel, DeltaFunction3D (interpolation.F90:7-30)
// ---------------------------------------------------------------------------
//   r <= 1/2:       (1 + sqrt(1 - 3 r^2)) / 3
//   1/2 < r <= 3/2: (5 - 3r - sqrt(1 - 3 (1-r)^2)) / 6
// Both branches are one square root and one division of the same shape, so
// they are evaluated once with per-lane operands (no divergent second
// path); each branch keeps its own operation order, so the value is the
// branch's value bit for bit.
// Branching form of the same function (lower register use where many
// factors are live at once; the value is identical).
__host__ __device__ __forceinline__ double phi1_branch(double r) {
    if (r <= 0.5) return (1.0 + sqrt(1.0 - (3.0 * r) * r)) / 3.0;
    if (r <= 1.5) return ((5.0 - 3.0 * r) - sqrt(1.0 - 3.0 * ((1.0 - r) * (1.0 - r)))) / 6.0;
    return 0.0;
}
__host__ __device__ __forceinline__ double phi1(double r) {
    const bool inner = r <= 0.5;
    const double q = inner ? (3.0 * r) * r : 3.0 * ((1.0 - r) * (1.0 - r));
    const double s = sqrt(1.0 - q);
    const double num = inner ? 1.0 + s : (5.0 - 3.0 * r) - s;
    const double v = num / (inner ? 3.0 : 6.0);
    return r <= 1.5 ? v : 0.0;
}

// cell of a position, PutParcelsIntoCellsDEM (ZJS_UPDATE_LIST_DEM.F90:41-43)
__device__ __forceinline__ int cell_of(double x, double d) { return (int)floor((x - 0.0) / d) + 1; }

// ---------------------------------------------------------------------------
// multigrid level (one precision)
// ---------------------------------------------------------------------------
struct LevelDev {
    int nx, ny, nz;
    int bt[6];
    double ihx2, ihy2, ihz2;     // 1/h^2 per dimension
    double idiag[64];            // 1/diag indexed by the per-dimension counts
    double diag[64];             // (cx,cy,cz) in 0..3, see diag_index
    int rb_valid;                // red-black colouring consistent (no odd periodic dim)
    int dirichlet;               // an outflow side: the operator is non-singular
    int split;                   // colour-split storage (nx even), see lidx
    int half;                    // nx*ny*nz / 2 when split
    // udiv_small multipliers for nx / 2 (split) or nx, ny, nx*ny
    unsigned m_row, m_ny, m_nxy;
};

// ceil(2^32 / d) (0 for d = 1): with it, udiv_small(p, d, m) = p / d exactly
// for p * d < 2^32 (the error term p * (m d - 2^32) / 2^32 stays below p d /
// 2^32 < 1 of the fractional part's room), in one multiply-high instead of
// the integer division sequence
inline unsigned udiv_magic(unsigned d) {
    return d <= 1 ? 0u : (unsigned)(((1ull << 32) + d - 1) / d);
}
__device__ __forceinline__ int udiv_small(int p, int d, unsigned m) {
    return d == 1 ? p : (int)__umulhi((unsigned)p, m);
}
// lcell for levels of at most 2^16 cells (and dimensions): the divisions
// through udiv_small
__device__ __forceinline__ void lcell_small(const LevelDev &L, int q, int &i, int &j, int &k) {
    if (!L.split) {
        const int r = udiv_small(q, L.nx, L.m_row);
        i = q - r * L.nx;
        k = udiv_small(q, L.nx * L.ny, L.m_nxy);
        j = r - k * L.ny;
        return;
    }
    const int c = q >= L.half ? 1 : 0;
    const int r = q - c * L.half, h = L.nx >> 1;
    const int row = udiv_small(r, h, L.m_row);
    k = udiv_small(row, L.ny, L.m_ny);
    j = row - k * L.ny;
    i = 2 * (r - row * h) + ((j + k + c) & 1);
}

// Storage of a level's cell (i,j,k), 0-based. With nx even the level is
// stored colour-split: the (i+j+k)-even cells first, then the odd ones, each
// half in (k, j, i/2) order, so a red-black half sweep reads and writes
// whole cache lines of one colour; otherwise x fastest.
__host__ __device__ __forceinline__ int lidx(const LevelDev &L, int i, int j, int k) {
    const int row = k * L.ny + j;
    if (!L.split) return row * L.nx + i;
    return ((i + j + k) & 1) * L.half + row * (L.nx >> 1) + (i >> 1);
}
// the cell stored at position q
__host__ __device__ __forceinline__ void lcell(const LevelDev &L, int q, int &i, int &j, int &k) {
    if (!L.split) {
        i = q % L.nx; j = (q / L.nx) % L.ny; k = q / (L.nx * L.ny);
        return;
    }
    const int c = q >= L.half ? 1 : 0;
    const int r = q - c * L.half, h = L.nx >> 1;
    const int row = r / h;
    j = row % L.ny;
    k = row / L.ny;
    i = 2 * (r - row * h) + ((j + k + c) & 1);
}

template <typename T>
struct MGLevel {
    LevelDev d;
    long n;
    T *x, *b, *r;                // level solution, RHS, residual
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Profile {
    int kernel = 0;
    std::vector<hipEvent_t> ev;  // pairs
    double total_ms = 0.0;
    int64_t launches = 0;
    double bytes_total = 0.0;    // algorithmic bytes summed over the launches
};

// host-side state that the captured launches depend on or change
struct HostState {
    int vcur, pcur;
    bool bulk_pending, first_mdf, spread_valid, rmax_dirty;
    bool operator==(const HostState &o) const {
        return vcur == o.vcur && pcur == o.pcur && bulk_pending == o.bulk_pending && first_mdf == o.first_mdf &&
               spread_valid == o.spread_valid && rmax_dirty == o.rmax_dirty;
    }
};
struct CapturedPart {
    hipGraphExec_t exec = nullptr;
    HostState before{}, after{};
    int cg_last = 0;             // CG iterations captured (first part only)
};
struct StepGraphs {
    bool enabled = true;         // IBMDEM_NO_GRAPH=1 in the environment turns them off
    int cg_k = -1;               // CG iterations to capture (-1: not yet known)
    CapturedPart pre[4], post[4];
    CapturedPart pre_rhs[4];     // pre without the solve's preconditioning and iterations (batched solve)
    CapturedPart pre_a[4], pre_b[4];   // pre_rhs before and after the forcing loop (batched forcing)
    int64_t launches = 0;        // graph launches (for tests)
    int64_t epoch = 0;           // drop_graphs calls (a batch re-plans when it changes)
};

struct ibmdem_ctx {
    ibmdem_case cs;
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    GridDev g;
    long n = 0;                  // interior cells
    double dV = 0.0;             // dx*dy*dz as the reference forms it
    double eta = 0.0;
    // fluid (FP64, interior-only)
    double *u[2] = {}, *v[2] = {}, *w[2] = {};
    int vcur = 0;
    double *P[2] = {};
    int pcur = 0;
    double *pcorr = nullptr, *rhs = nullptr;
    double *rhsl[3] = {};        // RHSul, RHSvl, RHSwl (RK3 contexts)
    double *uold[3] = {};        // uold, vold, wold (RK3 contexts)
    double *ffull[3] = {};       // the IBM force as full fields (RK3 contexts)
    // PCG (FP64)
    double *cg_x = nullptr, *cg_r = nullptr, *cg_z = nullptr, *cg_p = nullptr, *cg_Ap = nullptr;
    int nlev = 0;
    int tail_first = IBMDEM_MAX_LEVELS;  // levels from here down run as one work-group
    LevelDev *dlev = nullptr;            // level descriptors in device memory (for that work-group)
    MGLevel<double> lev64[IBMDEM_MAX_LEVELS];
    MGLevel<float> lev32[IBMDEM_MAX_LEVELS];
    // reductions
    double *partial = nullptr;   // per-block partial sums
    int max_partial = 0;
    double *scal = nullptr;      // device scalars (see SC_* in kernels)
    double *scal_host = nullptr; // pinned mirror
    double *conv_host = nullptr; // pinned, one convergence flag per CG iteration
    hipEvent_t iter_ev[2] = {};   // end of the last two queued CG iterations
    int *flags = nullptr;        // device error flags
    int *flags_host = nullptr;
    bool bulk_pending = false;   // bulk shift computed but not yet applied
    // markers
    int M = 0, B = 0;
    double *mx = nullptr, *mv = nullptr, *fdrag = nullptr, *fcol = nullptr;
    double *mvol = nullptr, *mmass = nullptr, *mrad = nullptr;
    int *mbody = nullptr;
    int *mcell = nullptr;        // 0-based linear interior cell
    // per-marker record, 8 doubles (64 B): x, y, z, bits of the cell index
    // (set once per step), fdrag*vol (3, set by every interpolation), pad
    double *mrec = nullptr;
    double *mfac = nullptr;              // per marker: 12 (face, centre) delta-factor pairs
    // bins over interior cells
    int *bin_count = nullptr, *bin_start = nullptr, *bin_items = nullptr, *bin_tmp = nullptr;
    int *scan_blocks = nullptr;
    int scan_nblocks = 0;
    // IBM spread footprint (per step): active 4x4x4 tiles and, per tile, its
    // source entries (marker << 9 | source cell) in (kc,jc,ic,id) order
    unsigned char *act_flag = nullptr;   // one flag per tile
    int *act_list = nullptr;
    int *act_meta = nullptr;             // [0] active tiles (device-side count)
    long act_cap = 0;                    // tiles
    long act_grid = 0;                   // launch width: act_cap rounded up to 8
    int *tile_ecount = nullptr;          // act_cap + 1
    int *tile_estart = nullptr;          // act_cap + 1 (exclusive scan)
    uint32_t *tile_ent = nullptr;        // ent_cap
    long ent_cap = 0;
    unsigned long long *tile_mask = nullptr;  // cells with a contribution, IBM_MASK_WORDS per tile
    double *fact = nullptr;              // f of the tiles' cells, 3 * 64 * act_cap
    bool first_mdf = true;
    bool spread_valid = false;           // fact/tile_mask hold the last spread
    long footprint_cells = 0;            // cells the spread writes (measured when profiling)
    // bodies
    int *body_start = nullptr;   // B+1
    double *bx = nullptr, *bv = nullptr, *bvold = nullptr, *bom = nullptr, *bL = nullptr, *bLold = nullptr;
    double *brot = nullptr, *brotg = nullptr, *bI = nullptr, *bm = nullptr, *brho = nullptr;
    int *bfixed = nullptr;
    double *bfdrag = nullptr, *bfmom = nullptr, *bfcol = nullptr, *bfmomcol = nullptr;
    double *bfhydro = nullptr, *bfmomhydro = nullptr;
    double *brmax = nullptr;     // max |x_m - X_b| over the body's markers
    int *bnear = nullptr;        // number of other bodies within contact reach; [B]: any
                                 // body has one
    int *blist = nullptr;        // their ids, up to 32 per body
    bool rmax_dirty = true;
    bool bin_gate = true;        // rebuilds between sub-steps gated by the culling
                                 // (IBMDEM_NO_BIN_GATE=1 at creation: always bin)
    // pair list for parity (allocated on demand)
    int *pair_count = nullptr, *pair_partner = nullptr, *pair_shift = nullptr;
    int pair_cap = 0;
    Profile prof;
    // per-body probe ring (ibmdem_probe_bodies): after every step the listed
    // bodies' 18 values are appended on the device; probe_count (device) is
    // the number of steps recorded so far
    int probe_n = 0, probe_cap = 0, probe_list_cap = 0;
    long probe_ring_cap = 0;
    int *probe_list = nullptr;
    double *probe_ring = nullptr;
    long long *probe_count = nullptr;
    // the coupled step as two captured HIP graphs (see ctx.hip), one pair
    // per (vcur, pcur) parity of the step
    StepGraphs graphs;
    int last_cg_iterations = 0;
    int64_t bytes = 0;
    std::vector<void *> allocs;
};

// device scalar slots
enum {
    SC_MODP = 0,      // 3: modifyPx,y,z * dt (the bulk shift)
    SC_BULKSUM = 3,   // 3: sums of u,v,w
    SC_RZ = 6, SC_PAP = 7, SC_RR = 8, SC_ALPHA = 9, SC_BETA = 10, SC_BNORM2 = 11,
    SC_RZ_NEW = 12, SC_PIN = 13, SC_X0 = 14,
    SC_CONV = 15,     // 1 once ||r|| <= tol: every later CG kernel returns at once
    SC_TOL = 16,      // rtol * ||b||
    SC_ITER = 17,     // iteration at which SC_CONV was set
    SC_DONE = 18,     // last CG iteration that ran (its x update is deferred)
    SC_COUNT = 32
};
// device error flags: contact candidates beyond maxCollisionNo, a marker
// outside the domain, a non-finite marker or body state, the IBM spread's
// tile entry list beyond its capacity
enum { FL_CAPACITY = 0, FL_OUTSIDE = 1, FL_NAN = 2, FL_SPREAD = 3, FL_COUNT = 8 };

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
int set_error(ibmdem_ctx *c, int code, const std::string &msg);
#define HIPX(c, call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return set_error((c), IBMDEM_ERR_HIP,                                  \
                             std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)
#define CHECK_LAUNCH(c) HIPX(c, hipGetLastError())

template <typename T>
int dev_alloc(ibmdem_ctx *c, T **p, size_t count) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    void *q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) {
        *p = nullptr;
        return set_error(c, IBMDEM_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    if (hipMemsetAsync(q, 0, bytes, c->stream) != hipSuccess)
        return set_error(c, IBMDEM_ERR_HIP, "hipMemsetAsync failed");
    c->allocs.push_back(q);
    c->bytes += (int64_t)bytes;
    *p = (T *)q;
    return IBMDEM_OK;
}

// launch-time profiling of one kernel id
void prof_begin(ibmdem_ctx *c, int kid);
void prof_end(ibmdem_ctx *c, int kid, double bytes, int launches = 1);

// ---------------------------------------------------------------------------
// launchers (one per translation unit)
// ---------------------------------------------------------------------------
// ctx.hip: the coupled step in phases (for batches of contexts)
int step_phase1(ibmdem_ctx *c, ibmdem_solve_info *info, bool *done);
int step_phase2(ibmdem_ctx *c, ibmdem_solve_info *info);
int step_phase3(ibmdem_ctx *c);
void drop_graphs(ibmdem_ctx *c);
// a batch step whose solve runs batched: phase 1 queues the step up to the
// solve's right-hand side, phase 2 takes the solver scalars of the batched
// solve and queues the rest of the step (phase 3 as above)
bool step_batchable(const ibmdem_ctx *c);
int step_phase1_batched(ibmdem_ctx *c);
// pre_rhs without the forcing loop, which the batch runs between the two
int step_phase1a_batched(ibmdem_ctx *c);
int step_phase1b_batched(ibmdem_ctx *c);
void step_mdf_batched_state(ibmdem_ctx *c);
// the host state after a first part (up to the solve's right-hand side) that
// the batch ran for the context; the context's host state (for planning
// batched parts at the state a step's later launches see)
void step_pre_rhs_batched_state(ibmdem_ctx *c);
HostState host_state(const ibmdem_ctx *c);
void set_host_state(ibmdem_ctx *c, const HostState &h);
// the solve's report from the batch's copy of the solver scalars
int step_check_batched(ibmdem_ctx *c, const double *scal, ibmdem_solve_info *info);
// the rest of the step on the context's own stream (its captured graph)
int step_post_own(ibmdem_ctx *c);
// the host state after a post part that the batch ran for the context
void step_post_batched_state(ibmdem_ctx *c);
// dem.hip: exclusive scan of n counts into start[0..n] (start[n] = total)
int scan_counts(ibmdem_ctx *c, const int *cnt, int *start, long n);
// fluid.hip
int fluid_predictor(ibmdem_ctx *c, double a, double b, double cc);
int fluid_bulk(ibmdem_ctx *c, bool apply_now);
int fluid_divergence(ibmdem_ctx *c, double cc);
int fluid_project(ibmdem_ctx *c, double cc);
int fluid_upload_bc(ibmdem_ctx *c, int reset_planes);
int fluid_save_uold(ibmdem_ctx *c);
int fluid_predictor_general(ibmdem_ctx *c, double a, double b, double cc);
// poisson.hip
// Batched plans (batched.hpp): each appends the launches of one part of a
// step, for every context of a group, to a plan.
namespace bk {
struct BatchPlan;
}
size_t dem_plan_cases(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs);
int dem_plan_body_hydro(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, std::string &err);
int dem_plan_substep(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, bool first, bool last,
                     std::string &err);
int dem_plan_probe(bk::BatchPlan &P, size_t cases, const std::vector<ibmdem_ctx *> &cs, std::string &err);
// the step's error status from a host copy of the context's flags (clears them)
int dem_flags_status(ibmdem_ctx *c, const int *flags);
// poisson_gauge and fluid_project(c, 1) of every context (the current
// parity of each)
int poisson_plan_gauge(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
int fluid_plan_project(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
// the n_mdf iterations of ibm_interpolate and ibm_spread of every context
int ibm_plan_mdf(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
// ibm_begin_step of every context
int ibm_plan_begin_step(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
// fluid_predictor(c, 1, 0, 1) with its bulk sums, and fluid_divergence(c, 1),
// of every context (at the parity and pending bulk shift it has when planned)
int fluid_plan_predictor(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
int fluid_plan_divergence(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);
// reduce_sum's launches for several sums: out = sum_{q >= skip} a[q] (b[q])
struct ReduceJob {
    const double *a, *b;
    long n, skip;
    double *partial, *out;
};
int plan_reduce_sum(bk::BatchPlan &P, const std::vector<ReduceJob> &jobs, std::string &err);
// poisson_begin_rhs of every context
int poisson_plan_rhs(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err);

int poisson_setup(ibmdem_ctx *c);
long poisson_partials(int nx, int ny, int nz);
int poisson_solve(ibmdem_ctx *c, ibmdem_solve_info *info);
int poisson_begin(ibmdem_ctx *c);
int poisson_begin_rhs(ibmdem_ctx *c);
int poisson_begin_precondition(ibmdem_ctx *c);
int poisson_iterations(ibmdem_ctx *c, int first, int last);
int poisson_iterate_to_convergence(ibmdem_ctx *c, int first);
int poisson_check(ibmdem_ctx *c, ibmdem_solve_info *info);
int poisson_check_scalars(ibmdem_ctx *c, const double *scal, ibmdem_solve_info *info);
int poisson_gauge(ibmdem_ctx *c);

// The batched solve of a group of contexts (same precision, smoothing counts
// and iteration cap): one launch per solver operation over all of them.
struct PoissonBatch;
bool poisson_batchable(const ibmdem_ctx *c);
int poisson_batch_create(const std::vector<ibmdem_ctx *> &cs, PoissonBatch **out, std::string &err);
void poisson_batch_destroy(PoissonBatch *pb);
// z = M^-1 r, r.z, p = z of every case (after each case's poisson_begin_rhs)
int poisson_batch_begin(PoissonBatch *pb, hipStream_t s, std::string &err);
// CG iteration it (1-based) of every case
int poisson_batch_iteration(PoissonBatch *pb, hipStream_t s, int it, std::string &err);
// every case's solver scalars to the host (read after s has drained)
int poisson_batch_gather(PoissonBatch *pb, hipStream_t s, std::string &err);
const double *poisson_batch_scalars(const PoissonBatch *pb, int q);
// launches of the begin part and k iterations
int poisson_batch_launches(const PoissonBatch *pb, int k);
// Profiling (IBMDEM_BK_* kinds, 0..8): while on, a solve queued between
// arm(true) and arm(false) records an event after every launch; collect
// (after the solve has drained) adds the launches' times, counts and
// algorithmic bytes to ms, launches and bytes, indexed by kind.
void poisson_batch_set_profile(PoissonBatch *pb, bool on);
bool poisson_batch_profiling(const PoissonBatch *pb);
void poisson_batch_profile_arm(PoissonBatch *pb, bool armed);
int poisson_batch_profile_collect(PoissonBatch *pb, double *ms, int64_t *launches, double *bytes);
// ibm.hip
int ibm_begin_step(ibmdem_ctx *c);
int ibm_interpolate(ibmdem_ctx *c);
int ibm_spread(ibmdem_ctx *c);
int ibm_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz);
// the IBM force of the last spread as full device fields (zero elsewhere)
int ibm_force_field(ibmdem_ctx *c, double *fx, double *fy, double *fz);
long ibm_tile_count(const GridDev &g);
// spread tiles are IBM_TL^3 cells; each keeps IBM_TL^3 / 64 mask words
constexpr int IBM_TL = 8;
constexpr int IBM_TC = IBM_TL * IBM_TL * IBM_TL;
constexpr int IBM_MASK_WORDS = IBM_TC / 64;
// dem.hip
int dem_setup(ibmdem_ctx *c);
int dem_rebuild(ibmdem_ctx *c);
int dem_contact(ibmdem_ctx *c);
int dem_body_assemble(ibmdem_ctx *c, bool all_terms);
int dem_body_integrate(ibmdem_ctx *c);
int dem_markers_follow(ibmdem_ctx *c);
int dem_body_hydro(ibmdem_ctx *c);
int dem_pairs(ibmdem_ctx *c, int cap);
int dem_check_flags(ibmdem_ctx *c);
// the probe record of the step just completed (no host wait)
int dem_probe_append(ibmdem_ctx *c);

// deterministic reductions (fluid.hip)
// sum over n values of f(i) into scal[slot]; fixed tree, independent of timing
int reduce_sum(ibmdem_ctx *c, const double *a, const double *b, long n, int slot, long skip_first);

static inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }
