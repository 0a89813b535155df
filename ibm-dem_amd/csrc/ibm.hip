// ibm.hip — regularized-delta interpolation and force spreading
// (ZJS_VOLUMEFRACTION_DEM.F90:2-67, ZJS_INTERPOLATE_DEM.F90:2-17,
// interpolation.F90:2-33) with the multi-direct-forcing velocity update
// (fluid.F90:135-153) fused into the spread.
//
// Markers do not move during the fluid part of a step, so at the start of a
// step every marker gets its 24 one-dimensional delta factors (per
// dimension: staggered-face and cell-centre variants at cell offsets
// -1..+2 from its own cell), stored marker-major, and the spread footprint
// is compacted into a sorted list of active 4x4x4 cell tiles.
//
// Interpolation is marker-parallel and sums the taps in the reference's
// (kk,jj,ii) order, so it is bit-identical to the reference's 125-cell loop
// (zero-weight taps add an exact zero there).
//
// Spreading is a deterministic gather, no float atomics. One wavefront owns
// a 4x4x4 tile of target cells. It stages the markers of the tile's 7x7x7
// source cells in LDS (their delta factors, fdrag*vol and cell), in source
// order (kc,jc,ic) and ascending id inside a cell; the reference's periodic
// ghost copies (x + coordShift) and wall-mirror copies (2*coordMirror - x,
// fdrag unnegated; ZJS_MPI_DEM.F90:173-269) enter as wrapped or mirrored
// source cells with their factors evaluated at the image position. Each lane
// then sums, for its own cell, f = 0 - sum (fdrag*vol)*w over its 4x4x4
// source cells in that order, and applies u += dt*(f - f_old). The order a
// cell's contributions are summed in does not depend on the tile size.
#include "common.hpp"

namespace {

constexpr int T = 256;
constexpr int CHUNK = 4096;   // flags per compaction block
constexpr int TL = 4;         // tile edge (cells)
constexpr int SRC = TL + 3;   // source cells per tile edge
constexpr int NSRC = SRC * SRC * SRC;
constexpr int CAP = 64;       // LDS marker entries per pass (one per lane)

__device__ __forceinline__ double wt_face(double xp, int i, double d) {
    const double xc = ((double)i - 0.5) * d;
    const double dl = xp - xc;
    return phi1(fabs((dl + 0.5 * d) / d));
}
__device__ __forceinline__ double wt_cent(double xp, int i, double d) {
    const double xc = ((double)i - 0.5) * d;
    const double dl = xp - xc;
    return phi1(fabs(dl / d));
}

// weights layout (marker-major): wts[m*24 + dim*8 + variant*4 + o],
// variant 0 staggered face, 1 cell centre, o = 0..3 for cells ci-1..ci+2
__global__ __launch_bounds__(T) void k_weights(int M, GridDev g, const double *__restrict__ mx,
                                               const int *__restrict__ mcell, double *__restrict__ wts,
                                               double *__restrict__ fdrag) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int cc[3] = {cl % g.nx + 1, (cl / g.nx) % g.ny + 1, cl / (g.nx * g.ny) + 1};
    const double dd[3] = {g.dx, g.dy, g.dz};
    double *w = wts + 24L * m;
    for (int d = 0; d < 3; ++d) {
        const double xp = mx[(long)d * M + m];
        for (int o = 0; o < 4; ++o) {
            const int cell = cc[d] + o - 1;
            w[d * 8 + o] = wt_face(xp, cell, dd[d]);
            w[d * 8 + 4 + o] = wt_cent(xp, cell, dd[d]);
        }
        fdrag[(long)d * M + m] = 0.0;
    }
}

// UpdateFluidVelocityOnLagrangianPoints: U_p = sum w u dx dy dz,
// fdrag += (U_p - v_p)/dt; also stores fdrag*vol for the spread
template <bool P2>
__global__ __launch_bounds__(T) void k_interp(int M, GridDev g, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ w,
                                              const int *__restrict__ mcell, const double *__restrict__ wts,
                                              const double *__restrict__ mv, double *__restrict__ fdrag,
                                              const double *__restrict__ mvol, double *__restrict__ fv, double dV,
                                              double idV, double dt) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    const double *wm = wts + 24L * m;
    double fx[4], cx[4], fy[4], cy[4], fz[4], cz[4];
    for (int o = 0; o < 4; ++o) {
        fx[o] = wm[o]; cx[o] = wm[4 + o];
        fy[o] = wm[8 + o]; cy[o] = wm[12 + o];
        fz[o] = wm[16 + o]; cz[o] = wm[20 + o];
    }
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    double ulp = 0.0, vlp = 0.0, wlp = 0.0;
    // u: x-face taps -1..2, y/z centre taps -1..1 (the nonzero ranges)
    for (int kk = 0; kk < 3; ++kk)
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 4; ++ii) {
                const double wt = qdiv<P2>(fx[ii] * cy[jj] * cz[kk], dV, idV);
                ulp = ulp + wt * vel_at<0>(u, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    for (int kk = 0; kk < 3; ++kk)
        for (int jj = 0; jj < 4; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const double wt = qdiv<P2>(cx[ii] * fy[jj] * cz[kk], dV, idV);
                vlp = vlp + wt * vel_at<1>(v, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    for (int kk = 0; kk < 4; ++kk)
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const double wt = qdiv<P2>(cx[ii] * cy[jj] * fz[kk], dV, idV);
                wlp = wlp + wt * vel_at<2>(w, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    const double f0 = fdrag[m] + 1.0 * (ulp - mv[m]) / dt;
    const double f1 = fdrag[(long)M + m] + 1.0 * (vlp - mv[(long)M + m]) / dt;
    const double f2 = fdrag[2L * M + m] + 1.0 * (wlp - mv[2L * M + m]) / dt;
    fdrag[m] = f0;
    fdrag[(long)M + m] = f1;
    fdrag[2L * M + m] = f2;
    const double vol = mvol[m];
    fv[4L * m + 0] = f0 * vol;
    fv[4L * m + 1] = f1 * vol;
    fv[4L * m + 2] = f2 * vol;
}

struct TileGeom {
    int ntx, nty, ntz;
};

// mark the tiles holding an interior cell that a marker (or an image) can reach
__global__ __launch_bounds__(T) void k_mark_tiles(int M, GridDev g, TileGeom tg, const int *__restrict__ mcell,
                                                  unsigned char *__restrict__ flag) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int cc[3] = {cl % g.nx + 1, (cl / g.nx) % g.ny + 1, cl / (g.nx * g.ny) + 1};
    const int n[3] = {g.nx, g.ny, g.nz};
    const int nt[3] = {tg.ntx, tg.nty, tg.ntz};
    int lo[3], hi[3];
    bool wrap[3];
    for (int d = 0; d < 3; ++d) {
        int a = cc[d] - 1, b = cc[d] + 2;
        wrap[d] = g.bt[2 * d] == IBMDEM_BC_PERIODIC;
        if (!wrap[d]) { a = a < 1 ? 1 : a; b = b > n[d] ? n[d] : b; }
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

// per-block counts of set flags (block b owns flags [b*CHUNK, (b+1)*CHUNK))
__global__ __launch_bounds__(T) void k_flag_count(long n, const unsigned char *__restrict__ flag, int *__restrict__ cnt) {
    __shared__ int red[T];
    const long lo = (long)blockIdx.x * CHUNK;
    int s = 0;
    for (long q = lo + threadIdx.x; q < min(n, lo + CHUNK); q += T) s += flag[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = T / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = red[0];
}

// exclusive scan of nb ints in one block; total into a[nb]
__global__ __launch_bounds__(1024) void k_scan_small(int nb, int *__restrict__ a) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int q = base + threadIdx.x;
        int v = q < nb ? a[q] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int st = 1; st < 1024; st <<= 1) {
            int t = threadIdx.x >= st ? buf[threadIdx.x - st] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (q < nb) a[q] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[nb] = carry;
}

// ordered compaction of set flags into list; clears the flags
__global__ __launch_bounds__(T) void k_flag_compact(long n, unsigned char *__restrict__ flag,
                                                    const int *__restrict__ off, int *__restrict__ list) {
    __shared__ int buf[T];
    constexpr int PER = CHUNK / T;   // 16 consecutive flags per thread
    const long lo = (long)blockIdx.x * CHUNK + (long)threadIdx.x * PER;
    int s = 0;
    for (int q = 0; q < PER; ++q)
        if (lo + q < n) s += flag[lo + q];
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int st = 1; st < T; st <<= 1) {
        int t = threadIdx.x >= st ? buf[threadIdx.x - st] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    int pos = off[blockIdx.x] + buf[threadIdx.x] - s;
    for (int q = 0; q < PER; ++q)
        if (lo + q < n && flag[lo + q]) {
            list[pos++] = (int)(lo + q);
            flag[lo + q] = 0;
        }
}

__global__ __launch_bounds__(T) void k_zero(long n, double *__restrict__ a) {
    for (long q = (long)blockIdx.x * T + threadIdx.x; q < n; q += (long)gridDim.x * T) a[q] = 0.0;
}

// one source-cell coordinate mapped to the interior with its image kind:
// 0 direct, 1 periodic (image = x + sh), 2 mirror (image = 2*sh - x)
__device__ __forceinline__ bool src_map(int &i, int n, int lo, int hi, double L, int &kind, double &sh) {
    kind = 0; sh = 0.0;
    if (i < 1) {
        if (lo == IBMDEM_BC_PERIODIC) { i += n; kind = 1; sh = -L; return true; }
        i = 1 - i; kind = 2; sh = 0.0; return i >= 1 && i <= n;
    }
    if (i > n) {
        if (hi == IBMDEM_BC_PERIODIC) { i -= n; kind = 1; sh = L; return true; }
        i = 2 * n + 1 - i; kind = 2; sh = L; return i >= 1 && i <= n;
    }
    return true;
}

struct SpreadArgs {
    GridDev g;
    TileGeom tg;
    int M;
    const int *tiles;
    const int *bin_start, *bin_items;
    const double *mx, *wts, *fv;
    double *fact;      // 3 x (64 * cap_tiles), tile slot major
    long fstride;      // 64 * cap_tiles
    double *u, *v, *w;
    double dV, idV, dt;
};

template <bool P2>
__global__ __launch_bounds__(64) void k_spread_tiles(SpreadArgs a) {
    __shared__ int s_off[NSRC + 1];
    __shared__ int s_gstart[NSRC];       // first item of the source cell in bin_items
    __shared__ int s_kind[NSRC];         // image kinds, 2 bits per dimension
    __shared__ double s_w[CAP][24];
    __shared__ double s_fv[CAP][3];
    __shared__ int s_cc[CAP][3];
    const GridDev &g = a.g;
    const int lane = threadIdx.x;
    const int tile = a.tiles[blockIdx.x];
    const int tx = tile % a.tg.ntx, ty = (tile / a.tg.ntx) % a.tg.nty, tz = tile / (a.tg.ntx * a.tg.nty);
    const int ib = tx * TL - 1, jb = ty * TL - 1, kb = tz * TL - 1;   // 1-based source origin
    const double L3[3] = {g.L[0], g.L[1], g.L[2]};
    // 1. source cells: first item, image kind, marker count. The bin
    // boundaries of all of the lane's cells are loaded before any is used.
    constexpr int SPL = (NSRC + 63) / 64;
    int sbin[SPL], skind[SPL], sfirst[SPL], slast[SPL];
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        const int s = lane + 64 * r;
        sbin[r] = -1; skind[r] = 0;
        if (s < NSRC) {
            int ic = ib + s % SRC, jc = jb + (s / SRC) % SRC, kc = kb + s / (SRC * SRC);
            int ki, kj, kk; double si, sj, sk;
            if (src_map(ic, g.nx, g.bt[0], g.bt[1], L3[0], ki, si) && src_map(jc, g.ny, g.bt[2], g.bt[3], L3[1], kj, sj) &&
                src_map(kc, g.nz, g.bt[4], g.bt[5], L3[2], kk, sk)) {
                sbin[r] = (int)cidx(g, ic, jc, kc);
                skind[r] = ki | (kj << 2) | (kk << 4);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        sfirst[r] = sbin[r] >= 0 ? a.bin_start[sbin[r]] : 0;
        slast[r] = sbin[r] >= 0 ? a.bin_start[sbin[r] + 1] : 0;
    }
#pragma unroll
    for (int r = 0; r < SPL; ++r) {
        const int s = lane + 64 * r;
        if (s < NSRC) {
            s_gstart[s] = sfirst[r];
            s_kind[s] = skind[r];
            s_off[s + 1] = slast[r] - sfirst[r];
        }
    }
    __syncthreads();
    {
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
    }
    __syncthreads();
    const int total = s_off[NSRC];
    const int li = lane % TL, lj = (lane / TL) % TL, lk = lane / (TL * TL);
    const int i = tx * TL + 1 + li, j = ty * TL + 1 + lj, k = tz * TL + 1 + lk;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int e0 = 0; e0 < total; e0 += CAP) {
        const int e1 = min(total, e0 + CAP);
        // 2. stage entries [e0, e1), one lane per entry
        for (int e = e0 + lane; e < e1; e += 64) {
            int lo = 0, hi = NSRC;   // last source cell with s_off <= e
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= e) lo = mid; else hi = mid - 1;
            }
            const int s = lo;
            const int m = a.bin_items[s_gstart[s] + (e - s_off[s])];
            const int kinds = s_kind[s];
            const int scc[3] = {ib + s % SRC, jb + (s / SRC) % SRC, kb + s / (SRC * SRC)};
            const int el = e - e0;
            const double *fvm = a.fv + 4L * m;
            s_fv[el][0] = fvm[0]; s_fv[el][1] = fvm[1]; s_fv[el][2] = fvm[2];
            if (kinds == 0) {
                const double2 *wm = reinterpret_cast<const double2 *>(a.wts + 24L * m);
                double2 *dst = reinterpret_cast<double2 *>(&s_w[el][0]);
                for (int r = 0; r < 12; ++r) dst[r] = wm[r];
                for (int d = 0; d < 3; ++d) s_cc[el][d] = scc[d];
            } else {
                const double dd[3] = {g.dx, g.dy, g.dz};
                for (int d = 0; d < 3; ++d) {
                    const int kind = (kinds >> (2 * d)) & 3;
                    double xp = a.mx[(long)d * a.M + m];
                    if (kind == 1) xp = xp + (scc[d] < 1 ? -L3[d] : L3[d]);
                    else if (kind == 2) xp = 2.0 * (scc[d] < 1 ? 0.0 : L3[d]) - xp;
                    const int cimg = cell_of(xp, dd[d]);
                    s_cc[el][d] = cimg;
                    for (int o = 0; o < 4; ++o) {
                        s_w[el][d * 8 + o] = wt_face(xp, cimg + o - 1, dd[d]);
                        s_w[el][d * 8 + 4 + o] = wt_cent(xp, cimg + o - 1, dd[d]);
                    }
                }
            }
        }
        __syncthreads();
        // 3. gather for this lane's cell: 16 rows of 4 consecutive source
        // cells, each a contiguous entry range, in (kc,jc,ic,id) order
        for (int sk = lk; sk < lk + 4; ++sk)
            for (int sj = lj; sj < lj + 4; ++sj) {
                const int s0 = (sk * SRC + sj) * SRC + li;
                const int q0 = max(s_off[s0], e0), q1 = min(s_off[s0 + 4], e1);
                for (int q = q0; q < q1; ++q) {
                    const int e = q - e0;
                    const int ox = i - s_cc[e][0] + 1, oy = j - s_cc[e][1] + 1, oz = k - s_cc[e][2] + 1;
                    if (ox < 0 || ox > 3 || oy < 0 || oy > 3 || oz < 0 || oz > 3) continue;
                    const double wu = qdiv<P2>(s_w[e][ox] * s_w[e][12 + oy] * s_w[e][20 + oz], a.dV, a.idV);
                    const double wv = qdiv<P2>(s_w[e][4 + ox] * s_w[e][8 + oy] * s_w[e][20 + oz], a.dV, a.idV);
                    const double ww = qdiv<P2>(s_w[e][4 + ox] * s_w[e][12 + oy] * s_w[e][16 + oz], a.dV, a.idV);
                    f0 = f0 - s_fv[e][0] * wu;
                    f1 = f1 - s_fv[e][1] * wv;
                    f2 = f2 - s_fv[e][2] * ww;
                }
            }
        __syncthreads();
    }
    // 4. multi-direct-forcing update u += dt*(f - f_old) on the lane's cell
    if (i > g.nx || j > g.ny || k > g.nz) return;
    const int tgt[3] = {i, j, k};
    const long slot = (long)blockIdx.x * 64 + lane;
    const long o = cidx(g, i, j, k);
    double *vel[3] = {a.u, a.v, a.w};
    const double fn[3] = {f0, f1, f2};
    for (int c = 0; c < 3; ++c) {
        const double fo = a.fact[c * a.fstride + slot];
        if (fn[c] == 0.0 && fo == 0.0) continue;
        a.fact[c * a.fstride + slot] = fn[c];
        const bool face = tgt[c] == 1 && g.bt[2 * c] == IBMDEM_BC_WALL;
        if (!face) vel[c][o] = vel[c][o] + a.dt * (fn[c] - fo);
    }
}

__global__ __launch_bounds__(64) void k_scatter_force(const int *__restrict__ tiles, TileGeom tg, GridDev g,
                                                      const double *__restrict__ fact, long fstride,
                                                      double *__restrict__ fx, double *__restrict__ fy,
                                                      double *__restrict__ fz) {
    const int tile = tiles[blockIdx.x];
    const int lane = threadIdx.x;
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int i = tx * TL + 1 + lane % TL, j = ty * TL + 1 + (lane / TL) % TL, k = tz * TL + 1 + lane / (TL * TL);
    if (i > g.nx || j > g.ny || k > g.nz) return;
    const long slot = (long)blockIdx.x * 64 + lane, o = cidx(g, i, j, k);
    fx[o] = fact[slot];
    fy[o] = fact[fstride + slot];
    fz[o] = fact[2 * fstride + slot];
}

TileGeom tile_geom(const GridDev &g) { return TileGeom{div_up(g.nx, TL), div_up(g.ny, TL), div_up(g.nz, TL)}; }

}  // namespace

long ibm_tile_count(const GridDev &g) {
    TileGeom t = tile_geom(g);
    return (long)t.ntx * t.nty * t.ntz;
}

int ibm_begin_step(ibmdem_ctx *c) {
    const int M = c->M;
    c->nact = 0;
    c->first_mdf = true;
    if (M == 0) return IBMDEM_OK;
    hipStream_t s = c->stream;
    const TileGeom tg = tile_geom(c->g);
    const long nt = (long)tg.ntx * tg.nty * tg.ntz;
    hipLaunchKernelGGL(k_weights, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mx, c->mcell, c->wts, c->fdrag);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_mark_tiles, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, tg, c->mcell, c->act_flag);
    CHECK_LAUNCH(c);
    const int nb = div_up(nt, CHUNK);
    int *cnt = c->scan_blocks;   // sized >= n/CHUNK + 2 at create
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(T), 0, s, nt, c->act_flag, cnt);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, nb, cnt);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_flag_compact, dim3(nb), dim3(T), 0, s, nt, c->act_flag, cnt, c->act_list);
    CHECK_LAUNCH(c);
    HIPX(c, hipMemcpyAsync(c->flags_host + FL_COUNT, cnt + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(c, hipStreamSynchronize(s));
    const int nact = c->flags_host[FL_COUNT];
    if (nact > c->act_cap) return set_error(c, IBMDEM_ERR_CAPACITY, "spread footprint exceeds its buffer");
    c->nact = nact;
    const long nz = 3L * 64 * std::max(nact, 1);
    hipLaunchKernelGGL(k_zero, dim3(std::max(1, std::min(4096, div_up(nz, T)))), dim3(T), 0, s, nz, c->fact);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_interpolate(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    prof_begin(c, IBMDEM_K_INTERP);
    if (is_pow2(c->dV))
        hipLaunchKernelGGL((k_interp<true>), dim3(div_up(c->M, T)), dim3(T), 0, c->stream, c->M, c->g, c->u[cur],
                           c->v[cur], c->w[cur], c->mcell, c->wts, c->mv, c->fdrag, c->mvol, c->fv, c->dV,
                           1.0 / c->dV, c->cs.dt);
    else
        hipLaunchKernelGGL((k_interp<false>), dim3(div_up(c->M, T)), dim3(T), 0, c->stream, c->M, c->g, c->u[cur],
                           c->v[cur], c->w[cur], c->mcell, c->wts, c->mv, c->fdrag, c->mvol, c->fv, c->dV,
                           1.0 / c->dV, c->cs.dt);
    prof_end(c, IBMDEM_K_INTERP, 0.0);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_spread(ibmdem_ctx *c) {
    if (c->M == 0 || c->nact == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    SpreadArgs a;
    a.g = c->g;
    a.tg = tile_geom(c->g);
    a.M = c->M;
    a.tiles = c->act_list;
    a.bin_start = c->bin_start;
    a.bin_items = c->bin_items;
    a.mx = c->mx;
    a.wts = c->wts;
    a.fv = c->fv;
    a.fact = c->fact;
    a.fstride = 64L * c->nact;
    a.u = c->u[cur]; a.v = c->v[cur]; a.w = c->w[cur];
    a.dV = c->dV;
    a.idV = 1.0 / c->dV;
    a.dt = c->cs.dt;
    prof_begin(c, IBMDEM_K_SPREAD);
    if (is_pow2(c->dV))
        hipLaunchKernelGGL((k_spread_tiles<true>), dim3(c->nact), dim3(64), 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_spread_tiles<false>), dim3(c->nact), dim3(64), 0, c->stream, a);
    prof_end(c, IBMDEM_K_SPREAD, 0.0);
    CHECK_LAUNCH(c);
    c->first_mdf = false;
    return IBMDEM_OK;
}

int ibm_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    const long n = c->n;
    double *buf = nullptr;
    HIPX(c, hipMalloc((void **)&buf, 3 * n * sizeof(double)));
    hipStream_t s = c->stream;
    int rc = IBMDEM_OK;
    if (hipMemsetAsync(buf, 0, 3 * n * sizeof(double), s) != hipSuccess) rc = IBMDEM_ERR_HIP;
    if (!rc && c->nact > 0) {
        hipLaunchKernelGGL(k_scatter_force, dim3(c->nact), dim3(64), 0, s, c->act_list, tile_geom(c->g), c->g,
                           c->fact, 64L * c->nact, buf, buf + n, buf + 2 * n);
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
