// ibm.hip — regularized-delta interpolation and force spreading
// (ZJS_VOLUMEFRACTION_DEM.F90:2-67, ZJS_INTERPOLATE_DEM.F90:2-17,
// interpolation.F90:2-33) with the multi-direct-forcing velocity update
// (fluid.F90:135-153) fused into the spread.
//
// Markers do not move during the fluid part of a step, so at the start of a
// step every marker gets its 24 one-dimensional delta factors (per
// dimension: staggered-face and cell-centre variants at cell offsets
// -1..+2) and the spread footprint (interior cells within reach of a marker
// or of one of its periodic/mirror images) is compacted into a sorted list.
//
// Interpolation is marker-parallel and sums the taps in the reference's
// (kk,jj,ii) order, so it is bit-identical to the reference's 125-cell loop
// (zero-weight taps add an exact zero there).
//
// Spreading is a deterministic gather, no float atomics: one thread per
// footprint cell visits the 4x4x4 source cells that can reach it (in kk,jj,ii
// order; markers of a cell in ascending id), including the reference's
// periodic ghost copies (x + coordShift) and wall-mirror copies
// (2*coordMirror - x, fdrag copied unnegated; ZJS_MPI_DEM.F90:173-269),
// which appear as wrapped or mirrored source cells. The result is
// f = 0 - sum fdrag*vol*w per cell (the reference's f(cell) -= ... chain in
// another fixed order), then u += dt*(f - f_old) on the same cell.
#include "common.hpp"

namespace {

constexpr int T = 256;
constexpr int CHUNK = 4096;   // cells per compaction block

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

// weights layout: wts[(dim*8 + variant*4 + o)*M + m], variant 0 face, 1 centre,
// o = 0..3 for cells ci-1..ci+2
__global__ __launch_bounds__(T) void k_weights(int M, GridDev g, const double *__restrict__ mx,
                                               const int *__restrict__ mcell, double *__restrict__ wts,
                                               double *__restrict__ fdrag) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    const int cc[3] = {ci, cj, ck};
    const double dd[3] = {g.dx, g.dy, g.dz};
    for (int d = 0; d < 3; ++d) {
        const double xp = mx[(long)d * M + m];
        for (int o = 0; o < 4; ++o) {
            const int cell = cc[d] + o - 1;
            wts[(long)(d * 8 + o) * M + m] = wt_face(xp, cell, dd[d]);
            wts[(long)(d * 8 + 4 + o) * M + m] = wt_cent(xp, cell, dd[d]);
        }
        fdrag[(long)d * M + m] = 0.0;
    }
}

// UpdateFluidVelocityOnLagrangianPoints: U_p = sum w u dx dy dz,
// fdrag += (U_p - v_p)/dt
__global__ __launch_bounds__(T) void k_interp(int M, GridDev g, const double *__restrict__ u,
                                              const double *__restrict__ v, const double *__restrict__ w,
                                              const int *__restrict__ mcell, const double *__restrict__ wts,
                                              const double *__restrict__ mv, double *__restrict__ fdrag, double dV,
                                              double dt) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    double fx[4], cx[4], fy[4], cy[4], fz[4], cz[4];
    for (int o = 0; o < 4; ++o) {
        fx[o] = wts[(long)(0 + o) * M + m]; cx[o] = wts[(long)(4 + o) * M + m];
        fy[o] = wts[(long)(8 + o) * M + m]; cy[o] = wts[(long)(12 + o) * M + m];
        fz[o] = wts[(long)(16 + o) * M + m]; cz[o] = wts[(long)(20 + o) * M + m];
    }
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    double ulp = 0.0, vlp = 0.0, wlp = 0.0;
    // u: x-face taps -1..2, y/z centre taps -1..1 (nonzero ranges)
    for (int kk = 0; kk < 3; ++kk)
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 4; ++ii) {
                const double wt = fx[ii] * cy[jj] * cz[kk] / dV;
                ulp = ulp + wt * vel_at<0>(u, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    for (int kk = 0; kk < 3; ++kk)
        for (int jj = 0; jj < 4; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const double wt = cx[ii] * fy[jj] * cz[kk] / dV;
                vlp = vlp + wt * vel_at<1>(v, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    for (int kk = 0; kk < 4; ++kk)
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const double wt = cx[ii] * cy[jj] * fz[kk] / dV;
                wlp = wlp + wt * vel_at<2>(w, g, ci + ii - 1, cj + jj - 1, ck + kk - 1) * dx * dy * dz;
            }
    fdrag[m] = fdrag[m] + 1.0 * (ulp - mv[m]) / dt;
    fdrag[(long)M + m] = fdrag[(long)M + m] + 1.0 * (vlp - mv[(long)M + m]) / dt;
    fdrag[2L * M + m] = fdrag[2L * M + m] + 1.0 * (wlp - mv[2L * M + m]) / dt;
}

// mark the interior cells a marker (or an image of it) can reach
__global__ __launch_bounds__(T) void k_mark(int M, GridDev g, const int *__restrict__ mcell,
                                            unsigned char *__restrict__ flag) {
    const int m = blockIdx.x * T + threadIdx.x;
    if (m >= M) return;
    const int cl = mcell[m];
    const int ci = cl % g.nx + 1, cj = (cl / g.nx) % g.ny + 1, ck = cl / (g.nx * g.ny) + 1;
    for (int c = -1; c <= 2; ++c) {
        int k = ck + c;
        if (k < 1 || k > g.nz) { if (g.bt[4] != IBMDEM_BC_PERIODIC) continue; k = k < 1 ? k + g.nz : k - g.nz; }
        for (int b = -1; b <= 2; ++b) {
            int j = cj + b;
            if (j < 1 || j > g.ny) { if (g.bt[2] != IBMDEM_BC_PERIODIC) continue; j = j < 1 ? j + g.ny : j - g.ny; }
            for (int a = -1; a <= 2; ++a) {
                int i = ci + a;
                if (i < 1 || i > g.nx) { if (g.bt[0] != IBMDEM_BC_PERIODIC) continue; i = i < 1 ? i + g.nx : i - g.nx; }
                flag[cidx(g, i, j, k)] = 1;
            }
        }
    }
}

// per-block counts of set flags (block b owns cells [b*CHUNK, (b+1)*CHUNK))
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

// exclusive scan of nb ints in one block; total into out[nb]
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
    constexpr int PER = CHUNK / T;   // 16 consecutive cells per thread
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

// one source-cell coordinate: map to the interior, record the image kind.
// kind 0: direct, 1: periodic (image = x + sh), 2: mirror (image = 2*sh - x)
__device__ __forceinline__ bool src_map(int &i, int n, int lo, int hi, double L, int &kind, double &sh) {
    kind = 0; sh = 0.0;
    if (i < 1) {
        if (lo == IBMDEM_BC_PERIODIC) { i += n; kind = 1; sh = -L; return true; }
        i = 1 - i; kind = 2; sh = 0.0; return i <= n;
    }
    if (i > n) {
        if (hi == IBMDEM_BC_PERIODIC) { i -= n; kind = 1; sh = L; return true; }
        i = 2 * n + 1 - i; kind = 2; sh = L; return i >= 1;
    }
    return true;
}

__global__ __launch_bounds__(T) void k_spread(int nact, const int *__restrict__ act, GridDev g, int M,
                                              const int *__restrict__ bin_start, const int *__restrict__ bin_items,
                                              const double *__restrict__ mx, const int *__restrict__ mcell,
                                              const double *__restrict__ wts, const double *__restrict__ fdrag,
                                              const double *__restrict__ mvol, double *__restrict__ fact, long cap,
                                              double *__restrict__ u, double *__restrict__ v,
                                              double *__restrict__ w, double dV, double dt) {
    const int s = blockIdx.x * T + threadIdx.x;
    if (s >= nact) return;
    const int cl = act[s];
    const int i = cl % g.nx + 1, j = (cl / g.nx) % g.ny + 1, k = cl / (g.nx * g.ny) + 1;
    const int tgt[3] = {i, j, k};
    const double dd[3] = {g.dx, g.dy, g.dz};
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int kc = k - 2; kc <= k + 1; ++kc) {
        int kk = kc, kkind; double ksh;
        if (!src_map(kk, g.nz, g.bt[4], g.bt[5], g.L[2], kkind, ksh)) continue;
        for (int jc = j - 2; jc <= j + 1; ++jc) {
            int jj = jc, jkind; double jsh;
            if (!src_map(jj, g.ny, g.bt[2], g.bt[3], g.L[1], jkind, jsh)) continue;
            for (int ic = i - 2; ic <= i + 1; ++ic) {
                int ii = ic, ikind; double ish;
                if (!src_map(ii, g.nx, g.bt[0], g.bt[1], g.L[0], ikind, ish)) continue;
                const long b = cidx(g, ii, jj, kk);
                const int q0 = bin_start[b], q1 = bin_start[b + 1];
                if (q0 == q1) continue;
                const bool direct = (ikind | jkind | kkind) == 0;
                for (int q = q0; q < q1; ++q) {
                    const int m = bin_items[q];
                    double wf[3][4], wc[3][4];   // [dim][offset]
                    int off[3];
                    if (direct) {
                        off[0] = i - ic + 1; off[1] = j - jc + 1; off[2] = k - kc + 1;
                        for (int d = 0; d < 3; ++d) {
                            wf[d][0] = wts[(long)(d * 8 + off[d]) * M + m];
                            wc[d][0] = wts[(long)(d * 8 + 4 + off[d]) * M + m];
                        }
                    } else {
                        const int kind[3] = {ikind, jkind, kkind};
                        const double sh[3] = {ish, jsh, ksh};
                        bool skip = false;
                        for (int d = 0; d < 3; ++d) {
                            double xp = mx[(long)d * M + m];
                            if (kind[d] == 1) xp = xp + sh[d];
                            else if (kind[d] == 2) xp = 2.0 * sh[d] - xp;
                            const int cimg = cell_of(xp, dd[d]);
                            off[d] = tgt[d] - cimg + 1;
                            if (off[d] < 0 || off[d] > 3) { skip = true; break; }
                            wf[d][0] = wt_face(xp, tgt[d], dd[d]);
                            wc[d][0] = wt_cent(xp, tgt[d], dd[d]);
                        }
                        if (skip) continue;
                    }
                    const double fv0 = fdrag[m] * mvol[m];
                    const double fv1 = fdrag[(long)M + m] * mvol[m];
                    const double fv2 = fdrag[2L * M + m] * mvol[m];
                    const double wu = wf[0][0] * wc[1][0] * wc[2][0] / dV;
                    const double wv = wc[0][0] * wf[1][0] * wc[2][0] / dV;
                    const double ww = wc[0][0] * wc[1][0] * wf[2][0] / dV;
                    f0 = f0 - fv0 * wu;
                    f1 = f1 - fv1 * wv;
                    f2 = f2 - fv2 * ww;
                }
            }
        }
    }
    const long o = cl;
    double *vel[3] = {u, v, w};
    const double fn[3] = {f0, f1, f2};
    for (int c = 0; c < 3; ++c) {
        const double fo = fact[c * cap + s];
        fact[c * cap + s] = fn[c];
        const bool face = tgt[c] == 1 && g.bt[2 * c] == IBMDEM_BC_WALL;
        if (!face) vel[c][o] = vel[c][o] + dt * (fn[c] - fo);
    }
}

__global__ __launch_bounds__(T) void k_scatter_force(int nact, const int *__restrict__ act,
                                                     const double *__restrict__ fact, long cap,
                                                     double *__restrict__ fx, double *__restrict__ fy,
                                                     double *__restrict__ fz) {
    const int s = blockIdx.x * T + threadIdx.x;
    if (s >= nact) return;
    const int cl = act[s];
    fx[cl] = fact[s];
    fy[cl] = fact[cap + s];
    fz[cl] = fact[2 * cap + s];
}

}  // namespace

int ibm_begin_step(ibmdem_ctx *c) {
    const int M = c->M;
    const long n = c->n;
    c->nact = 0;
    c->first_mdf = true;
    if (M == 0) return IBMDEM_OK;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_weights, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mx, c->mcell, c->wts, c->fdrag);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_mark, dim3(div_up(M, T)), dim3(T), 0, s, M, c->g, c->mcell, c->act_flag);
    CHECK_LAUNCH(c);
    const int nb = div_up(n, CHUNK);
    int *cnt = c->scan_blocks;   // sized >= n/CHUNK + 1 at create
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(T), 0, s, n, c->act_flag, cnt);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, nb, cnt);
    CHECK_LAUNCH(c);
    hipLaunchKernelGGL(k_flag_compact, dim3(nb), dim3(T), 0, s, n, c->act_flag, cnt, c->act_list);
    CHECK_LAUNCH(c);
    int nact = 0;
    HIPX(c, hipMemcpyAsync(c->flags_host + FL_COUNT, cnt + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(c, hipStreamSynchronize(s));
    nact = c->flags_host[FL_COUNT];
    if (nact > c->act_cap) return set_error(c, IBMDEM_ERR_CAPACITY, "spread footprint exceeds its buffer");
    c->nact = nact;
    hipLaunchKernelGGL(k_zero, dim3(std::max(1, std::min(4096, div_up(3L * c->act_cap, T)))), dim3(T), 0, s,
                       3L * c->act_cap, c->fact);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_interpolate(ibmdem_ctx *c) {
    if (c->M == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    prof_begin(c, IBMDEM_K_INTERP);
    hipLaunchKernelGGL(k_interp, dim3(div_up(c->M, T)), dim3(T), 0, c->stream, c->M, c->g, c->u[cur], c->v[cur],
                       c->w[cur], c->mcell, c->wts, c->mv, c->fdrag, c->dV, c->cs.dt);
    prof_end(c, IBMDEM_K_INTERP, 0.0);
    CHECK_LAUNCH(c);
    return IBMDEM_OK;
}

int ibm_spread(ibmdem_ctx *c) {
    if (c->M == 0 || c->nact == 0) return IBMDEM_OK;
    const int cur = c->vcur;
    prof_begin(c, IBMDEM_K_SPREAD);
    hipLaunchKernelGGL(k_spread, dim3(div_up(c->nact, T)), dim3(T), 0, c->stream, c->nact, c->act_list, c->g,
                       c->M, c->bin_start, c->bin_items, c->mx, c->mcell, c->wts, c->fdrag, c->mvol, c->fact,
                       c->act_cap, c->u[cur], c->v[cur], c->w[cur], c->dV, c->cs.dt);
    prof_end(c, IBMDEM_K_SPREAD, 0.0);
    CHECK_LAUNCH(c);
    c->first_mdf = false;
    return IBMDEM_OK;
}

int ibm_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    const long n = c->n;
    double *buf = nullptr;
    int rc;
    if ((rc = dev_alloc(c, &buf, 3 * n))) return rc;
    hipStream_t s = c->stream;
    if (c->nact > 0) {
        hipLaunchKernelGGL(k_scatter_force, dim3(div_up(c->nact, T)), dim3(T), 0, s, c->nact, c->act_list, c->fact,
                           c->act_cap, buf, buf + n, buf + 2 * n);
        CHECK_LAUNCH(c);
    }
    double *dst[3] = {fx, fy, fz};
    for (int q = 0; q < 3; ++q)
        if (dst[q]) HIPX(c, hipMemcpyAsync(dst[q], buf + q * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPX(c, hipStreamSynchronize(s));
    // release the temporary
    for (size_t q = 0; q < c->allocs.size(); ++q)
        if (c->allocs[q] == buf) { c->allocs.erase(c->allocs.begin() + q); break; }
    c->bytes -= (int64_t)(3 * n * sizeof(double));
    HIPX(c, hipFree(buf));
    return IBMDEM_OK;
}
