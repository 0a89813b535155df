// scan.hpp — integer counting scans shared by the marker binning and the
// IBM tile lists: per-chunk sums, an exclusive scan of the chunk sums in one
// work-group, and the per-chunk exclusive scan into starts. Integer sums are
// exact in any order, so every schedule gives the same starts.
#pragma once

#include "batched.hpp"

namespace scan {

using bk::Blk;
constexpr int SCAN_T = 256;        // threads of the chunk kernels
constexpr int SCAN_CHUNK = 4096;   // counts per chunk

// sum of the counts of each SCAN_CHUNK cells (integers: any order is exact), by
// wavefront shuffles and the SCAN_T/64 wavefront totals
__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ void chunk_sum_body(const Blk &kb, long n, const int *__restrict__ a, int *__restrict__ out) {
    __shared__ int red[SCAN_T / 64];
    const long lo = (long)kb.bx * SCAN_CHUNK;
    int s = 0;
    for (long q = lo + kb.tx; q < min(n, lo + SCAN_CHUNK); q += SCAN_T) s += a[q];
    s = wave_sum_int(s);
    if ((kb.tx & 63) == 0) red[kb.tx >> 6] = s;
    __syncthreads();
    if (kb.tx == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_T / 64; ++w) t += red[w];
        out[kb.bx] = t;
    }
}

// exclusive scan of the nb chunk sums in place (a[nb] = total), one
// work-group: 4096 values per pass, four consecutive per thread, a scan
// across each wavefront by shuffles and the 16 wavefront totals from LDS
__device__ __forceinline__ void scan_blocks_body(const Blk &kb, int nb, int *__restrict__ a) {
    __shared__ int wsum[16];
    const int t = kb.tx, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int base = 0; base < nb; base += 4096) {
        int v[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + 4 * t + q;
            v[q] = idx < nb ? a[idx] : 0;
            s += v[q];
        }
        int x = s;   // inclusive scan of s over the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int before = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int ws = wsum[q];
            before += q < w ? ws : 0;
            tot += ws;
        }
        int run = carry + before + x - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + 4 * t + q;
            if (idx < nb) a[idx] = run;
            run += v[q];
        }
        carry += tot;
        __syncthreads();   // wsum is rewritten by the next pass
    }
    if (t == 0) a[nb] = carry;
}

// exclusive scan of counts into start (SCAN_CHUNK cells per block, 16 per
// thread). Counts and starts cross HBM in coalesced rows through a padded
// LDS tile (element e at e + e/32, so a thread's 16 consecutive elements
// fall in distinct banks across the wavefront).
__device__ __forceinline__ void chunk_scan_body(const Blk &kb, long n, const int *__restrict__ cnt, const int *__restrict__ off,
                                                  int *__restrict__ start, int nb) {
    __shared__ int tile[SCAN_CHUNK + SCAN_CHUNK / 32];
    __shared__ int buf[SCAN_T / 64];
    constexpr int PER = SCAN_CHUNK / SCAN_T;
    const long base = (long)kb.bx * SCAN_CHUNK;
    for (int q = 0; q < PER; ++q) {
        const int e = q * SCAN_T + kb.tx;
        tile[e + (e >> 5)] = base + e < n ? cnt[base + e] : 0;
    }
    __syncthreads();
    int loc[PER];
    int s = 0;
    for (int q = 0; q < PER; ++q) {
        const int e = kb.tx * PER + q;
        loc[q] = tile[e + (e >> 5)];
        s += loc[q];
    }
    // exclusive prefix of the thread totals: a scan across each wavefront
    // by shuffles, then the SCAN_T/64 wavefront totals
    const int lane = kb.tx & 63, wv = kb.tx >> 6;
    int x = s;
#pragma unroll
    for (int st = 1; st < 64; st <<= 1) {
        const int y = __shfl_up(x, st, 64);
        if (lane >= st) x += y;
    }
    if (lane == 63) buf[wv] = x;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < SCAN_T / 64; ++w) before += w < wv ? buf[w] : 0;
    int pos = off[kb.bx] + before + x - s;
    for (int q = 0; q < PER; ++q) {
        const int e = kb.tx * PER + q;
        tile[e + (e >> 5)] = pos;
        pos += loc[q];
    }
    __syncthreads();
    for (int q = 0; q < PER; ++q) {
        const int e = q * SCAN_T + kb.tx;
        if (base + e < n) start[base + e] = tile[e + (e >> 5)];
    }
    if (kb.bx == 0 && kb.tx == 0) start[n] = off[nb];
}

}  // namespace scan
