// batched.hpp — launches that cover many independent cases at once.
//
// A batched launch runs, for each case of a group, the blocks of that case's
// own launch of the same kernel body, with the case's own grid, block shape
// and arguments. The bodies take their block coordinates explicitly (Blk):
// from the hardware in a launch for one case (hw_blk), from batch_for in a
// batched launch. A case's arithmetic is therefore that of its own launch,
// bit for bit. The cases' arguments live in a device arena built once per
// plan (the pointers they hold do not change between steps).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibmdem.h"

namespace bk {

// Block coordinates a kernel body works with.
struct Blk {
    int bx, by, bz;   // block index
    int gx, gy, gz;   // grid dimensions
    int tx, ty;       // thread index in the block
    int dx, dy;       // block dimensions
};
__device__ __forceinline__ Blk hw_blk() {
    return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z,
               (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.x, (int)blockDim.y};
}

struct BGrid {
    int gx, gy, gz;   // the case's grid of blocks
    int dx, dy;       // its block shape: dx a power of two, dx * dy the launch's block size
    int lgdx;         // log2(dx)
    int pb;           // physical blocks of the case in the launch (each takes every pb-th block)
    int ident;        // 1: physical block p is the case's block p (the body maps blocks to XCDs itself)
};

inline BGrid bgrid(dim3 grid, dim3 block) {
    int lg = 0;
    while ((1 << lg) < (int)block.x) ++lg;
    return BGrid{(int)grid.x, (int)grid.y, (int)grid.z, (int)block.x, (int)block.y, lg,
                 (int)(grid.x * grid.y * grid.z), 0};
}

// Pointers read from a launch's tables are generic to the compiler, which
// would then access memory through them with flat instructions (both wait
// counters, 64-bit vector addresses). Device arrays are global memory, and
// the tables and the descriptors they point to are read-only while a kernel
// runs, so they are marked as such: global (AS 1) and constant (AS 4, scalar
// loads where the address is uniform). (Through an integer: a
// generic-to-global-to-generic cast pair would be folded away.)
template <typename T>
__device__ __forceinline__ T *glb(T *p) {
    return (T *)((__attribute__((address_space(1))) T *)(uintptr_t)p);
}
template <typename T>
__device__ __forceinline__ const T *cst(const T *p) {
    return (const T *)((const __attribute__((address_space(4))) T *)(uintptr_t)p);
}

template <class A>
struct BTable {
    const int *first;   // first physical block of each case, ascending
    const A *arg;       // the case's arguments
    int ncase;
};

// Runs f(args, kb) for each of the case's own blocks this block takes. The
// block's case is the number of cases whose first block is at most
// blockIdx.x, less one (a wavefront compares 64 first blocks per round).
// Blocks go to the XCDs round robin (block i to XCD i % 8), and a case with
// 8 or more blocks starts on XCD 0 with a multiple of 8 physical blocks, so
// physical block p runs on XCD p % 8: XCD s takes the s-th of 8 contiguous
// ranges of the case's blocks (whole planes where the case has 8 or more),
// its physical blocks striding through the range together. Planes and their
// z neighbours are then read through one XCD's L2, whatever a plane's block
// count. A barrier separates a physical block's blocks, as bodies reuse
// their LDS.
template <class A, class F>
__device__ __forceinline__ void batch_for(const BTable<A> &tab, F &&f) {
    const int bid = blockIdx.x, lane = threadIdx.x & 63;
    int q = -1;
    for (int c0 = 0; c0 < tab.ncase; c0 += 64) {
        const int cq = c0 + lane;
        const int cnt = __popcll(__ballot(cq < tab.ncase && tab.first[cq] <= bid));
        q += cnt;
        if (cnt < 64) break;
    }
    q = __builtin_amdgcn_readfirstlane(q);
    const A &a = *cst(tab.arg + q);
    const int p = bid - *cst(tab.first + q);
    const BGrid g = a.g;
    if (p >= g.pb) return;
    const int plane = g.gx * g.gy, nb = plane * g.gz;
    int lb0, hi, step;
    if (g.ident) {
        lb0 = p;
        hi = nb;
        step = g.pb;
    } else if (nb >= 8) {
        const int xs = p & 7, per = g.pb >> 3;
        int lo;
        if (g.gz >= 8) {
            lo = (xs * g.gz >> 3) * plane;
            hi = ((xs + 1) * g.gz >> 3) * plane;
        } else {
            lo = (int)((long)xs * nb >> 3);
            hi = (int)((long)(xs + 1) * nb >> 3);
        }
        lb0 = lo + (p >> 3);
        step = per;
    } else {
        lb0 = p;
        hi = nb;
        step = g.pb;
    }
    Blk kb;
    kb.gx = g.gx; kb.gy = g.gy; kb.gz = g.gz;
    kb.dx = g.dx; kb.dy = g.dy;
    kb.tx = threadIdx.x & (g.dx - 1);
    kb.ty = threadIdx.x >> g.lgdx;
    for (int lb = lb0; lb < hi; lb += step) {
        kb.bx = lb % g.gx;
        kb.by = (lb / g.gx) % g.gy;
        kb.bz = lb / plane;
        f(a, kb);
        if (lb + step < hi) __syncthreads();
    }
}

// One batched launch: its kernel (launched by fn with the uniform arguments
// below) and its table in the plan's device arena.
struct BOp;
using BLaunch = void (*)(const BOp &, hipStream_t, int it);
struct BOp {
    BLaunch fn = nullptr;
    size_t first_off = 0, arg_off = 0;   // table offsets in the arena
    const int *first = nullptr;
    const void *arg = nullptr;
    int ncase = 0, blocks = 0;
    int color = 0, slot = 0, op = 0;   // uniform arguments (their meaning is the kernel's)
    bool with_it = false;      // the kernel's iteration argument is the iteration's number
    bool skip_first = false;   // not in iteration 1
    int kind = 0;              // IBMDEM_BK_* (profiling)
    double bytes = 0.0;        // algorithmic HBM bytes of the launch (0: not stated)
    size_t aux_off = SIZE_MAX; // a shared per-case array in the arena (SIZE_MAX: none)
    const void *aux = nullptr;
};
template <class A>
BTable<A> table(const BOp &o) {
    return BTable<A>{o.first, (const A *)o.arg, o.ncase};
}

// Host bytes of the tables, copied to the device once the plan is built.
struct Arena {
    std::vector<char> bytes;
    size_t put(const void *p, size_t n) {
        const size_t off = (bytes.size() + 15) & ~size_t(15);
        bytes.resize(off + n);
        if (n) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};

// One launch of `block` threads per block over the cases with arguments
// args (a case's grid and block shape in args[i].g). A physical block takes
// up to BLOCKS_PER_PHYS of its case's blocks (fewer where the launch would
// otherwise have under 4096 physical blocks), so its start (finding its
// case and loading the case's arguments) is paid once for several. With
// ident, physical block p is the case's block p (every case's blocks start
// on XCD 0, as in its own launch).
constexpr int BLOCKS_PER_PHYS = 8;
template <class A>
int add_op(Arena &ar, std::vector<BOp> &ops, BLaunch fn, int block, std::vector<A> args, BOp o,
           std::string &err, bool ident = false) {
    if (args.empty()) return IBMDEM_OK;
    long total = 0;
    for (const A &a : args) total += (long)a.g.gx * a.g.gy * a.g.gz;
    const int per =
        ident ? 1 : (int)std::max<long>(1, std::min<long>(BLOCKS_PER_PHYS, total / 4096));
    std::vector<int> first(args.size());
    long tot = 0;
    for (size_t i = 0; i < args.size(); ++i) {
        BGrid &g = args[i].g;
        if (g.dx * g.dy != block || (1 << g.lgdx) != g.dx) {
            err = "batched launch: a case's block shape does not fill the launch's blocks";
            return IBMDEM_ERR_ARG;
        }
        const long nb = (long)g.gx * g.gy * g.gz;
        long pb = (nb + per - 1) / per;
        g.ident = ident ? 1 : 0;
        if (ident) {
            tot = (tot + 7) & ~7L;
        } else if (nb >= 8) {
            // the case starts on XCD 0 and its blocks keep their XCD
            pb = (pb + 7) & ~7L;
            tot = (tot + 7) & ~7L;
        }
        g.pb = (int)pb;
        first[i] = (int)tot;
        tot += pb;
        if (tot >= (1L << 31)) {
            err = "batched launch of more than 2^31 blocks";
            return IBMDEM_ERR_ARG;
        }
    }
    o.fn = fn;
    o.ncase = (int)args.size();
    o.blocks = (int)tot;
    o.first_off = ar.put(first.data(), first.size() * sizeof(int));
    o.arg_off = ar.put(args.data(), args.size() * sizeof(A));
    ops.push_back(o);
    return IBMDEM_OK;
}

// A sequence of batched launches with its arena: built on the host, then
// finalize copies the arena to the device and points the ops at it.
struct BatchPlan {
    Arena ar;
    std::vector<BOp> ops;
    char *dev = nullptr;
    int finalize(std::string &err) {
        if (hipMalloc((void **)&dev, std::max<size_t>(ar.bytes.size(), 16)) != hipSuccess ||
            hipMemcpy(dev, ar.bytes.data(), ar.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
            err = "batched plan: device arena allocation or copy failed";
            return IBMDEM_ERR_HIP;
        }
        for (BOp &o : ops) {
            o.first = (const int *)(dev + o.first_off);
            o.arg = dev + o.arg_off;
            if (o.aux_off != SIZE_MAX) o.aux = dev + o.aux_off;
        }
        return IBMDEM_OK;
    }
    // a device pointer into the arena (after finalize)
    template <typename P>
    P *at(size_t off) const { return (P *)(dev + off); }
    void release() {
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        ops.clear();
        ar.bytes.clear();
    }
    int run(hipStream_t s, std::string &err) const {
        for (const BOp &o : ops) {
            o.fn(o, s, 0);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                err = std::string("batched launch: ") + hipGetErrorString(e);
                return IBMDEM_ERR_HIP;
            }
        }
        return IBMDEM_OK;
    }
};

}  // namespace bk
