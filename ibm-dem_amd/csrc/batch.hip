// batch.hip — the batched form: many independent cases of uneven grid size
// and particle count advanced together on one GPU.
//
// Each case is a full context with its own HIP stream. A batch step runs the
// three phases of every case's step interleaved (queue the first captured
// graph of all cases, then wait for the solves and queue each case's second
// graph, then collect each case's error flags), so the cases' kernels overlap
// on the GPU and the host never waits for one case while another could be
// queued.
//
// The Poisson solve, most of a step's launches, runs batched: cases that
// share the precision, the smoothing counts and the iteration cap form a
// group whose solve is one captured graph of launches that each cover every
// member (poisson_batch_*), on the group's own stream after the members'
// first parts. The host reads all members' solver scalars at once, runs
// further batched iterations while a member has not converged, and queues
// the members' second parts. A group whose members cannot all run from
// graphs this step (first step, fresh markers, profiling) steps them one by
// one, as ibmdem_batch_set_fused(b, 0) does for every case.
//
// The autotuner picks the multigrid cycle (pre/post smoothing sweeps,
// coarsest-level sweeps) with the shortest measured batch step, then the
// number of solve groups, then, per case, the predictor's tile schedule
// (case.stencil_tile).
#include "common.hpp"
#include "batched.hpp"

#include <algorithm>
#include <chrono>
#include <map>
#include <new>
#include <tuple>
#include <vector>

namespace {

// A batched part of a step for every member of a group, for one parity of
// the members (they share it), as one captured graph: the first part up to
// the solve's right-hand side (keyed by the velocity and pressure
// parities), the forcing loop alone (keyed by the velocity parity), or the
// rest of the step after the solve (gauge, projection, the DEM phase,
// probes; keyed by the velocity and pressure parities and the pending bulk
// shift).
struct PostPlan {
    bk::BatchPlan plan;
    hipGraphExec_t exec = nullptr;
};

struct BatchGroup {
    std::vector<int> member;          // case indices, ascending
    PoissonBatch *pb = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;    // begin + exec_k iterations
    int exec_k = -1;
    bool active = false;              // solves batched this step
    hipEvent_t t0 = nullptr, t1 = nullptr;   // around the solve (GPU time)
    std::map<int, PostPlan> post;     // by parity key (vcur, pcur, bulk shift)
    bool post_ran = false;            // this step's post part ran batched
    std::map<int, PostPlan> mdf;      // by velocity parity
    bool mdf_now = false;             // this step's forcing loop runs batched
    hipEvent_t mdf_done = nullptr;
    std::map<int, PostPlan> pre;      // by velocity and pressure parity
    bool pre_now = false;             // this step's first part runs batched
    bool synced = false;              // no member has work on its own stream that the group stream does not follow
    const int **flag_ptrs = nullptr;  // device: each member's error flags
    int *flags_dev = nullptr, *flags_host = nullptr;   // gathered flags (host: pinned)
};

}  // namespace

struct ibmdem_batch {
    std::vector<ibmdem_ctx *> ctx;
    std::string err;
    bool fused = true;
    bool fused_post = true;          // the post part batched too
    bool fused_mdf = true;           // the forcing loop batched too
    bool fused_pre = true;           // the whole first part batched (IBM set-up, forcing loop, predictor,
                                     // divergence, right-hand side)
    std::vector<BatchGroup> groups;
    std::vector<int32_t> plan_key;   // the cases' parameters the groups were planned with
    std::vector<hipEvent_t> ev;      // per case: end of its step's first part
    std::vector<hipEvent_t> ev_a;    // per case: end of the first part's IBM set-up (batched forcing)
    int split = 1;                   // solve groups per set of shared parameters
    // statistics: group solves run batched; GPU ms of those solves; host ms
    // queueing first parts, waiting for solves, queueing second parts,
    // waiting for the cases' ends
    int64_t fused_solves = 0;
    double ms[5] = {0, 0, 0, 0, 0};
    // profiling of the batched solves, per IBMDEM_BK_* kind
    bool prof = false;
    double prof_ms[9] = {}, prof_bytes[9] = {};
    int64_t prof_n[9] = {};
};

namespace {

int batch_error(ibmdem_batch *b, int q, int rc) {
    if (b && q >= 0 && q < (int)b->ctx.size() && b->ctx[q]) b->err = "case " + std::to_string(q) + ": " + b->ctx[q]->err;
    return rc;
}

// the whole state a step reads, on the host
struct HostCopy {
    std::vector<double> f[4];
    std::vector<double> mx, mv, fd, fc, vol, mass, rad;
    std::vector<int32_t> body;
    std::vector<double> bx, bv, bvold, bom, bL, bLold, brot, brotg, bI, bm, brho;
    std::vector<int32_t> bfixed;
    std::vector<double> plane[3];   // outflow planes (sides 1, 3, 5)
};

ibmdem_markers marker_view(HostCopy &h) {
    return ibmdem_markers{h.mx.data(), h.mv.data(), h.fd.data(), h.fc.data(),
                          h.vol.data(), h.mass.data(), h.rad.data(), h.body.data()};
}
ibmdem_bodies body_view(HostCopy &h) {
    ibmdem_bodies b{};
    b.x = h.bx.data(); b.v = h.bv.data(); b.vold = h.bvold.data();
    b.omega = h.bom.data(); b.angmom = h.bL.data(); b.angmom_old = h.bLold.data();
    b.rot = h.brot.data(); b.rot_glo = h.brotg.data(); b.inertia = h.bI.data();
    b.mass = h.bm.data(); b.rho = h.brho.data(); b.fixed = h.bfixed.data();
    return b;
}

int save_state(ibmdem_ctx *c, HostCopy &h) {
    const size_t n = (size_t)c->n, M = (size_t)c->M, B = (size_t)c->B;
    for (auto &v : h.f) v.resize(n);
    for (auto *v : {&h.mx, &h.mv, &h.fd, &h.fc}) v->resize(3 * M);
    for (auto *v : {&h.vol, &h.mass, &h.rad}) v->resize(M);
    h.body.resize(M);
    for (auto *v : {&h.bx, &h.bv, &h.bvold, &h.bom, &h.bL, &h.bLold}) v->resize(3 * B);
    for (auto *v : {&h.brot, &h.brotg, &h.bI}) v->resize(9 * B);
    h.bm.resize(B); h.brho.resize(B); h.bfixed.resize(B);
    int rc = ibmdem_download_fields(c, h.f[0].data(), h.f[1].data(), h.f[2].data(), h.f[3].data());
    if (rc) return rc;
    if (B) {
        ibmdem_bodies bv = body_view(h);
        if ((rc = ibmdem_download_bodies(c, &bv))) return rc;
    }
    if (M) {
        ibmdem_markers mk = marker_view(h);
        if ((rc = ibmdem_download_markers(c, &mk))) return rc;
    }
    const long pn[3] = {(long)c->g.ny * c->g.nz, (long)c->g.nx * c->g.nz, (long)c->g.nx * c->g.ny};
    for (int d = 0; d < 3; ++d) {
        if (c->g.bt[2 * d + 1] != IBMDEM_BC_OUTFLOW) continue;
        h.plane[d].resize(pn[d]);
        if ((rc = ibmdem_download_outflow(c, 2 * d + 1, h.plane[d].data()))) return rc;
    }
    return IBMDEM_OK;
}

int restore_state(ibmdem_ctx *c, HostCopy &h) {
    int rc;
    if (c->B) {
        ibmdem_bodies bv = body_view(h);
        if ((rc = ibmdem_upload_bodies(c, &bv))) return rc;
    }
    if (c->M) {
        ibmdem_markers mk = marker_view(h);
        if ((rc = ibmdem_upload_markers(c, &mk))) return rc;
    }
    if ((rc = ibmdem_upload_fields(c, h.f[0].data(), h.f[1].data(), h.f[2].data(), h.f[3].data()))) return rc;
    // after the fields: their upload resets the outflow planes
    for (int d = 0; d < 3; ++d)
        if (!h.plane[d].empty() && (rc = ibmdem_upload_outflow(c, 2 * d + 1, h.plane[d].data()))) return rc;
    return ibmdem_sync(c);
}

void free_groups(ibmdem_batch *b) {
    for (BatchGroup &g : b->groups) {
        for (auto *m : {&g.post, &g.mdf, &g.pre})
            for (auto &kv : *m) {
                if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
                kv.second.plan.release();
            }
        if (g.mdf_done) (void)hipEventDestroy(g.mdf_done);
        if (g.flag_ptrs) (void)hipFree(g.flag_ptrs);
        if (g.flags_dev) (void)hipFree(g.flags_dev);
        if (g.flags_host) (void)hipHostFree(g.flags_host);
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.stream) (void)hipStreamDestroy(g.stream);
        if (g.t0) (void)hipEventDestroy(g.t0);
        if (g.t1) (void)hipEventDestroy(g.t1);
        poisson_batch_destroy(g.pb);
    }
    b->groups.clear();
    b->plan_key.clear();
}

// a group failure is each member's failure, with the same message
int group_error(ibmdem_batch *b, const BatchGroup &g, int rc, const std::string &msg) {
    for (int q : g.member) set_error(b->ctx[q], rc, msg);
    return rc;
}

// Groups of the cases whose solve can run batched, rebuilt when a case's
// cycle, precision or iteration cap, or the split, has changed since they
// were planned. The cases that share those parameters are split into
// b->split groups of about equal cell count (largest case first to the
// lightest group), so one group's solve overlaps other groups' step parts.
int plan_groups(ibmdem_batch *b) {
    const int n = (int)b->ctx.size();
    std::vector<int32_t> key{b->split};
    for (ibmdem_ctx *c : b->ctx)
        key.insert(key.end(), {poisson_batchable(c) ? 1 : 0, c->cs.precision, c->cs.mg_pre, c->cs.mg_post,
                               c->cs.mg_coarse_sweeps, c->cs.poisson_max_iter, c->cs.couple_interval,
                               c->cs.n_mdf, (int32_t)c->graphs.epoch});
    if (key == b->plan_key) return IBMDEM_OK;
    free_groups(b);
    std::map<std::tuple<int, int, int, int, int, int>, std::vector<int>> by;
    for (int q = 0; q < n; ++q) {
        const ibmdem_ctx *c = b->ctx[q];
        if (poisson_batchable(c))
            by[std::make_tuple(c->cs.precision, c->cs.mg_pre, c->cs.mg_post, c->cs.poisson_max_iter,
                               c->cs.couple_interval, c->cs.n_mdf)]
                .push_back(q);
    }
    std::vector<std::vector<int>> parts;
    for (auto &kv : by) {
        const int k = std::min<int>(b->split, (int)kv.second.size());
        std::vector<int> order = kv.second;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->ctx[x]->n > b->ctx[y]->n; });
        std::vector<std::vector<int>> part(k);
        std::vector<long> load(k, 0);
        for (int q : order) {
            const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            part[j].push_back(q);
            load[j] += b->ctx[q]->n;
        }
        for (auto &pm : part) {
            std::sort(pm.begin(), pm.end());
            parts.push_back(pm);
        }
    }
    for (auto &pm : parts) {
        b->groups.emplace_back();
        BatchGroup &g = b->groups.back();
        g.member = pm;
        std::vector<ibmdem_ctx *> cs;
        for (int q : g.member) cs.push_back(b->ctx[q]);
        std::string err;
        int rc = poisson_batch_create(cs, &g.pb, err);
        if (!rc) poisson_batch_set_profile(g.pb, b->prof);
        if (!rc && (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreate(&g.t0) != hipSuccess || hipEventCreate(&g.t1) != hipSuccess ||
                    hipEventCreateWithFlags(&g.mdf_done, hipEventDisableTiming) != hipSuccess)) {
            rc = IBMDEM_ERR_HIP;
            err = "batch stream or event creation failed";
        }
        if (!rc) {
            std::vector<const int *> fp;
            for (ibmdem_ctx *c : cs) fp.push_back(c->flags);
            const size_t nf = cs.size() * FL_COUNT * sizeof(int);
            if (hipMalloc((void **)&g.flag_ptrs, fp.size() * sizeof(int *)) != hipSuccess ||
                hipMemcpy(g.flag_ptrs, fp.data(), fp.size() * sizeof(int *), hipMemcpyHostToDevice) != hipSuccess ||
                hipMalloc((void **)&g.flags_dev, nf) != hipSuccess ||
                hipHostMalloc((void **)&g.flags_host, nf, hipHostMallocDefault) != hipSuccess) {
                rc = IBMDEM_ERR_HIP;
                err = "batch flag buffers: allocation failed";
            }
        }
        if (rc) {
            b->err = err;
            free_groups(b);
            return rc;
        }
    }
    for (auto *evs : {&b->ev, &b->ev_a})
        for (int q = (int)evs->size(); q < n; ++q) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                b->err = "event creation failed";
                free_groups(b);
                return IBMDEM_ERR_HIP;
            }
            evs->push_back(e);
        }
    b->plan_key = key;
    return IBMDEM_OK;
}

// the group's solve as one graph: the begin part and k iterations
int capture_group(ibmdem_batch *b, BatchGroup &g, int k) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.exec_k = -1;
    hipGraph_t graph = nullptr;
    std::string err;
    if (hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched solve capture failed to begin");
    int rc = poisson_batch_begin(g.pb, g.stream, err);
    for (int it = 1; it <= k && !rc; ++it) rc = poisson_batch_iteration(g.pb, g.stream, it, err);
    const hipError_t e = hipStreamEndCapture(g.stream, &graph);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return group_error(b, g, rc ? rc : IBMDEM_ERR_HIP,
                           rc ? err : std::string("batched solve capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        g.exec = nullptr;
        return group_error(b, g, IBMDEM_ERR_HIP, std::string("batched solve instantiation: ") + hipGetErrorString(ei));
    }
    g.exec_k = k;
    return IBMDEM_OK;
}

// Queue the group's solve after its members' first parts: k = the largest
// iteration count a member has needed, then the scalars to the host.
int group_solve_queue(ibmdem_batch *b, BatchGroup &g) {
    int k = 0;
    for (int q : g.member) {
        // a first part the group ran is already on its stream
        if (!g.pre_now && hipStreamWaitEvent(g.stream, b->ev[q], 0) != hipSuccess)
            return group_error(b, g, IBMDEM_ERR_HIP, "batched solve: stream wait failed");
        k = std::max(k, b->ctx[q]->graphs.cg_k);
    }
    int rc;
    std::string err;
    // t0 completes once every member's first part has
    if (hipEventRecord(g.t0, g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched solve event record failed");
    if (b->prof) {
        // profiled: the same launches queued one by one, an event after each
        poisson_batch_profile_arm(g.pb, true);
        rc = poisson_batch_begin(g.pb, g.stream, err);
        for (int it = 1; it <= k && !rc; ++it) rc = poisson_batch_iteration(g.pb, g.stream, it, err);
        poisson_batch_profile_arm(g.pb, false);
        if (rc) return group_error(b, g, rc, err);
        g.exec_k = k;
    } else {
        if (!g.exec || g.exec_k != k)
            if ((rc = capture_group(b, g, k))) return rc;
        if (hipGraphLaunch(g.exec, g.stream) != hipSuccess)
            return group_error(b, g, IBMDEM_ERR_HIP, "batched solve graph launch failed");
    }
    if ((rc = poisson_batch_gather(g.pb, g.stream, err))) return group_error(b, g, rc, err);
    if (hipEventRecord(g.t1, g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched solve event record failed");
    b->fused_solves++;
    return IBMDEM_OK;
}

// Wait for the group's solve; while a member has not converged, one more
// batched iteration (a member that has converged skips it), up to the cap.
int group_solve_finish(ibmdem_batch *b, BatchGroup &g) {
    std::string err;
    int it = g.exec_k;
    const int maxit = b->ctx[g.member[0]]->cs.poisson_max_iter;
    bool first = true;
    while (true) {
        if (hipStreamSynchronize(g.stream) != hipSuccess)
            return group_error(b, g, IBMDEM_ERR_HIP, "batched solve: stream synchronization failed");
        if (first) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, g.t0, g.t1) == hipSuccess) b->ms[0] += ms;
            if (poisson_batch_profile_collect(g.pb, b->prof_ms, b->prof_n, b->prof_bytes))
                return group_error(b, g, IBMDEM_ERR_HIP, "batched solve profiling: event times unavailable");
            first = false;
        }
        // a member with a non-finite right-hand side or residual fails its
        // check whatever further iterations do, so it does not hold the
        // others' loop
        bool all = true;
        for (size_t i = 0; i < g.member.size(); ++i) {
            const double *sc = poisson_batch_scalars(g.pb, (int)i);
            const bool finite = sc[SC_BNORM2] - sc[SC_BNORM2] == 0.0 && sc[SC_RR] == sc[SC_RR];
            all = all && (sc[SC_CONV] != 0.0 || !finite);
        }
        if (all || it >= maxit) return IBMDEM_OK;
        ++it;
        int rc = poisson_batch_iteration(g.pb, g.stream, it, err);
        if (!rc) rc = poisson_batch_gather(g.pb, g.stream, err);
        if (rc) return group_error(b, g, rc, err);
    }
}

// every member's error flags into one array (member i at i * FL_COUNT)
__global__ __launch_bounds__(64) void k_gather_flags(const int *const *__restrict__ src, int n, int *__restrict__ out) {
    const int i = blockIdx.x;
    if (i < n && threadIdx.x < FL_COUNT) out[i * FL_COUNT + threadIdx.x] = src[i][threadIdx.x];
}

// the members' parity key, or -1 where they differ
int parity_key(const ibmdem_batch *b, const BatchGroup &g) {
    int key = -1;
    for (int q : g.member) {
        const ibmdem_ctx *c = b->ctx[q];
        const int k = c->vcur | (c->pcur << 1) | ((c->bulk_pending ? 1 : 0) << 2);
        if (key >= 0 && k != key) return -1;
        key = k;
    }
    return key;
}

// A batched part for the members' current parity (key in `plans`): the
// plan that fill builds and its captured graph, built at the first use of
// that parity.
template <class F>
int graph_plan(ibmdem_batch *b, BatchGroup &g, std::map<int, PostPlan> &plans, int key, F fill, PostPlan **out) {
    auto it = plans.find(key);
    if (it != plans.end()) {
        *out = &it->second;
        return IBMDEM_OK;
    }
    PostPlan &pp = plans[key];
    std::vector<ibmdem_ctx *> cs;
    for (int q : g.member) cs.push_back(b->ctx[q]);
    std::string err;
    bk::BatchPlan &P = pp.plan;
    int rc = fill(P, cs, err);
    if (!rc) rc = P.finalize(err);
    hipGraph_t graph = nullptr;
    if (!rc && hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = IBMDEM_ERR_HIP;
        err = "batched post part: capture failed to begin";
    } else if (!rc) {
        rc = P.run(g.stream, err);
        const hipError_t e = hipStreamEndCapture(g.stream, &graph);
        if (!rc && e != hipSuccess) {
            rc = IBMDEM_ERR_HIP;
            err = std::string("batched post part capture: ") + hipGetErrorString(e);
        }
        if (!rc && hipGraphInstantiate(&pp.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            pp.exec = nullptr;
            rc = IBMDEM_ERR_HIP;
            err = "batched post part: graph instantiation failed";
        }
        if (graph) (void)hipGraphDestroy(graph);
    }
    if (rc) {
        if (pp.exec) (void)hipGraphExecDestroy(pp.exec);
        pp.plan.release();
        plans.erase(key);
        return group_error(b, g, rc, err);
    }
    *out = &pp;
    return IBMDEM_OK;
}

// the launches of step_post for every member, in step_post's order
int fill_post(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    const size_t cases = dem_plan_cases(P, cs);
    const int ci = cs[0]->cs.couple_interval;
    int rc = poisson_plan_gauge(P, cs, err);
    if (!rc) rc = fluid_plan_project(P, cs, err);
    if (!rc) rc = dem_plan_body_hydro(P, cases, cs, err);
    for (int s = 0; s < ci && !rc; ++s) rc = dem_plan_substep(P, cases, cs, s == 0, s == ci - 1, err);
    if (!rc) rc = dem_plan_probe(P, cases, cs, err);
    return rc;
}

// the launches of step_pre_rhs for every member, in its order; the
// divergence and the right-hand side are planned at the host state the
// predictor leaves
int fill_pre(bk::BatchPlan &P, const std::vector<ibmdem_ctx *> &cs, std::string &err) {
    int rc = ibm_plan_begin_step(P, cs, err);
    if (!rc) rc = ibm_plan_mdf(P, cs, err);
    if (!rc) rc = fluid_plan_predictor(P, cs, err);
    if (rc) return rc;
    std::vector<HostState> before;
    for (ibmdem_ctx *c : cs) {
        before.push_back(host_state(c));
        step_pre_rhs_batched_state(c);
    }
    rc = fluid_plan_divergence(P, cs, err);
    if (!rc) rc = poisson_plan_rhs(P, cs, err);
    for (size_t i = 0; i < cs.size(); ++i) set_host_state(cs[i], before[i]);
    return rc;
}

// the velocity and pressure parities the members share, or -1
int pre_key(const ibmdem_batch *b, const BatchGroup &g) {
    int key = -1;
    for (int q : g.member) {
        const ibmdem_ctx *c = b->ctx[q];
        const int k = c->vcur | (c->pcur << 1);
        if (key >= 0 && k != key) return -1;
        key = k;
    }
    return key;
}

// Queue the group's first part on its stream (after any work still queued
// on the members' own streams); the members' host state moves on as
// step_pre_rhs would move it.
int group_pre_queue(ibmdem_batch *b, BatchGroup &g, int key) {
    PostPlan *pp = nullptr;
    int rc;
    if ((rc = graph_plan(b, g, g.pre, key, fill_pre, &pp))) return rc;
    if (!g.synced) {
        for (int q : g.member)
            if (hipEventRecord(b->ev_a[q], b->ctx[q]->stream) != hipSuccess ||
                hipStreamWaitEvent(g.stream, b->ev_a[q], 0) != hipSuccess)
                return group_error(b, g, IBMDEM_ERR_HIP, "batched first part: stream wait failed");
        g.synced = true;
    }
    if (hipGraphLaunch(pp->exec, g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched first part: graph launch failed");
    for (int q : g.member) step_pre_rhs_batched_state(b->ctx[q]);
    return IBMDEM_OK;
}

// the velocity parity the members share, or -1
int velocity_key(const ibmdem_batch *b, const BatchGroup &g) {
    int key = -1;
    for (int q : g.member) {
        const int k = b->ctx[q]->vcur;
        if (key >= 0 && k != key) return -1;
        key = k;
    }
    return key;
}

// Queue the group's forcing loop after its members' IBM set-up; the members'
// host state moves on as ibm_spread would move it.
int group_mdf_queue(ibmdem_batch *b, BatchGroup &g, int key) {
    PostPlan *pp = nullptr;
    int rc;
    if ((rc = graph_plan(b, g, g.mdf, key, ibm_plan_mdf, &pp))) return rc;
    for (int q : g.member)
        if (hipStreamWaitEvent(g.stream, b->ev_a[q], 0) != hipSuccess)
            return group_error(b, g, IBMDEM_ERR_HIP, "batched forcing: stream wait failed");
    if (hipGraphLaunch(pp->exec, g.stream) != hipSuccess || hipEventRecord(g.mdf_done, g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched forcing: graph launch failed");
    for (int q : g.member) step_mdf_batched_state(b->ctx[q]);
    return IBMDEM_OK;
}

// Queue the group's post part (all members converged), then its flags to
// the host; the members' host state moves on as step_post would move it.
int group_post_queue(ibmdem_batch *b, BatchGroup &g, int key) {
    PostPlan *pp = nullptr;
    int rc;
    if ((rc = graph_plan(b, g, g.post, key, fill_post, &pp))) return rc;
    const int n = (int)g.member.size();
    if (hipGraphLaunch(pp->exec, g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched post part: graph launch failed");
    hipLaunchKernelGGL(k_gather_flags, dim3(n), dim3(64), 0, g.stream, g.flag_ptrs, n, g.flags_dev);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(g.flags_host, g.flags_dev, (size_t)n * FL_COUNT * sizeof(int), hipMemcpyDeviceToHost,
                       g.stream) != hipSuccess)
        return group_error(b, g, IBMDEM_ERR_HIP, "batched post part: flag copy failed");
    for (int q : g.member) step_post_batched_state(b->ctx[q]);
    return IBMDEM_OK;
}

// `n_steps` batch steps. A case that fails stops there (its context keeps
// its own message, as a single context's failed step does); every other case
// still completes the step in progress, so no context is left between its
// two step graphs. Returns the first failure (lowest case index).
int run_steps(ibmdem_batch *b, int n_steps, ibmdem_solve_info *inf) {
    const int n = (int)b->ctx.size();
    int rc;
    if (b->fused) {
        if ((rc = plan_groups(b))) return rc;
    } else {
        free_groups(b);
    }
    std::vector<char> done(n), failed(n, 0), batched(n, 0);
    int first_rc = IBMDEM_OK, first_q = -1;
    // the members' streams may hold work queued outside this run
    for (BatchGroup &g : b->groups) g.synced = false;
    auto fail = [&](int q, int r) {
        failed[q] = 1;
        if (first_q < 0 || q < first_q) { first_q = q; first_rc = r; }
    };
    using clk = std::chrono::steady_clock;
    auto lap = [&](clk::time_point &t, int slot) {
        const clk::time_point now = clk::now();
        b->ms[slot] += std::chrono::duration<double, std::milli>(now - t).count();
        t = now;
    };
    for (int s = 0; s < n_steps && first_q < 0; ++s) {
        clk::time_point t = clk::now();
        std::fill(batched.begin(), batched.end(), 0);
        for (BatchGroup &g : b->groups) {
            g.active = true;
            for (int q : g.member) g.active = g.active && step_batchable(b->ctx[q]);
            if (g.active)
                for (int q : g.member) batched[q] = 1;
        }
        // the whole first part batched where the members share their
        // parities, else the forcing loop alone where they share their
        // velocity parity
        std::vector<char> mdf(n, 0), pre(n, 0);
        for (BatchGroup &g : b->groups) {
            g.pre_now = g.active && b->fused_pre && pre_key(b, g) >= 0;
            g.mdf_now = g.active && !g.pre_now && b->fused_mdf && velocity_key(b, g) >= 0;
            if (!g.pre_now) g.synced = false;   // members queue on their own streams
            for (int q : g.member) {
                pre[q] = g.pre_now;
                mdf[q] = g.mdf_now;
            }
        }
        for (BatchGroup &g : b->groups) {
            if (!g.pre_now) continue;
            if ((rc = group_pre_queue(b, g, pre_key(b, g)))) {
                for (int q : g.member) if (!failed[q]) fail(q, rc);
                g.pre_now = false;
                g.active = false;
            }
        }
        for (int q = 0; q < n; ++q) {
            ibmdem_ctx *c = b->ctx[q];
            bool d = false;
            if (pre[q]) {
                // queued with its group above
            } else if (mdf[q]) {
                rc = step_phase1a_batched(c);
                if (!rc && hipEventRecord(b->ev_a[q], c->stream) != hipSuccess)
                    rc = set_error(c, IBMDEM_ERR_HIP, "event record failed");
            } else if (batched[q]) {
                rc = step_phase1_batched(c);
                if (!rc && hipEventRecord(b->ev[q], c->stream) != hipSuccess)
                    rc = set_error(c, IBMDEM_ERR_HIP, "event record failed");
            } else {
                rc = step_phase1(c, &inf[q], &d);
            }
            done[q] = d;
            if (rc) fail(q, rc);
        }
        for (BatchGroup &g : b->groups) {
            if (!g.mdf_now) continue;
            if ((rc = group_mdf_queue(b, g, velocity_key(b, g)))) {
                for (int q : g.member) if (!failed[q]) fail(q, rc);
                g.mdf_now = false;
                g.active = false;
                continue;
            }
            for (int q : g.member) {
                if (failed[q]) continue;
                ibmdem_ctx *c = b->ctx[q];
                if (hipStreamWaitEvent(c->stream, g.mdf_done, 0) != hipSuccess)
                    rc = set_error(c, IBMDEM_ERR_HIP, "stream wait failed");
                else rc = step_phase1b_batched(c);
                if (!rc && hipEventRecord(b->ev[q], c->stream) != hipSuccess)
                    rc = set_error(c, IBMDEM_ERR_HIP, "event record failed");
                if (rc) fail(q, rc);
            }
        }
        for (BatchGroup &g : b->groups) {
            if (!g.active) continue;
            if ((rc = group_solve_queue(b, g))) {
                for (int q : g.member) if (!failed[q]) fail(q, rc);
                g.active = false;
            }
        }
        lap(t, 1);
        for (BatchGroup &g : b->groups) {
            g.post_ran = false;
            if (!g.active) continue;
            rc = group_solve_finish(b, g);
            lap(t, 2);
            if (rc) {
                for (int q : g.member) if (!failed[q]) fail(q, rc);
                continue;
            }
            bool all = true;
            for (size_t i = 0; i < g.member.size(); ++i) {
                const int q = g.member[i];
                if (!failed[q] && (rc = step_check_batched(b->ctx[q], poisson_batch_scalars(g.pb, (int)i), &inf[q])))
                    fail(q, rc);
                all = all && !failed[q];
            }
            // the post part batched when every member goes on and they
            // share their parity; else on each member's own stream
            const int key = all && b->fused_post ? parity_key(b, g) : -1;
            if (key >= 0) {
                if ((rc = group_post_queue(b, g, key))) {
                    for (int q : g.member) if (!failed[q]) fail(q, rc);
                } else {
                    g.post_ran = true;
                }
            } else {
                g.synced = false;
                for (int q : g.member)
                    if (!failed[q] && (rc = step_post_own(b->ctx[q]))) fail(q, rc);
            }
            lap(t, 3);
        }
        for (int q = 0; q < n; ++q) {
            if (batched[q] || done[q] || failed[q]) continue;
            if ((rc = step_phase2(b->ctx[q], &inf[q]))) fail(q, rc);
        }
        lap(t, 3);
        std::fill(batched.begin(), batched.end(), 0);
        for (BatchGroup &g : b->groups) {
            if (!g.post_ran) continue;
            for (int q : g.member) batched[q] = 1;
            if (hipStreamSynchronize(g.stream) != hipSuccess) {
                group_error(b, g, IBMDEM_ERR_HIP, "batched post part: stream synchronization failed");
                for (int q : g.member) if (!failed[q]) fail(q, IBMDEM_ERR_HIP);
                continue;
            }
            for (size_t i = 0; i < g.member.size(); ++i) {
                const int q = g.member[i];
                if (!failed[q] && (rc = dem_flags_status(b->ctx[q], g.flags_host + i * FL_COUNT))) fail(q, rc);
            }
        }
        for (int q = 0; q < n; ++q) {
            if (batched[q] || done[q] || failed[q]) continue;
            if ((rc = step_phase3(b->ctx[q]))) fail(q, rc);
        }
        lap(t, 4);
    }
    return first_q < 0 ? IBMDEM_OK : batch_error(b, first_q, first_rc);
}

// seconds per step of `steps` graph-replayed steps after one warm-up step
int time_steps(ibmdem_ctx *c, int steps, double *sec) {
    int rc = ibmdem_step(c, nullptr);   // eager: learns the iteration count
    if (rc) return rc;
    if ((rc = ibmdem_step(c, nullptr))) return rc;   // captures the graphs
    if ((rc = ibmdem_sync(c))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s)
        if ((rc = ibmdem_step(c, nullptr))) return rc;
    if ((rc = ibmdem_sync(c))) return rc;
    *sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / steps;
    return IBMDEM_OK;
}

}  // namespace

extern "C" {

int ibmdem_get_mg_params(ibmdem_ctx *c, int32_t *mg) {
    if (!c || !mg) return IBMDEM_ERR_ARG;
    mg[0] = c->cs.mg_pre;
    mg[1] = c->cs.mg_post;
    mg[2] = c->cs.mg_coarse_sweeps;
    return IBMDEM_OK;
}

int ibmdem_set_mg_params(ibmdem_ctx *c, const int32_t *mg) {
    if (!c || !mg) return IBMDEM_ERR_ARG;
    if (mg[0] < 1 || mg[0] > 8 || mg[1] < 1 || mg[1] > 8 || mg[2] < 1 || mg[2] > 256)
        return set_error(c, IBMDEM_ERR_ARG, "multigrid sweeps out of range (pre/post 1..8, coarse 1..256)");
    int rc = ibmdem_sync(c);
    if (rc) return rc;
    c->cs.mg_pre = mg[0];
    c->cs.mg_post = mg[1];
    c->cs.mg_coarse_sweeps = mg[2];
    // captured graphs hold the old cycle; the iteration count changes too
    drop_graphs(c);
    c->graphs.cg_k = -1;
    return IBMDEM_OK;
}

int ibmdem_get_stencil_tile(ibmdem_ctx *c, int32_t *tile) {
    if (!c || !tile) return IBMDEM_ERR_ARG;
    *tile = c->cs.stencil_tile;
    return IBMDEM_OK;
}

int ibmdem_set_stencil_tile(ibmdem_ctx *c, int32_t tile) {
    if (!c) return IBMDEM_ERR_ARG;
    if (tile < 0 || tile > 3) return set_error(c, IBMDEM_ERR_ARG, "stencil_tile must be 0..3");
    int rc = ibmdem_sync(c);
    if (rc) return rc;
    c->cs.stencil_tile = tile;
    drop_graphs(c);   // captured graphs hold the old schedule
    return IBMDEM_OK;
}

int ibmdem_batch_create(int32_t n, const ibmdem_case *cases, ibmdem_batch **out) {
    if (!out || !cases || n < 1) return IBMDEM_ERR_ARG;
    *out = nullptr;
    ibmdem_batch *b = new (std::nothrow) ibmdem_batch();
    if (!b) return IBMDEM_ERR_ALLOC;
    b->ctx.assign(n, nullptr);
    for (int q = 0; q < n; ++q) {
        const int rc = ibmdem_create(&cases[q], &b->ctx[q]);
        if (rc) {
            ibmdem_batch_destroy(b);
            return rc;
        }
    }
    *out = b;
    return IBMDEM_OK;
}

int ibmdem_batch_destroy(ibmdem_batch *b) {
    if (!b) return IBMDEM_OK;
    for (BatchGroup &g : b->groups)
        if (g.stream) (void)hipStreamSynchronize(g.stream);
    free_groups(b);
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ev_a) (void)hipEventDestroy(e);
    for (ibmdem_ctx *c : b->ctx) ibmdem_destroy(c);
    delete b;
    return IBMDEM_OK;
}

int ibmdem_batch_context(ibmdem_batch *b, int32_t q, ibmdem_ctx **ctx) {
    if (!b || !ctx || q < 0 || q >= (int)b->ctx.size()) return IBMDEM_ERR_ARG;
    *ctx = b->ctx[q];
    return IBMDEM_OK;
}

const char *ibmdem_batch_last_error(const ibmdem_batch *b) { return b ? b->err.c_str() : "no batch"; }

int ibmdem_batch_run(ibmdem_batch *b, int32_t n_steps, ibmdem_solve_info *info) {
    if (!b || n_steps < 0) return IBMDEM_ERR_ARG;
    std::vector<ibmdem_solve_info> tmp(b->ctx.size());
    return run_steps(b, n_steps, info ? info : tmp.data());
}

int ibmdem_batch_set_fused(ibmdem_batch *b, int32_t on) {
    if (!b || on < 0 || on > 4) return IBMDEM_ERR_ARG;
    for (BatchGroup &g : b->groups)
        if (g.stream) (void)hipStreamSynchronize(g.stream);
    b->fused = on != 0;
    b->fused_post = on == 1 || on == 3 || on == 4;
    b->fused_mdf = on == 1 || on == 4;
    b->fused_pre = on == 1;
    if (!b->fused) free_groups(b);
    return IBMDEM_OK;
}

int ibmdem_batch_set_solve_groups(ibmdem_batch *b, int32_t groups) {
    if (!b) return IBMDEM_ERR_ARG;
    if (groups < 1 || groups > 64) {
        b->err = "solve groups must be 1..64";
        return IBMDEM_ERR_ARG;
    }
    for (BatchGroup &g : b->groups)
        if (g.stream) (void)hipStreamSynchronize(g.stream);
    b->split = groups;   // the groups are re-planned at the next run
    return IBMDEM_OK;
}

int ibmdem_batch_get_solve_groups(const ibmdem_batch *b, int32_t *groups) {
    if (!b || !groups) return IBMDEM_ERR_ARG;
    *groups = b->split;
    return IBMDEM_OK;
}

int ibmdem_batch_profile(ibmdem_batch *b, int32_t on) {
    if (!b) return IBMDEM_ERR_ARG;
    b->prof = on != 0;
    for (BatchGroup &g : b->groups) {
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        poisson_batch_set_profile(g.pb, b->prof);
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.exec_k = -1;
    }
    return IBMDEM_OK;
}

int ibmdem_batch_profile_read(const ibmdem_batch *b, int32_t kind, double *total_ms, int64_t *launches,
                              double *bytes) {
    if (!b || kind < 1 || kind > 8 || !total_ms || !launches || !bytes) return IBMDEM_ERR_ARG;
    *total_ms = b->prof_ms[kind];
    *launches = b->prof_n[kind];
    *bytes = b->prof_bytes[kind];
    return IBMDEM_OK;
}

int ibmdem_batch_stats(const ibmdem_batch *b, int64_t *fused_solves, double *ms) {
    if (!b) return IBMDEM_ERR_ARG;
    if (fused_solves) *fused_solves = b->fused_solves;
    if (ms)
        for (int i = 0; i < 5; ++i) ms[i] = b->ms[i];
    return IBMDEM_OK;
}

int ibmdem_batch_autotune(ibmdem_batch *b, int32_t trial_steps) {
    if (!b) return IBMDEM_ERR_ARG;
    const int steps = trial_steps > 0 ? trial_steps : 3;
    const int n = (int)b->ctx.size();
    // candidate cycles: pre, post, coarsest-level sweeps
    static const int32_t cand[][3] = {{1, 1, 24}, {2, 2, 24}, {3, 3, 24}, {2, 2, 12}, {1, 2, 24}};
    const int ncand = (int)(sizeof cand / sizeof cand[0]);
    std::vector<HostCopy> h(n);
    std::vector<int32_t> orig(3 * n);
    int rc;
    for (int q = 0; q < n; ++q) {
        if ((rc = save_state(b->ctx[q], h[q]))) return batch_error(b, q, rc);
        ibmdem_get_mg_params(b->ctx[q], &orig[3 * q]);
    }
    auto restore_all = [&]() {
        for (int q = 0; q < n; ++q) {
            const int r = restore_state(b->ctx[q], h[q]);
            if (r) return batch_error(b, q, r);
        }
        return (int)IBMDEM_OK;
    };
    // The cycle, for the whole batch (its solve runs batched): two steps
    // under the candidate (the first learns the iteration counts, the
    // second captures the graphs), then `steps` timed batch steps. A cycle
    // under which a case fails (e.g. no convergence within the cap) is not a
    // candidate.
    std::vector<ibmdem_solve_info> info(n);
    double best = 1e300;
    int best_i = -1;
    for (int i = 0; i < ncand; ++i) {
        for (int q = 0; q < n; ++q)
            if ((rc = ibmdem_set_mg_params(b->ctx[q], cand[i]))) return batch_error(b, q, rc);
        double sec = 0.0;
        rc = run_steps(b, 2, info.data());
        if (rc == IBMDEM_OK) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = run_steps(b, steps, info.data());
            sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / steps;
        }
        if (rc == IBMDEM_OK && sec < best) { best = sec; best_i = i; }
        if ((rc = restore_all())) return rc;
    }
    for (int q = 0; q < n; ++q)
        if ((rc = ibmdem_set_mg_params(b->ctx[q], best_i >= 0 ? cand[best_i] : &orig[3 * q])))
            return batch_error(b, q, rc);
    if ((rc = restore_all())) return rc;
    // then the number of solve groups, timed the same way
    if (b->fused) {
        const int split0 = b->split;
        double best_s = 1e300;
        int best_split = split0;
        for (int sp : {1, 2, 4}) {
            b->split = sp;
            double sec = 0.0;
            rc = run_steps(b, 2, info.data());
            if (rc == IBMDEM_OK) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = run_steps(b, steps, info.data());
                sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / steps;
            }
            if (rc == IBMDEM_OK && sec < best_s) { best_s = sec; best_split = sp; }
            if ((rc = restore_all())) return rc;
        }
        b->split = best_split;
    }
    // then, per case, the predictor schedule (every schedule gives the same bits)
    for (int q = 0; q < n; ++q) {
        ibmdem_ctx *c = b->ctx[q];
        const int tile0 = c->cs.stencil_tile;
        double best_t = 1e300;
        int best_tile = tile0;
        for (int t = 0; t < 4; ++t) {
            if ((rc = ibmdem_set_stencil_tile(c, t))) return batch_error(b, q, rc);
            double sec = 0.0;
            rc = time_steps(c, steps, &sec);
            if (rc == IBMDEM_OK && sec < best_t) { best_t = sec; best_tile = t; }
            if ((rc = restore_state(c, h[q]))) return batch_error(b, q, rc);
        }
        if ((rc = ibmdem_set_stencil_tile(c, best_tile))) return batch_error(b, q, rc);
    }
    b->err.clear();
    return IBMDEM_OK;
}

}  // extern "C"
