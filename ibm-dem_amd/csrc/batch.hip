// batch.hip — the batched form: many independent cases of uneven grid size
// and particle count advanced together on one GPU.
//
// Each case is a full context with its own HIP stream. A batch step runs the
// three phases of every case's step interleaved (queue the first captured
// graph of all cases, then wait for each case's solve and queue its second
// graph, then collect each case's error flags), so the cases' kernels overlap
// on the GPU and the host never waits for one case while another could be
// queued. The autotuner picks, per case, the multigrid cycle (pre/post
// smoothing sweeps, coarsest-level sweeps) with the shortest measured step,
// then the predictor's tile schedule (case.stencil_tile) under that cycle.
#include "common.hpp"

#include <chrono>
#include <new>
#include <vector>

struct ibmdem_batch {
    std::vector<ibmdem_ctx *> ctx;
    std::string err;
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

// A case that fails stops there (its context keeps its own message, as a
// single context's failed step does); every other case still completes the
// step in progress, so no context is left between its two step graphs. The
// run then stops and returns the first failure (lowest case index).
int ibmdem_batch_run(ibmdem_batch *b, int32_t n_steps, ibmdem_solve_info *info) {
    if (!b || n_steps < 0) return IBMDEM_ERR_ARG;
    const int n = (int)b->ctx.size();
    std::vector<char> done(n), failed(n, 0);
    std::vector<ibmdem_solve_info> tmp(n);
    ibmdem_solve_info *inf = info ? info : tmp.data();
    int first_rc = IBMDEM_OK, first_q = -1;
    auto fail = [&](int q, int rc) {
        failed[q] = 1;
        if (first_q < 0 || q < first_q) { first_q = q; first_rc = rc; }
    };
    for (int s = 0; s < n_steps && first_q < 0; ++s) {
        for (int q = 0; q < n; ++q) {
            bool d = false;
            const int rc = step_phase1(b->ctx[q], &inf[q], &d);
            done[q] = d;
            if (rc) fail(q, rc);
        }
        for (int q = 0; q < n; ++q) {
            if (done[q] || failed[q]) continue;
            const int rc = step_phase2(b->ctx[q], &inf[q]);
            if (rc) fail(q, rc);
        }
        for (int q = 0; q < n; ++q) {
            if (done[q] || failed[q]) continue;
            const int rc = step_phase3(b->ctx[q]);
            if (rc) fail(q, rc);
        }
    }
    return first_q < 0 ? IBMDEM_OK : batch_error(b, first_q, first_rc);
}

int ibmdem_batch_autotune(ibmdem_batch *b, int32_t trial_steps) {
    if (!b) return IBMDEM_ERR_ARG;
    const int steps = trial_steps > 0 ? trial_steps : 3;
    // candidate cycles: pre, post, coarsest-level sweeps
    static const int32_t cand[][3] = {{1, 1, 24}, {2, 2, 24}, {3, 3, 24}, {2, 2, 12}, {1, 2, 24}};
    for (int q = 0; q < (int)b->ctx.size(); ++q) {
        ibmdem_ctx *c = b->ctx[q];
        HostCopy h;
        int rc = save_state(c, h);
        if (rc) return batch_error(b, q, rc);
        int32_t orig[3];
        ibmdem_get_mg_params(c, orig);
        double best = 1e300;
        int best_i = -1;
        for (int i = 0; i < (int)(sizeof cand / sizeof cand[0]); ++i) {
            if ((rc = ibmdem_set_mg_params(c, cand[i]))) return batch_error(b, q, rc);
            double sec = 0.0;
            rc = time_steps(c, steps, &sec);
            // a cycle that fails (e.g. no convergence within the cap) is not a candidate
            if (rc == IBMDEM_OK && sec < best) { best = sec; best_i = i; }
            if ((rc = restore_state(c, h))) return batch_error(b, q, rc);
        }
        if ((rc = ibmdem_set_mg_params(c, best_i >= 0 ? cand[best_i] : orig))) return batch_error(b, q, rc);
        if ((rc = restore_state(c, h))) return batch_error(b, q, rc);
        // then the predictor schedule (every schedule gives the same bits)
        const int tile0 = c->cs.stencil_tile;
        double best_t = 1e300;
        int best_tile = tile0;
        for (int t = 0; t < 4; ++t) {
            if ((rc = ibmdem_set_stencil_tile(c, t))) return batch_error(b, q, rc);
            double sec = 0.0;
            rc = time_steps(c, steps, &sec);
            if (rc == IBMDEM_OK && sec < best_t) { best_t = sec; best_tile = t; }
            if ((rc = restore_state(c, h))) return batch_error(b, q, rc);
        }
        if ((rc = ibmdem_set_stencil_tile(c, best_tile))) return batch_error(b, q, rc);
    }
    return IBMDEM_OK;
}

}  // extern "C"
