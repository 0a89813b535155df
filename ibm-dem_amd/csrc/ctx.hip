// ctx.hip — the C ABI of libibmdem (include/ibmdem.h): context lifetime,
// state transfer, operator entry points and the coupled step
// (FirstOrderForwardEuler, fluid.F90:252-327, with CloudEvolutionMS,
// couplesolverCylinder.F90:189-217).
#include "common.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

int dem_substep_fused(ibmdem_ctx *c, bool first, bool last);

int set_error(ibmdem_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

void prof_begin(ibmdem_ctx *c, int kid) {
    if (c->prof.kernel != kid) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->prof.ev.push_back(e);
}
// the interval since the matching prof_begin held `launches` launches of
// the profiled kernel moving `bytes` together
void prof_end(ibmdem_ctx *c, int kid, double bytes, int launches) {
    if (c->prof.kernel != kid) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->prof.ev.push_back(e);
    c->prof.launches += launches;
    c->prof.bytes_total += bytes;
}

namespace {

double eta_of(const ibmdem_case *cs) {
    const double le = std::log(cs->restitution);
    if (cs->eta_textbook) return -2.0 * le / std::sqrt(3.14159265358979323846 * 3.14159265358979323846 + le * le);
    // 3.1415926536**2: a default-REAL (binary32) literal squared in binary32
    // (ZJS_COLLISION_DEM.F90:30)
    volatile float pf = 3.1415926536f;
    volatile float pf2 = pf * pf;
    return -2 * le / std::sqrt((double)pf2 + le);
}

int validate(const ibmdem_case *cs) {
    if (cs->nx < 2 || cs->ny < 2 || cs->nz < 2 || cs->dx <= 0 || cs->dy <= 0 || cs->dz <= 0 || cs->dt <= 0 ||
        cs->n_markers < 0 || cs->n_bodies < 0 || cs->couple_interval < 1)
        return IBMDEM_ERR_ARG;
    for (int s = 0; s < 6; ++s) {
        const int b = cs->btype[s];
        if (b != IBMDEM_BC_PERIODIC && b != IBMDEM_BC_WALL && b != IBMDEM_BC_INFLOW && b != IBMDEM_BC_OUTFLOW)
            return IBMDEM_ERR_UNSUPPORTED;
        // inflow on low sides only, outflow on high sides only
        if ((b == IBMDEM_BC_INFLOW && (s & 1)) || (b == IBMDEM_BC_OUTFLOW && !(s & 1))) return IBMDEM_ERR_ARG;
        // bulk forcing is a periodic-channel device
        if ((b == IBMDEM_BC_INFLOW || b == IBMDEM_BC_OUTFLOW) &&
            (cs->bulk_switch[0] == 1 || cs->bulk_switch[1] == 1 || cs->bulk_switch[2] == 1))
            return IBMDEM_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3; ++d)
        if ((cs->btype[2 * d] == IBMDEM_BC_PERIODIC) != (cs->btype[2 * d + 1] == IBMDEM_BC_PERIODIC))
            return IBMDEM_ERR_ARG;
    if (cs->precision != IBMDEM_FP64 && cs->precision != IBMDEM_MIXED) return IBMDEM_ERR_ARG;
    if (cs->time_scheme != IBMDEM_SCHEME_EULER && cs->time_scheme != IBMDEM_SCHEME_RK3) return IBMDEM_ERR_ARG;
    if (cs->stencil_tile < 0 || cs->stencil_tile > 3) return IBMDEM_ERR_ARG;
    if ((long)cs->nx * cs->ny * cs->nz >= (1L << 31)) return IBMDEM_ERR_ARG;
    if (cs->n_markers > 0 && cs->n_bodies == 0) return IBMDEM_ERR_ARG;
    // a marker's 5x5x5 neighbourhood must not wrap onto itself, and marker
    // ids are packed in 23 bits in the spread's entry lists
    if (cs->n_markers > 0 && (cs->nx < 5 || cs->ny < 5 || cs->nz < 5)) return IBMDEM_ERR_ARG;
    if (cs->n_markers >= (1 << 23)) return IBMDEM_ERR_UNSUPPORTED;
    return IBMDEM_OK;
}

int alloc_all(ibmdem_ctx *c) {
    const long n = c->n;
    int rc;
#define A(p, cnt) if ((rc = dev_alloc(c, &(p), (cnt)))) return rc
    for (int q = 0; q < 2; ++q) { A(c->u[q], n); A(c->v[q], n); A(c->w[q], n); A(c->P[q], n); }
    A(c->pcorr, n); A(c->rhs, n);
    {
        const long pn[3] = {(long)c->g.ny * c->g.nz, (long)c->g.nx * c->g.nz, (long)c->g.nx * c->g.ny};
        for (int d = 0; d < 3; ++d)
            if (c->g.bt[2 * d + 1] == IBMDEM_BC_OUTFLOW) A(c->g.plane[d], pn[d]);
    }
    A(c->cg_x, n); A(c->cg_r, n); A(c->cg_z, n); A(c->cg_p, n); A(c->cg_Ap, n);
    // bulk-forcing partials: one per 64 x 4 row group and plane; the
    // matrix-vector product's: one per block
    const long pred_blocks = (long)div_up(c->g.nx, 64) * div_up(c->g.ny, 4) * c->g.nz;
    c->max_partial = (int)std::max<long>(std::max<long>(4096, 3 * pred_blocks + 3 * 256 + 16),
                                         poisson_partials(c->g.nx, c->g.ny, c->g.nz));
    A(c->partial, c->max_partial);
    A(c->scal, SC_COUNT);
    A(c->flags, FL_COUNT + 8);
    HIPX(c, hipHostMalloc((void **)&c->scal_host, SC_COUNT * sizeof(double), hipHostMallocDefault));
    HIPX(c, hipHostMalloc((void **)&c->flags_host, (FL_COUNT + 8) * sizeof(int), hipHostMallocDefault));
    HIPX(c, hipEventCreateWithFlags(&c->iter_ev[0], hipEventDisableTiming));
    HIPX(c, hipEventCreateWithFlags(&c->iter_ev[1], hipEventDisableTiming));
    HIPX(c, hipHostMalloc((void **)&c->conv_host, (c->cs.poisson_max_iter + 2) * sizeof(double), hipHostMallocDefault));
    const int M = c->M, B = c->B;
    A(c->mx, 3L * M); A(c->mv, 3L * M); A(c->fdrag, 3L * M); A(c->fcol, 3L * M);
    if (c->cs.time_scheme == IBMDEM_SCHEME_RK3)
        for (int q = 0; q < 3; ++q) { A(c->rhsl[q], n); A(c->uold[q], n); A(c->ffull[q], n); }
    A(c->mvol, M); A(c->mmass, M); A(c->mrad, M); A(c->mbody, M); A(c->mcell, M);
    A(c->mrec, 8L * M);
    A(c->mfac, 24L * M);
    A(c->bin_count, n); A(c->bin_start, n + 1); A(c->bin_items, M); A(c->bin_tmp, M);
    c->scan_nblocks = div_up(n, 4096) + 2;
    A(c->scan_blocks, c->scan_nblocks);
    const long ntiles = ibm_tile_count(c->g);
    A(c->act_flag, ntiles);
    c->act_cap = std::min<long>(ntiles, 8L * M + 8);   // a marker reaches at most 2x2x2 tiles
    c->act_grid = (c->act_cap + 7) / 8 * 8;
    A(c->act_list, c->act_cap);
    A(c->act_meta, 8);
    A(c->tile_ecount, c->act_cap + 1);
    A(c->tile_estart, c->act_cap + 1);
    A(c->tile_mask, IBM_MASK_WORDS * c->act_cap);
    // a cell lies in the source window of at most 3 tiles per dimension
    // (2 directly, 1 more through a wall mirror or periodic wrap)
    c->ent_cap = 27L * M + 64;
    A(c->tile_ent, c->ent_cap);
    A(c->fact, 3L * IBM_TC * c->act_cap);
    A(c->body_start, B + 1);
    A(c->bx, 3L * B); A(c->bv, 3L * B); A(c->bvold, 3L * B); A(c->bom, 3L * B); A(c->bL, 3L * B);
    A(c->bLold, 3L * B); A(c->brot, 9L * B); A(c->brotg, 9L * B); A(c->bI, 9L * B); A(c->bm, B);
    A(c->brho, B); A(c->bfixed, B);
    A(c->bfdrag, 3L * B); A(c->bfmom, 3L * B); A(c->bfcol, 3L * B); A(c->bfmomcol, 3L * B);
    A(c->bfhydro, 3L * B); A(c->bfmomhydro, 3L * B);
    A(c->brmax, B); A(c->bnear, B + 1); A(c->blist, 32L * B);
    const char *ng = std::getenv("IBMDEM_NO_BIN_GATE");
    c->bin_gate = !(ng && *ng && *ng != '0');
#undef A
    return poisson_setup(c);
}

int h2d(ibmdem_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!src || bytes == 0) return IBMDEM_OK;
    HIPX(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return IBMDEM_OK;
}
int d2h(ibmdem_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!dst || bytes == 0) return IBMDEM_OK;
    HIPX(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return IBMDEM_OK;
}
int finish(ibmdem_ctx *c) {
    HIPX(c, hipStreamSynchronize(c->stream));
    return IBMDEM_OK;
}

}  // namespace

HostState host_state(const ibmdem_ctx *c) {
    return HostState{c->vcur, c->pcur, c->bulk_pending, c->first_mdf, c->spread_valid, c->rmax_dirty};
}
void set_host_state(ibmdem_ctx *c, const HostState &h) {
    c->vcur = h.vcur; c->pcur = h.pcur;
    c->bulk_pending = h.bulk_pending; c->first_mdf = h.first_mdf;
    c->spread_valid = h.spread_valid; c->rmax_dirty = h.rmax_dirty;
}

#define CTX_OR_ARG(c) if (!(c)) return IBMDEM_ERR_ARG
#define TRY(expr) do { int rc_ = (expr); if (rc_) return rc_; } while (0)

namespace {

// ---------------------------------------------------------------------------
// The coupled step. Eagerly it is a sequence of launches with two host waits
// inside the Poisson solve (the lagged convergence test) and one at the end
// (the DEM error flags). Once the number of CG iterations a step needs is
// known, the step runs as two captured HIP graphs instead:
//   pre  = IBM begin-step tables, the multi-direct-forcing loop, predictor,
//          divergence, the start of the solve and K CG iterations;
//   post = gauge, projection and the particle sub-steps.
// The host waits once between them to read the convergence flag (and runs
// further iterations eagerly if K did not suffice, raising K for the next
// step) and once after post for the error flags. The captured launches are
// the eager ones, so the results are bitwise the same. The predictor and the
// projection swap ping-pong buffers, so a step's launches depend on the
// (vcur, pcur) parity: each parity has its own pair of graphs.
// ---------------------------------------------------------------------------

// the step up to the solve's right-hand side (r = b, ||b||), in three
// parts: the IBM set-up of the step, the multi-direct-forcing loop, then the
// predictor, bulk forcing, divergence and the solve's right-hand side
int step_pre_a(ibmdem_ctx *c) {
    TRY(fluid_save_uold(c));
    return ibm_begin_step(c);
}

int step_mdf(ibmdem_ctx *c) {
    for (int it = 0; it < c->cs.n_mdf; ++it) {
        TRY(ibm_interpolate(c));
        TRY(ibm_spread(c));
    }
    return IBMDEM_OK;
}

int step_pre_b(ibmdem_ctx *c) {
    TRY(fluid_predictor(c, 1.0, 0.0, 1.0));
    TRY(fluid_bulk(c, false));
    TRY(fluid_divergence(c, 1.0));
    return poisson_begin_rhs(c);
}

int step_pre_rhs(ibmdem_ctx *c) {
    TRY(step_pre_a(c));
    TRY(step_mdf(c));
    return step_pre_b(c);
}

int step_pre(ibmdem_ctx *c) {
    TRY(step_pre_rhs(c));
    return poisson_begin_precondition(c);
}

int step_post(ibmdem_ctx *c) {
    TRY(poisson_gauge(c));
    TRY(fluid_project(c, 1.0));
    TRY(dem_body_hydro(c));
    for (int s = 0; s < c->cs.couple_interval; ++s)
        TRY(dem_substep_fused(c, s == 0, s == c->cs.couple_interval - 1));
    return dem_probe_append(c);
}

void drop_part(CapturedPart &p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    p = CapturedPart();
}

template <class F>
int capture_part(ibmdem_ctx *c, CapturedPart &part, F launches) {
    drop_part(part);
    part.before = host_state(c);
    hipGraph_t graph = nullptr;
    HIPX(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launches();
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    part.after = host_state(c);
    set_host_state(c, part.before);   // run_part applies `after`
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : set_error(c, IBMDEM_ERR_HIP, std::string("step capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&part.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        part.exec = nullptr;
        return set_error(c, IBMDEM_ERR_HIP, std::string("step graph instantiation: ") + hipGetErrorString(ei));
    }
    return IBMDEM_OK;
}

int run_part(ibmdem_ctx *c, CapturedPart &part) {
    HIPX(c, hipGraphLaunch(part.exec, c->stream));
    set_host_state(c, part.after);
    c->graphs.launches++;
    return IBMDEM_OK;
}

// ThirdOrderLowStorageRungeKutta (fluid.F90:329-366): Wray's coefficients,
// c = a + b, three stages of predictor, divergence, solve and projection.
// Each stage advances from the previous stage's velocity (uold <- u before
// every VelocityGuess); the reference routine never sets uold, so as written
// its stages would all start from a stale velocity.
int step_rk3(ibmdem_ctx *c, ibmdem_solve_info *info) {
    const double aR[3] = {32.0 / 60.0, 25.0 / 60.0, 45.0 / 60.0};
    const double bR[3] = {0.0, -17.0 / 60.0, -25.0 / 60.0};
    for (int s = 0; s < 3; ++s) {
        const double cR = aR[s] + bR[s];
        TRY(fluid_save_uold(c));
        TRY(fluid_predictor(c, aR[s], bR[s], cR));
        TRY(fluid_divergence(c, cR));
        TRY(poisson_solve(c, info));
        TRY(fluid_project(c, cR));
    }
    return IBMDEM_OK;
}

int step_eager(ibmdem_ctx *c, ibmdem_solve_info *info) {
    TRY(step_pre(c));
    TRY(poisson_iterate_to_convergence(c, 1));
    TRY(poisson_check(c, info));
    if (c->last_cg_iterations > c->graphs.cg_k) c->graphs.cg_k = c->last_cg_iterations;
    TRY(step_post(c));
    return dem_check_flags(c);
}

}  // namespace

// A step in three phases so that a batch can interleave many contexts (one
// stream each): phase 1 queues the first captured part, phase 2 waits for
// the solve and queues the second part, phase 3 waits for the step's error
// flags. A context that cannot use its graphs yet (unknown iteration count,
// profiling, fresh marker upload) runs the whole step eagerly in phase 1 and
// reports it in *done.
int step_phase1(ibmdem_ctx *c, ibmdem_solve_info *info, bool *done) {
    if (c->cs.time_scheme == IBMDEM_SCHEME_RK3) {
        *done = true;
        return step_rk3(c, info);
    }
    if (!(c->graphs.enabled && c->graphs.cg_k >= 0 && c->prof.kernel == 0 && !c->rmax_dirty)) {
        *done = true;
        return step_eager(c, info);
    }
    *done = false;
    StepGraphs &G = c->graphs;
    const int K = G.cg_k;
    CapturedPart &pre = G.pre[c->vcur * 2 + c->pcur];
    if (!pre.exec || !(pre.before == host_state(c)) || pre.cg_last != K) {
        TRY(capture_part(c, pre, [&] {
            int rc = step_pre(c);
            return rc ? rc : poisson_iterations(c, 1, K);
        }));
        pre.cg_last = K;
    }
    TRY(run_part(c, pre));
    HIPX(c, hipMemcpyAsync(c->conv_host, c->scal + SC_CONV, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return IBMDEM_OK;
}

int step_phase2(ibmdem_ctx *c, ibmdem_solve_info *info) {
    StepGraphs &G = c->graphs;
    HIPX(c, hipStreamSynchronize(c->stream));
    if (c->conv_host[0] == 0.0) TRY(poisson_iterate_to_convergence(c, G.cg_k + 1));
    TRY(poisson_check(c, info));
    if (c->last_cg_iterations > G.cg_k) G.cg_k = c->last_cg_iterations;
    CapturedPart &post = G.post[c->vcur * 2 + c->pcur];
    if (!post.exec || !(post.before == host_state(c))) TRY(capture_part(c, post, [&] { return step_post(c); }));
    return run_part(c, post);
}

int step_phase3(ibmdem_ctx *c) { return dem_check_flags(c); }

// The batched form of phases 1 and 2: the solve's preconditioning and
// iterations are left to the batch (poisson_batch_*), which runs them for a
// group of contexts at once; the launches of each context are those of its
// own step, so the results are bitwise the same.
bool step_batchable(const ibmdem_ctx *c) {
    return c->cs.time_scheme != IBMDEM_SCHEME_RK3 && c->graphs.enabled && c->graphs.cg_k >= 0 &&
           c->prof.kernel == 0 && !c->rmax_dirty && poisson_batchable(c);
}

int step_phase1_batched(ibmdem_ctx *c) {
    CapturedPart &part = c->graphs.pre_rhs[c->vcur * 2 + c->pcur];
    if (!part.exec || !(part.before == host_state(c))) TRY(capture_part(c, part, [&] { return step_pre_rhs(c); }));
    return run_part(c, part);
}

// The same first phase with the forcing loop left to the batch as well:
// part a before it, part b after it.
int step_phase1a_batched(ibmdem_ctx *c) {
    CapturedPart &part = c->graphs.pre_a[c->vcur * 2 + c->pcur];
    if (!part.exec || !(part.before == host_state(c))) TRY(capture_part(c, part, [&] { return step_pre_a(c); }));
    return run_part(c, part);
}

// the host state ibm_spread leaves, for a forcing loop the batch ran
void step_mdf_batched_state(ibmdem_ctx *c) {
    if (c->M > 0 && c->cs.n_mdf > 0) {
        c->first_mdf = false;
        c->spread_valid = true;
    }
}

// the host state step_pre_rhs leaves: ibm_begin_step's, the forcing
// loop's, the predictor's parity swap and fluid_bulk's pending shift
void step_pre_rhs_batched_state(ibmdem_ctx *c) {
    c->first_mdf = true;
    c->spread_valid = false;
    step_mdf_batched_state(c);
    c->vcur = 1 - c->vcur;
    c->bulk_pending = c->cs.bulk_switch[0] == 1 || c->cs.bulk_switch[1] == 1 || c->cs.bulk_switch[2] == 1;
}

int step_phase1b_batched(ibmdem_ctx *c) {
    CapturedPart &part = c->graphs.pre_b[c->vcur * 2 + c->pcur];
    if (!part.exec || !(part.before == host_state(c))) TRY(capture_part(c, part, [&] { return step_pre_b(c); }));
    return run_part(c, part);
}

int step_check_batched(ibmdem_ctx *c, const double *scal, ibmdem_solve_info *info) {
    StepGraphs &G = c->graphs;
    TRY(poisson_check_scalars(c, scal, info));
    if (c->last_cg_iterations > G.cg_k) G.cg_k = c->last_cg_iterations;
    return IBMDEM_OK;
}

int step_post_own(ibmdem_ctx *c) {
    CapturedPart &post = c->graphs.post[c->vcur * 2 + c->pcur];
    if (!post.exec || !(post.before == host_state(c))) TRY(capture_part(c, post, [&] { return step_post(c); }));
    return run_part(c, post);
}

// The host state step_post leaves (fluid_project's pressure swap and the
// consumed bulk shift; the DEM part changes none, rmax_dirty being clear in
// a batchable context), for a post part that a batch ran for the context.
void step_post_batched_state(ibmdem_ctx *c) {
    c->pcur = 1 - c->pcur;
    c->bulk_pending = false;
}

void drop_graphs(ibmdem_ctx *c) {
    c->graphs.epoch++;
    for (int q = 0; q < 4; ++q) {
        drop_part(c->graphs.pre[q]);
        drop_part(c->graphs.post[q]);
        drop_part(c->graphs.pre_rhs[q]);
        drop_part(c->graphs.pre_a[q]);
        drop_part(c->graphs.pre_b[q]);
    }
}

namespace {

int step_impl(ibmdem_ctx *c, ibmdem_solve_info *info) {
    bool done = false;
    TRY(step_phase1(c, info, &done));
    if (done) return IBMDEM_OK;
    TRY(step_phase2(c, info));
    return step_phase3(c);
}

}  // namespace

extern "C" {

int ibmdem_create(const ibmdem_case *cs, ibmdem_ctx **out) {
    if (!cs || !out) return IBMDEM_ERR_ARG;
    *out = nullptr;
    int rc = validate(cs);
    if (rc) return rc;
    ibmdem_ctx *c = new (std::nothrow) ibmdem_ctx();
    if (!c) return IBMDEM_ERR_ALLOC;
    c->cs = *cs;
    if (c->cs.n_mdf <= 0) c->cs.n_mdf = 5;
    if (c->cs.max_pairs <= 0) c->cs.max_pairs = 200;
    if (c->cs.mg_pre <= 0) c->cs.mg_pre = 2;
    if (c->cs.mg_post <= 0) c->cs.mg_post = 2;
    if (c->cs.mg_coarse_sweeps <= 0) c->cs.mg_coarse_sweeps = 24;
    if (c->cs.poisson_max_iter <= 0) c->cs.poisson_max_iter = 200;
    if (c->cs.poisson_rtol <= 0) c->cs.poisson_rtol = 1e-10;
    GridDev &g = c->g;
    g.nx = cs->nx; g.ny = cs->ny; g.nz = cs->nz;
    for (int s = 0; s < 6; ++s) g.bt[s] = cs->btype[s];
    g.dx = cs->dx; g.dy = cs->dy; g.dz = cs->dz;
    g.L[0] = g.dx * g.nx; g.L[1] = g.dy * g.ny; g.L[2] = g.dz * g.nz;
    for (int d = 0; d < 3; ++d) {
        g.uin[d] = cs->btype[2 * d] == IBMDEM_BC_INFLOW ? cs->inflow[d] : 0.0;
        g.plane[d] = nullptr;
    }
    c->n = (long)g.nx * g.ny * g.nz;
    c->dV = g.dx * g.dy * g.dz;
    c->eta = eta_of(&c->cs);
    c->M = cs->n_markers;
    c->B = cs->n_bodies;
    const char *ng = std::getenv("IBMDEM_NO_GRAPH");
    c->graphs.enabled = !(ng && *ng && *ng != '0');
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return IBMDEM_ERR_HIP;
    }
    rc = alloc_all(c);
    // inflow faces hold the inflow velocity from the start
    if (rc == IBMDEM_OK) rc = fluid_upload_bc(c, 7);
    // the bin order is a permutation of the markers from the start (the
    // interpolation walks markers in it)
    std::vector<int> iota(c->M);
    for (int m = 0; m < c->M; ++m) iota[m] = m;
    if (rc == IBMDEM_OK) rc = h2d(c, c->bin_items, iota.data(), (size_t)c->M * sizeof(int));
    if (rc == IBMDEM_OK) rc = finish(c);
    if (rc) {
        ibmdem_destroy(c);
        return rc;
    }
    *out = c;
    return IBMDEM_OK;
}

int ibmdem_destroy(ibmdem_ctx *c) {
    if (!c) return IBMDEM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void *p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
    if (c->scal_host) (void)hipHostFree(c->scal_host);
    if (c->flags_host) (void)hipHostFree(c->flags_host);
    if (c->conv_host) (void)hipHostFree(c->conv_host);
    for (int q = 0; q < 2; ++q)
        if (c->iter_ev[q]) (void)hipEventDestroy(c->iter_ev[q]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return IBMDEM_OK;
}

const char *ibmdem_last_error(const ibmdem_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ibmdem_sync(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    return finish(c);
}

int ibmdem_upload_fields(ibmdem_ctx *c, const double *u, const double *v, const double *w, const double *p) {
    CTX_OR_ARG(c);
    const size_t b = c->n * sizeof(double);
    TRY(h2d(c, c->u[c->vcur], u, b));
    TRY(h2d(c, c->v[c->vcur], v, b));
    TRY(h2d(c, c->w[c->vcur], w, b));
    TRY(h2d(c, c->P[c->pcur], p, b));
    c->bulk_pending = false;
    TRY(fluid_upload_bc(c, (u ? 1 : 0) | (v ? 2 : 0) | (w ? 4 : 0)));
    return finish(c);
}

namespace {
long outflow_plane(ibmdem_ctx *c, int side, double **plane) {
    if (side != 1 && side != 3 && side != 5) return -1;
    const int d = side / 2;
    if (c->g.bt[side] != IBMDEM_BC_OUTFLOW) return -1;
    *plane = c->g.plane[d];
    return d == 0 ? (long)c->g.ny * c->g.nz : (d == 1 ? (long)c->g.nx * c->g.nz : (long)c->g.nx * c->g.ny);
}
}  // namespace

int ibmdem_download_outflow(ibmdem_ctx *c, int32_t side, double *plane) {
    CTX_OR_ARG(c);
    double *dp = nullptr;
    const long n = outflow_plane(c, side, &dp);
    if (n < 0 || !plane) return set_error(c, IBMDEM_ERR_ARG, "not an outflow side (1, 3 or 5 of type OUTFLOW)");
    TRY(d2h(c, plane, dp, n * sizeof(double)));
    return finish(c);
}

int ibmdem_upload_outflow(ibmdem_ctx *c, int32_t side, const double *plane) {
    CTX_OR_ARG(c);
    double *dp = nullptr;
    const long n = outflow_plane(c, side, &dp);
    if (n < 0 || !plane) return set_error(c, IBMDEM_ERR_ARG, "not an outflow side (1, 3 or 5 of type OUTFLOW)");
    TRY(h2d(c, dp, plane, n * sizeof(double)));
    return finish(c);
}

int ibmdem_download_fields(ibmdem_ctx *c, double *u, double *v, double *w, double *p) {
    CTX_OR_ARG(c);
    if (c->bulk_pending) TRY(fluid_bulk(c, true));
    const size_t b = c->n * sizeof(double);
    TRY(d2h(c, u, c->u[c->vcur], b));
    TRY(d2h(c, v, c->v[c->vcur], b));
    TRY(d2h(c, w, c->w[c->vcur], b));
    TRY(d2h(c, p, c->P[c->pcur], b));
    return finish(c);
}

int ibmdem_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    CTX_OR_ARG(c);
    return ibm_download_force(c, fx, fy, fz);
}

int ibmdem_download_poisson(ibmdem_ctx *c, double *pcorr, double *rhs) {
    CTX_OR_ARG(c);
    const size_t b = c->n * sizeof(double);
    TRY(d2h(c, pcorr, c->pcorr, b));
    TRY(d2h(c, rhs, c->rhs, b));
    return finish(c);
}

int ibmdem_upload_poisson_rhs(ibmdem_ctx *c, const double *rhs) {
    CTX_OR_ARG(c);
    if (!rhs) return IBMDEM_ERR_ARG;
    TRY(h2d(c, c->rhs, rhs, c->n * sizeof(double)));
    return finish(c);
}

int ibmdem_upload_markers(ibmdem_ctx *c, const ibmdem_markers *m) {
    CTX_OR_ARG(c);
    if (!m || !m->x || !m->v || !m->vol || !m->mass || !m->radeff || !m->body) return IBMDEM_ERR_ARG;
    const int M = c->M, B = c->B;
    std::vector<int> start(B + 1, 0);
    for (int i = 0; i < M; ++i) {
        const int b = m->body[i];
        if (b < 0 || b >= B || (i > 0 && b < m->body[i - 1]))
            return set_error(c, IBMDEM_ERR_ARG, "markers must be body-contiguous, bodies ascending");
        start[b + 1]++;
    }
    for (int b = 0; b < B; ++b) start[b + 1] += start[b];
    TRY(h2d(c, c->body_start, start.data(), (B + 1) * sizeof(int)));
    const size_t v3 = 3L * M * sizeof(double), v1 = (size_t)M * sizeof(double);
    TRY(h2d(c, c->mx, m->x, v3));
    TRY(h2d(c, c->mv, m->v, v3));
    if (m->fdrag) TRY(h2d(c, c->fdrag, m->fdrag, v3));
    else HIPX(c, hipMemsetAsync(c->fdrag, 0, v3, c->stream));
    if (m->fcol) TRY(h2d(c, c->fcol, m->fcol, v3));
    else HIPX(c, hipMemsetAsync(c->fcol, 0, v3, c->stream));
    TRY(h2d(c, c->mvol, m->vol, v1));
    TRY(h2d(c, c->mmass, m->mass, v1));
    TRY(h2d(c, c->mrad, m->radeff, v1));
    TRY(h2d(c, c->mbody, m->body, (size_t)M * sizeof(int)));
    // the contact scan's body radii, measured at its next run (a context
    // without markers or bodies never runs it)
    c->rmax_dirty = c->M > 0 && c->B > 0;
    TRY(dem_rebuild(c));
    TRY(dem_check_flags(c));
    return finish(c);
}

int ibmdem_download_markers(ibmdem_ctx *c, ibmdem_markers *m) {
    CTX_OR_ARG(c);
    if (!m) return IBMDEM_ERR_ARG;
    const int M = c->M;
    const size_t v3 = 3L * M * sizeof(double), v1 = (size_t)M * sizeof(double);
    TRY(d2h(c, m->x, c->mx, v3));
    TRY(d2h(c, m->v, c->mv, v3));
    TRY(d2h(c, m->fdrag, c->fdrag, v3));
    TRY(d2h(c, m->fcol, c->fcol, v3));
    TRY(d2h(c, m->vol, c->mvol, v1));
    TRY(d2h(c, m->mass, c->mmass, v1));
    TRY(d2h(c, m->radeff, c->mrad, v1));
    TRY(d2h(c, m->body, c->mbody, (size_t)M * sizeof(int)));
    return finish(c);
}

int ibmdem_upload_bodies(ibmdem_ctx *c, const ibmdem_bodies *b) {
    CTX_OR_ARG(c);
    if (!b) return IBMDEM_ERR_ARG;
    const int B = c->B;
    const size_t v3 = 3L * B * sizeof(double), v9 = 9L * B * sizeof(double), v1 = (size_t)B * sizeof(double);
    TRY(h2d(c, c->bx, b->x, v3)); TRY(h2d(c, c->bv, b->v, v3)); TRY(h2d(c, c->bvold, b->vold, v3));
    TRY(h2d(c, c->bom, b->omega, v3)); TRY(h2d(c, c->bL, b->angmom, v3));
    TRY(h2d(c, c->bLold, b->angmom_old, v3));
    TRY(h2d(c, c->brot, b->rot, v9)); TRY(h2d(c, c->brotg, b->rot_glo, v9)); TRY(h2d(c, c->bI, b->inertia, v9));
    TRY(h2d(c, c->bm, b->mass, v1)); TRY(h2d(c, c->brho, b->rho, v1));
    TRY(h2d(c, c->bfixed, b->fixed, (size_t)B * sizeof(int)));
    TRY(h2d(c, c->bfhydro, b->fhydro, v3)); TRY(h2d(c, c->bfmomhydro, b->fmomhydro, v3));
    c->rmax_dirty = c->M > 0 && c->B > 0;
    return finish(c);
}

int ibmdem_download_bodies(ibmdem_ctx *c, ibmdem_bodies *b) {
    CTX_OR_ARG(c);
    if (!b) return IBMDEM_ERR_ARG;
    const int B = c->B;
    const size_t v3 = 3L * B * sizeof(double), v9 = 9L * B * sizeof(double), v1 = (size_t)B * sizeof(double);
    TRY(d2h(c, b->x, c->bx, v3)); TRY(d2h(c, b->v, c->bv, v3)); TRY(d2h(c, b->vold, c->bvold, v3));
    TRY(d2h(c, b->omega, c->bom, v3)); TRY(d2h(c, b->angmom, c->bL, v3));
    TRY(d2h(c, b->angmom_old, c->bLold, v3));
    TRY(d2h(c, b->rot, c->brot, v9)); TRY(d2h(c, b->rot_glo, c->brotg, v9)); TRY(d2h(c, b->inertia, c->bI, v9));
    TRY(d2h(c, b->mass, c->bm, v1)); TRY(d2h(c, b->rho, c->brho, v1));
    TRY(d2h(c, b->fixed, c->bfixed, (size_t)B * sizeof(int)));
    TRY(d2h(c, b->fdrag, c->bfdrag, v3)); TRY(d2h(c, b->fmom, c->bfmom, v3)); TRY(d2h(c, b->fcol, c->bfcol, v3));
    TRY(d2h(c, b->fmomcol, c->bfmomcol, v3)); TRY(d2h(c, b->fhydro, c->bfhydro, v3));
    TRY(d2h(c, b->fmomhydro, c->bfmomhydro, v3));
    return finish(c);
}

int ibmdem_begin_step(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    TRY(fluid_save_uold(c));
    return ibm_begin_step(c);
}
int ibmdem_ibm_interpolate(ibmdem_ctx *c) { CTX_OR_ARG(c); return ibm_interpolate(c); }
int ibmdem_ibm_spread(ibmdem_ctx *c) { CTX_OR_ARG(c); return ibm_spread(c); }
int ibmdem_fluid_predictor(ibmdem_ctx *c, double a, double b, double cc) {
    CTX_OR_ARG(c);
    return fluid_predictor(c, a, b, cc);
}
int ibmdem_bulk_forcing(ibmdem_ctx *c) { CTX_OR_ARG(c); return fluid_bulk(c, true); }
int ibmdem_divergence(ibmdem_ctx *c, double cc) { CTX_OR_ARG(c); return fluid_divergence(c, cc); }
int ibmdem_poisson_solve(ibmdem_ctx *c, ibmdem_solve_info *info) { CTX_OR_ARG(c); return poisson_solve(c, info); }
int ibmdem_project(ibmdem_ctx *c, double cc) { CTX_OR_ARG(c); return fluid_project(c, cc); }
int ibmdem_apply_bc(ibmdem_ctx *c) { CTX_OR_ARG(c); return IBMDEM_OK; }

int ibmdem_body_hydro(ibmdem_ctx *c) { CTX_OR_ARG(c); return dem_body_hydro(c); }
int ibmdem_dem_contact(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    TRY(dem_contact(c));
    return dem_check_flags(c);
}
int ibmdem_body_assemble(ibmdem_ctx *c) { CTX_OR_ARG(c); return dem_body_assemble(c, true); }
int ibmdem_body_integrate(ibmdem_ctx *c) { CTX_OR_ARG(c); return dem_body_integrate(c); }
int ibmdem_markers_follow(ibmdem_ctx *c) { CTX_OR_ARG(c); return dem_markers_follow(c); }
int ibmdem_dem_rebuild_lists(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    TRY(dem_rebuild(c));
    return dem_check_flags(c);
}
int ibmdem_dem_substep(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    TRY(dem_substep_fused(c, true, true));
    return dem_check_flags(c);
}
int ibmdem_cloud_evolution(ibmdem_ctx *c) {
    CTX_OR_ARG(c);
    TRY(dem_body_hydro(c));
    for (int s = 0; s < c->cs.couple_interval; ++s)
        TRY(dem_substep_fused(c, s == 0, s == c->cs.couple_interval - 1));
    return dem_check_flags(c);
}

int ibmdem_step(ibmdem_ctx *c, ibmdem_solve_info *info) {
    CTX_OR_ARG(c);
    return step_impl(c, info);
}

int ibmdem_run(ibmdem_ctx *c, int32_t n_steps, ibmdem_solve_info *info) {
    CTX_OR_ARG(c);
    for (int s = 0; s < n_steps; ++s) TRY(step_impl(c, info));
    return IBMDEM_OK;
}

int ibmdem_get_cells(ibmdem_ctx *c, int32_t *ijk) {
    CTX_OR_ARG(c);
    if (!ijk) return IBMDEM_ERR_ARG;
    std::vector<int> cl(c->M);
    TRY(d2h(c, cl.data(), c->mcell, (size_t)c->M * sizeof(int)));
    TRY(finish(c));
    const GridDev &g = c->g;
    for (int m = 0; m < c->M; ++m) {
        ijk[m] = cl[m] % g.nx + 1;
        ijk[c->M + m] = (cl[m] / g.nx) % g.ny + 1;
        ijk[2L * c->M + m] = cl[m] / (g.nx * g.ny) + 1;
    }
    return IBMDEM_OK;
}

int ibmdem_get_pairs(ibmdem_ctx *c, int32_t cap, int32_t *count, int32_t *partner, int32_t *shift) {
    CTX_OR_ARG(c);
    if (!count || cap <= 0) return IBMDEM_ERR_ARG;
    TRY(dem_pairs(c, cap));
    TRY(d2h(c, count, c->pair_count, (size_t)c->M * sizeof(int)));
    if (partner) TRY(d2h(c, partner, c->pair_partner, (size_t)c->M * cap * sizeof(int)));
    if (shift) TRY(d2h(c, shift, c->pair_shift, (size_t)c->M * cap * sizeof(int)));
    TRY(finish(c));
    // the pair kernel also flags counts above maxCollisionNo; a listing only
    // reports overflow of the caller's capacity
    HIPX(c, hipMemsetAsync(c->flags, 0, FL_COUNT * sizeof(int), c->stream));
    for (int m = 0; m < c->M; ++m)
        if (count[m] > cap) return set_error(c, IBMDEM_ERR_CAPACITY, "pair list longer than the caller's capacity");
    return finish(c);
}

int ibmdem_profile_kernel(ibmdem_ctx *c, int32_t kid) {
    CTX_OR_ARG(c);
    TRY(finish(c));
    for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
    c->prof = Profile();
    c->prof.kernel = kid;
    return IBMDEM_OK;
}

int ibmdem_profile_read(ibmdem_ctx *c, double *total_ms, int64_t *launches, double *bytes_per_launch) {
    CTX_OR_ARG(c);
    TRY(finish(c));
    double tot = 0.0;
    for (size_t q = 0; q + 1 < c->prof.ev.size(); q += 2) {
        float ms = 0.f;
        HIPX(c, hipEventElapsedTime(&ms, c->prof.ev[q], c->prof.ev[q + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = c->prof.launches;
    if (bytes_per_launch) *bytes_per_launch = c->prof.launches ? c->prof.bytes_total / c->prof.launches : 0.0;
    return IBMDEM_OK;
}

int ibmdem_memory_bytes(const ibmdem_ctx *c, int64_t *bytes) {
    if (!c || !bytes) return IBMDEM_ERR_ARG;
    *bytes = c->bytes;
    return IBMDEM_OK;
}

}  // extern "C"
