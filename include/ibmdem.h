/*
 * ibmdem.h — C ABI of the MI355X coupled IBM-DEM time step.
 *
 * One context owns one case on one GPU: the staggered (MAC) velocity field,
 * the pressure, the Lagrangian markers ("parcels") and the rigid bodies
 * ("multi-spheres"). Every entry point returns an int status (IBMDEM_OK on
 * success) and never throws, aborts or STOPs across the boundary. Host buffers
 * passed in are borrowed for the duration of the call only. Calls are ordered
 * on the context's HIP stream; calls that return host data synchronize.
 *
 * Index conventions follow the reference (vars.F90:130-153):
 *   - cell (i,j,k), 1-based, i in 1..nx fastest; fields cross the boundary
 *     as interior-only arrays of nx*ny*nz doubles, linear index
 *     (k-1)*nx*ny + (j-1)*nx + (i-1);
 *   - u(i,j,k) lives on the x-face at x=(i-1)dx, v and w likewise on their
 *     faces, P at the cell centre x=(i-1/2)dx (initialization.F90:251,
 *     fluid.F90:80,97, ZJS_INTERPOLATE_DEM.F90:12-14);
 *   - boundary sides 0..5 = reference bType(1..6) = x-, x+, y-, y+, z-, z+
 *     (initialization.F90:341-377).
 * Marker and body vectors are component-major: x[0..n) then y[0..n) then
 * z[0..n). 3x3 matrices are row-major per body, as the reference's
 * mat(1:9) (couplesolverCylinder.F90:87-95): 9 consecutive doubles per body.
 */
#ifndef IBMDEM_H
#define IBMDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define IBMDEM_OK                 0
#define IBMDEM_ERR_HIP            1  /* a HIP runtime call failed            */
#define IBMDEM_ERR_ARG            2  /* invalid argument or case             */
#define IBMDEM_ERR_CAPACITY       3  /* pair list or bin capacity exceeded   */
#define IBMDEM_ERR_OUTSIDE        4  /* a marker left the domain             */
#define IBMDEM_ERR_NOT_CONVERGED  5  /* Poisson solve hit max iterations     */
#define IBMDEM_ERR_NAN            6  /* non-finite value detected            */
#define IBMDEM_ERR_ALLOC          7  /* device allocation failed             */
#define IBMDEM_ERR_UNSUPPORTED    8  /* feature not available in this form   */

/* ---- boundary types (reference bType: 0 periodic, 1 solid wall) --------
 * INFLOW and OUTFLOW have no counterpart in the reference (its fluid
 * boundaries are periodic or no-slip, initialization.F90:178-237):
 *   INFLOW  (low sides only): the velocity is cs.inflow at the face, the
 *           normal component stored on the face, the tangential ones
 *           through the ghost 2*U_t - u; pressure Neumann;
 *   OUTFLOW (high sides only): zero normal gradient of every velocity
 *           component, pressure (and its correction) zero at the face; the
 *           Poisson problem is then non-singular (no pinned cell).
 * Markers treat both like walls: the spread adds their mirror images across
 * the side (as the reference does for its outlet ghosts, ZJS_MPI_DEM.F90:
 * 203-209) and markers in the last cell layer get wall contact from it. */
#define IBMDEM_BC_PERIODIC 0
#define IBMDEM_BC_WALL     1
#define IBMDEM_BC_INFLOW   2
#define IBMDEM_BC_OUTFLOW  3

/* ---- precision forms ---------------------------------------------------- */
#define IBMDEM_FP64  0   /* everything in FP64                                   */
#define IBMDEM_MIXED 1   /* FP32 stencil predictor and FP32 multigrid inside FP64
                            CG (residual, inner products, projection, DEM in FP64) */

/* ---- time schemes --------------------------------------------------------
 * EULER: FirstOrderForwardEuler (fluid.F90:252-327) with the IBM/DEM
 *   coupling; the predictor takes the forward-Euler coefficients (1, 0, 1).
 * RK3: the context also keeps uold, RHSl and the full-field IBM force, so
 *   VelocityGuess(a, b, c) takes any coefficients, and ibmdem_step runs
 *   ThirdOrderLowStorageRungeKutta (fluid.F90:329-366): three stages of
 *   predictor, divergence, Poisson solve and projection, fluid only, as the
 *   reference; each stage starts from the previous stage's velocity (Wray's
 *   scheme; the reference routine never sets uold). */
#define IBMDEM_SCHEME_EULER 0
#define IBMDEM_SCHEME_RK3   1

/* ---- case description --------------------------------------------------- */
typedef struct ibmdem_case {
    int32_t nx, ny, nz;          /* interior cells                          */
    int32_t btype[6];            /* IBMDEM_BC_* per side                    */
    double  dx, dy, dz;          /* grid spacing                            */
    double  visc, rhof, dt;      /* nu, rho_f, fluid time step              */
    double  cp[3];               /* constant pressure gradient CPx,CPy,CPz  */
    int32_t bulk_switch[3];      /* DefAveBulk{U,V,W}Switch                 */
    double  bulk_target[3];      /* bulkU, bulkV, bulkW                     */
    double  gravity[3];          /* body gravity (reference: (0,0,-9.81))   */
    int32_t n_mdf;               /* multi-direct-forcing iterations Ns (5)  */
    int32_t couple_interval;     /* DEM sub-steps per fluid step (10)       */
    double  kn_coeff;            /* knCoeff (4e4)                           */
    double  restitution;         /* ResistCoeff e (1e-4)                    */
    double  collision_dist;      /* collisionDist = (0.75 dx)*2             */
    double  collision_buffer;    /* collisionBuffer = 0.1                   */
    int32_t eta_textbook;        /* 0: reference eta, 1: -2ln e/sqrt(pi^2+ln^2 e) */
    int32_t n_markers;           /* M                                       */
    int32_t n_bodies;            /* B                                       */
    int32_t max_pairs;           /* maxCollisionNo (200)                    */
    int32_t precision;           /* IBMDEM_FP64 / IBMDEM_MIXED              */
    double  poisson_rtol;        /* ||b-Ax||_2 <= rtol ||b||_2              */
    int32_t poisson_max_iter;    /* CG iteration cap                        */
    int32_t mg_pre, mg_post;     /* red-black GS sweeps (0 -> 2,2)          */
    int32_t mg_coarse_sweeps;    /* sweeps on the coarsest level (0 -> 24)  */
    int32_t mg_max_levels;       /* 0 -> as many as the grid allows         */
    int32_t stencil_tile;        /* predictor schedule, same result: 0 LDS  *
                                  * z-marching 64x4 x 32 planes (default),  *
                                  * 1 64x4 x 16, 2 64x8 x 32, 3 direct      */
    int32_t time_scheme;         /* IBMDEM_SCHEME_*                         */
    double  inflow[3];           /* normal velocity of the inflow side of
                                  * dimension d (x, y, z)                  */
    int32_t reserved[6];
} ibmdem_case;

/* Markers (reference type Parcel, ZJS_VARS.F90:6-12), SoA. Pointers may be
 * NULL in a download to skip that array; every array is required in an
 * upload except fdrag/fcol (NULL -> zero). */
typedef struct ibmdem_markers {
    double  *x;       /* 3*M positions                                    */
    double  *v;       /* 3*M velocities                                   */
    double  *fdrag;   /* 3*M IBM force density accumulator (fdrag)        */
    double  *fcol;    /* 3*M contact force                                */
    double  *vol;     /* M   marker volume                                */
    double  *mass;    /* M   marker mass m                                */
    double  *radeff;  /* M   contact radius radeff                        */
    int32_t *body;    /* M   0-based owning body (MSIdx-1); markers of one
                               body must be contiguous and bodies ascending */
} ibmdem_markers;

/* Bodies (reference type MultiSphere, ZJS_VARS.F90:97-102), SoA. */
typedef struct ibmdem_bodies {
    double  *x, *v, *vold;           /* 3*B                                */
    double  *omega, *angmom, *angmom_old; /* 3*B                           */
    double  *rot;                    /* 9*B incremental rotation rotMat    */
    double  *rot_glo;                /* 9*B accumulated rotation rotMatGlo */
    double  *inertia;                /* 9*B body-frame inertia tensor      */
    double  *mass, *rho;             /* B                                  */
    int32_t *fixed;                  /* B   1 -> body does not move        */
    /* diagnostics (download only; may be NULL) */
    double  *fdrag, *fmom, *fcol, *fmomcol, *fhydro, *fmomhydro; /* 3*B    */
} ibmdem_bodies;

typedef struct ibmdem_solve_info {
    int32_t iterations;     /* CG iterations performed                    */
    int32_t converged;      /* 1 if ||r|| <= rtol ||b||                   */
    double  rel_residual;   /* final ||r||_2 / ||b||_2 (recurrence)        */
    double  rhs_norm;       /* ||b||_2                                    */
} ibmdem_solve_info;

typedef struct ibmdem_ctx ibmdem_ctx;

/* ---- lifetime and state (replace allocate.F90, AllocateDEM, the
 *      Fluid/InitializationDEM state set-up and the output copies) ------- */
int ibmdem_create(const ibmdem_case *cs, ibmdem_ctx **out);
int ibmdem_destroy(ibmdem_ctx *ctx);
const char *ibmdem_last_error(const ibmdem_ctx *ctx);
int ibmdem_sync(ibmdem_ctx *ctx);
/* u, v, w, p: nx*ny*nz each; NULL skips. Upload applies the boundary
 * rules (wall-normal face velocities forced to zero, fluid.F90:243-248;
 * inflow faces set to the inflow velocity; the outflow plane of an uploaded
 * normal component reset to its last interior face). */
int ibmdem_upload_fields(ibmdem_ctx *ctx, const double *u, const double *v,
                         const double *w, const double *p);
int ibmdem_download_fields(ibmdem_ctx *ctx, double *u, double *v, double *w,
                           double *p);
/* Outflow plane of side 1, 3 or 5 (type IBMDEM_BC_OUTFLOW): the normal
 * velocity on the boundary face, which the interior-only field arrays do not
 * hold; ny*nz (side 1), nx*nz (side 3) or nx*ny (side 5) values, the lower
 * transverse index fastest. No counterpart in the reference (its fluid has
 * no outflow side); part of the state a snapshot must carry. */
int ibmdem_download_outflow(ibmdem_ctx *ctx, int32_t side, double *plane);
int ibmdem_upload_outflow(ibmdem_ctx *ctx, int32_t side, const double *plane);
/* IBM momentum source f = (fx,fy,fz) on the interior (fluid.F90 fx,fy,fz). */
int ibmdem_download_force(ibmdem_ctx *ctx, double *fx, double *fy, double *fz);
/* The same download without any device allocation, for callers that cannot
 * allocate (code that forbids hipMalloc, tight serving loops). No counterpart
 * in the reference. It stages the fields one slab of z planes at a time in a
 * fixed 16 MiB device array that each GPU holds from the first call until the
 * process ends; a grid whose single z plane holds more than 699,050 cells is
 * rejected with IBMDEM_ERR_UNSUPPORTED. Calls on one GPU share that array:
 * calls that overlap on one GPU race. */
int ibmdem_download_force_noalloc(ibmdem_ctx *ctx, double *fx, double *fy, double *fz);
/* Pressure correction P' of the last solve and the Poisson RHS velDiv. */
int ibmdem_download_poisson(ibmdem_ctx *ctx, double *pcorr, double *rhs);
int ibmdem_upload_poisson_rhs(ibmdem_ctx *ctx, const double *rhs);
int ibmdem_upload_markers(ibmdem_ctx *ctx, const ibmdem_markers *m);
int ibmdem_download_markers(ibmdem_ctx *ctx, ibmdem_markers *m);
int ibmdem_upload_bodies(ibmdem_ctx *ctx, const ibmdem_bodies *b);
int ibmdem_download_bodies(ibmdem_ctx *ctx, ibmdem_bodies *b);

/* ---- the coupled step, operator by operator (fluid.F90:252-327) -------- */
/* Start of FirstOrderForwardEuler (fluid.F90:258-263): fdrag = 0, f = 0,
 * and the per-step IBM tables (delta weights, spread footprint). */
int ibmdem_begin_step(ibmdem_ctx *ctx);
/* UpdateFluidVelocityOnLagrangianPoints (ZJS_VOLUMEFRACTION_DEM.F90:2-36)
 * with the weight lists of ZJS_UPDATE_LIST_DEM.F90:323-416 folded in. */
int ibmdem_ibm_interpolate(ibmdem_ctx *ctx);
/* UpdateParcelInfoInGhostRegion mirror images (ZJS_MPI_DEM.F90:222-269) +
 * UpdateMomentumSourceTerm (ZJS_VOLUMEFRACTION_DEM.F90:38-67) +
 * VelocityPostGuess (fluid.F90:135-153) + FillGhost/Boundary. */
int ibmdem_ibm_spread(ibmdem_ctx *ctx);
/* VelocityGuess(a,b,c) (fluid.F90:107-133), fused ConvectionTensor,
 * ViscousTensor and PressureGradient (fluid.F90:2-87). */
int ibmdem_fluid_predictor(ibmdem_ctx *ctx, double a, double b, double c);
/* ModifyBulkVelocity (fluid.F90:155-197). */
int ibmdem_bulk_forcing(ibmdem_ctx *ctx);
/* VelocityDivergence(c) (fluid.F90:89-105). */
int ibmdem_divergence(ibmdem_ctx *ctx, double c);
/* SetPressureVector + PoissonSolver + GetLocalArray (poisson.F90:107-186):
 * multigrid-preconditioned CG replacing PETSc KSPSolve. info may be NULL. */
int ibmdem_poisson_solve(ibmdem_ctx *ctx, ibmdem_solve_info *info);
/* PressureCorrection + VelocityCorrection(c) (fluid.F90:199-238). */
int ibmdem_project(ibmdem_ctx *ctx, double c);
/* FillGhostRegionForFluid + Boundary (boundary.F90:122-247,
 * fluid.F90:240-250): boundary rules are applied inside every kernel, so
 * this entry point only validates the context. */
int ibmdem_apply_bc(ibmdem_ctx *ctx);

/* ---- DEM and rigid bodies (couplesolverCylinder.F90) -------------------- */
/* CloudEvolutionMS:194-211: fhydro/fmomHydro from (v-vold),(L-Lold),
 * AssembleForceMS, then fhydro += fdrag/m, fmomHydro += fmom. */
int ibmdem_body_hydro(ibmdem_ctx *ctx);
/* CalculateCollisionForce + CalculateSolidBoundaryCollisionForce
 * (ZJS_COLLISION_DEM.F90:21-148) over the pair list of
 * RebuildParcelCollisionListDEM (ZJS_UPDATE_LIST_DEM.F90:248-286). */
int ibmdem_dem_contact(ibmdem_ctx *ctx);
/* AssembleLocalForceMS (ZJS_MS_DEM.F90:494-519). */
int ibmdem_body_assemble(ibmdem_ctx *ctx);
/* EvolutionMS (couplesolverCylinder.F90:2-42). */
int ibmdem_body_integrate(ibmdem_ctx *ctx);
/* ResetVelocityAndPositionLagrangianPoints (couplesolverCylinder.F90:44-74)
 * + AdjustParcelPosition (ZJS_EVOLUTION_DEM.F90:121-150). */
int ibmdem_markers_follow(ibmdem_ctx *ctx);
/* RebuildAllListOrNot (ZJS_EVOLUTION_DEM.F90:34-41): periodic migration
 * (CompressDEM), cell binning (PutParcelsIntoCellsDEM). */
int ibmdem_dem_rebuild_lists(ibmdem_ctx *ctx);
/* One CloudEvolutionStepMS (couplesolverCylinder.F90:99-187). */
int ibmdem_dem_substep(ibmdem_ctx *ctx);
/* CloudEvolutionMS (couplesolverCylinder.F90:189-217). */
int ibmdem_cloud_evolution(ibmdem_ctx *ctx);

/* ---- whole step ---------------------------------------------------------- */
/* FirstOrderForwardEuler (fluid.F90:252-327) with CloudEvolutionMS. */
int ibmdem_step(ibmdem_ctx *ctx, ibmdem_solve_info *info);
/* Run n steps; info (may be NULL) receives the last solve. */
int ibmdem_run(ibmdem_ctx *ctx, int32_t n_steps, ibmdem_solve_info *info);

/* ---- per-body probes (PostProcessing, fluid.F90:368-404) -----------------
 * After every coupled step the library appends, for each listed body, 18
 * doubles -- position, velocity, angular velocity, assembled hydrodynamic
 * force (rho_f sum fdrag vol) and torque, contact force, 3 each in that
 * order -- to a device ring holding the last `capacity` steps. Recording
 * takes no host wait (the step graphs include it); n_bodies = 0 stops it.
 * The reference writes these probes to files every step; here the caller
 * reads them when it wants. */
int ibmdem_probe_bodies(ibmdem_ctx *ctx, int32_t n_bodies, const int32_t *bodies, int32_t capacity);
/* *n_steps = steps held (at most capacity), *first_step = the 0-based index
 * of the oldest held step since probing started; records (may be NULL)
 * receives n_steps * n_bodies * 18 doubles, oldest step first. */
int ibmdem_probe_read(ibmdem_ctx *ctx, int32_t *n_steps, int64_t *first_step, double *records);

/* ---- integer lists for parity / diagnostics ------------------------------ */
/* Per-marker cell as the reference's PutParcelsIntoCellsDEM computes it
 * (ZJS_UPDATE_LIST_DEM.F90:41-43): 1-based (i,j,k), 3*M int32,
 * component-major. */
int ibmdem_get_cells(ibmdem_ctx *ctx, int32_t *cell_ijk);
/* Canonical contact candidate list: for marker m (id order), count[m]
 * partners; partner[m*cap + q] = partner id, shift[m*cap + q] = periodic
 * image code (sx+1) + 3(sy+1) + 9(sz+1), s in {-1,0,1}; order = the
 * reference scan order (125 cells kk,jj,ii; in-bin ascending id). */
int ibmdem_get_pairs(ibmdem_ctx *ctx, int32_t cap, int32_t *count,
                     int32_t *partner, int32_t *shift);

/* ---- multigrid cycle of a context --------------------------------------------
 * mg[0..2] = pre-smoothing sweeps, post-smoothing sweeps, coarsest-level
 * sweeps of the V-cycle preconditioner (no counterpart in the reference,
 * whose solver is PETSc's, poisson.F90:133-158). Setting them drops the
 * context's captured step graphs. */
int ibmdem_get_mg_params(ibmdem_ctx *ctx, int32_t *mg);
int ibmdem_set_mg_params(ibmdem_ctx *ctx, const int32_t *mg);
/* The predictor's schedule, case.stencil_tile (0..3; every schedule gives
 * the same result bit for bit); setting it drops the captured step graphs.
 * No counterpart in the reference. */
int ibmdem_get_stencil_tile(ibmdem_ctx *ctx, int32_t *tile);
int ibmdem_set_stencil_tile(ibmdem_ctx *ctx, int32_t tile);

/* ---- batched form -----------------------------------------------------------
 * Many independent cases of uneven size advanced together on one GPU, each a
 * full context with its own HIP stream; a batch step interleaves the cases'
 * captured step graphs so their kernels overlap, and runs the forcing loop,
 * the Poisson solve and the rest of the step after it for cases that share
 * the precision, the multigrid smoothing counts, the iteration cap, the
 * forcing iteration count and the DEM sub-step count as captured graphs of
 * launches that each cover all of them. Every case's
 * result is bitwise that of stepping it alone. No
 * counterpart in the reference, which runs one case per MPI job (main.F90). */
typedef struct ibmdem_batch ibmdem_batch;
int ibmdem_batch_create(int32_t n, const ibmdem_case *cases, ibmdem_batch **out);
int ibmdem_batch_destroy(ibmdem_batch *batch);
/* the context of case q, for uploads, downloads and single operators */
int ibmdem_batch_context(ibmdem_batch *batch, int32_t q, ibmdem_ctx **ctx);
const char *ibmdem_batch_last_error(const ibmdem_batch *batch);
/* Advance every case n_steps (FirstOrderForwardEuler + CloudEvolutionMS
 * each); info (may be NULL) receives one solve record per case. */
int ibmdem_batch_run(ibmdem_batch *batch, int32_t n_steps, ibmdem_solve_info *info);
/* Time trial_steps batch steps (<= 0: 3) with each candidate multigrid
 * cycle (the same for every case, so the solves stay batched) from the
 * current state and keep the fastest, then the number of solve groups
 * (1, 2, 4), then, per case, each predictor schedule (stencil_tile 0..3);
 * the state is restored after every trial. */
int ibmdem_batch_autotune(ibmdem_batch *batch, int32_t trial_steps);
/* mode 0: every case's step runs on its own stream, as a single context's
 * does (for comparison); 1 (the default): the whole step runs batched (the
 * IBM set-up, the multi-direct-forcing loop, predictor, divergence and the
 * solve's right-hand side; the solve; the rest of the step after it:
 * projection, the DEM phase, probes); 2: only the solve runs batched; 3: the
 * solve and the part after it; 4: the forcing loop, the solve and the part
 * after it (the rest of the first part on each case's own stream). */
int ibmdem_batch_set_fused(ibmdem_batch *batch, int32_t mode);
/* The cases that share precision, smoothing counts and iteration cap solve
 * in `groups` (1..64, default 1) batched groups of about equal cell count,
 * each on its own stream, so one group's solve overlaps the other groups'
 * step parts. */
int ibmdem_batch_set_solve_groups(ibmdem_batch *batch, int32_t groups);
int ibmdem_batch_get_solve_groups(const ibmdem_batch *batch, int32_t *groups);
/* Statistics since creation: group solves that ran batched; ms[0] = GPU
 * time of those solves, ms[1..4] = host time queueing the cases' first
 * parts and the solves, waiting for the solves, queueing the second parts,
 * waiting for the cases' ends (either pointer may be NULL). */
int ibmdem_batch_stats(const ibmdem_batch *batch, int64_t *fused_solves, double *ms);
/* Profiling of the batched solves: while on, each group's solve is queued
 * launch by launch (not from its graph) with an event after every launch,
 * and the launches' run times are added up per kind. profile_read gives a
 * kind's total ms, launches and total algorithmic HBM bytes (0 where not
 * stated). */
#define IBMDEM_BK_SMOOTH0   1   /* level-0 red-black half sweeps, all variants */
#define IBMDEM_BK_SMOOTH    2   /* half sweeps of the coarser levels           */
#define IBMDEM_BK_RESTRICT  3
#define IBMDEM_BK_PROLONG   4
#define IBMDEM_BK_BOTTOM    5   /* one-work-group tails, coarsest-level solves */
#define IBMDEM_BK_MATVEC    6
#define IBMDEM_BK_VECTOR    7   /* CG vector updates (p and x, r with r.r, p = z) */
#define IBMDEM_BK_SCALAR    8   /* CG scalar steps                             */
int ibmdem_batch_profile(ibmdem_batch *batch, int32_t on);
int ibmdem_batch_profile_read(const ibmdem_batch *batch, int32_t kind, double *total_ms, int64_t *launches,
                              double *bytes);

/* ---- measurement --------------------------------------------------------- */
/* Kernel ids for ibmdem_profile_* (HIP events around every launch of that
 * kernel on the context stream; the level-0 smoother's pre- and post-smoothing
 * sweeps are timed as two groups of back-to-back launches). */
#define IBMDEM_K_PREDICTOR   1
#define IBMDEM_K_MG_SMOOTH   2   /* level-0 red-black half sweeps      */
#define IBMDEM_K_MATVEC      3   /* CG matrix-vector product           */
#define IBMDEM_K_SPREAD      4
#define IBMDEM_K_INTERP      5
#define IBMDEM_K_CONTACT     6
#define IBMDEM_K_PROJECT     7
#define IBMDEM_K_DIVERGENCE  8
#define IBMDEM_K_CG_UPDATE   9   /* CG r -= alpha Ap with r.r          */
int ibmdem_profile_kernel(ibmdem_ctx *ctx, int32_t kernel_id);
int ibmdem_profile_read(ibmdem_ctx *ctx, double *total_ms, int64_t *launches,
                        double *bytes_per_launch);
/* Device bytes the context holds. */
int ibmdem_memory_bytes(const ibmdem_ctx *ctx, int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif /* IBMDEM_H */
