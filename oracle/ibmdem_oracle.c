/*
 * ibmdem_oracle.c — CPU FP64 restatement of the reference coupled time step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the CPU
 * baseline: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it, and only as the checker. The product never calls it.
 *
 * It exports the same C ABI as the GPU library (include/ibmdem.h) so the same
 * driver code can run either, but its insides follow the reference Fortran
 * line by line, in the reference's own data layout:
 *   - every field lives on the halo box (lg:ug) = (1-lap : n+lap), lap = 2
 *     (vars.F90:131, allocate.F90:10-66), and halos are filled by a literal
 *     restatement of FillGhostRegionForFluid + Boundary
 *     (boundary.F90:2-247, fluid.F90:240-250) for one rank;
 *   - periodic and wall-mirror ghost markers are real copies built by the
 *     reference's 6-side sweep (ZJS_UPDATE_LIST_DEM.F90:29-177,
 *     ZJS_MPI_DEM.F90:173-269) and refreshed before use;
 *   - markers are binned per halo cell in insertion order and the 125-cell
 *     neighbourhood (ZJS_INIT_DEM.F90:181-214) is scanned kk,jj,ii;
 *   - every floating-point expression keeps the reference's operand order
 *     (Fortran evaluates left to right); build with -ffp-contract=off.
 *
 * Deliberate, documented differences from the reference (one rank, fixed
 * marker order):
 *   - bodies follow the multi-rank semantics: every body is integrated once
 *     with the sum of its own markers (the 1-rank path at
 *     couplesolverCylinder.F90:114-121,130 evolves body 1 only);
 *   - markers keep a fixed order (global id); a marker that crosses a
 *     periodic face is wrapped in place with the same FP operation the
 *     reference's ghost copy uses (x + coordShift) instead of migrating to
 *     the end of the list (ZJS_UPDATE_LIST_DEM.F90:2-27);
 *   - a body's centre seen by a marker is its nearest periodic image, the
 *     copy the marker's rank holds in the multi-rank run (ZJS_MS_DEM.F90:105);
 *   - gravity is a case parameter (EvolutionMS hard-codes (0,0,-9.81),
 *     couplesolverCylinder.F90:12-13);
 *   - PETSc GMRES+BoomerAMG (poisson.F90:133-158) is replaced by the same
 *     multigrid-preconditioned CG the GPU runs (below); the pinned row
 *     (poisson.F90:78,97-100,126) is reproduced exactly by solving the
 *     singular system with b[0] := -sum_{i>0} b[i] and shifting P'(1,1,1)=0.
 * Inflow and outflow sides have no counterpart in the reference (its fluid
 * is periodic or walled). Their rules here are this library's own, the same
 * in the GPU library, and are pinned by closed forms, not by the reference:
 *   - inflow (low sides): the normal velocity is U_n on the face and on every
 *     ghost layer beyond it; tangential ghosts mirror with negation;
 *     pressure ghosts mirror (Neumann);
 *   - outflow (high sides): the normal velocity on the boundary face n+1 is
 *     state of its own (the outflow plane): set to the last interior face
 *     before the divergence, corrected by the projection with P' = 0 on the
 *     face, and copied, with the transverse mirror signs, into ghost layers
 *     n+1 and n+2; tangential ghosts mirror without negation; pressure ghosts
 *     mirror with negation (P = 0 on the face), which makes the Poisson
 *     operator non-singular (no pinned row, no gauge shift);
 *   - markers treat both like walls (mirror images, wall contact).
 * OpenMP: loops whose iterations write disjoint data run in parallel (the
 * spread gives each thread a slab of z planes and walks the parcels in
 * list order, so every cell still receives its terms in that order), and
 * every reduction sums fixed chunks of 4096 terms serially. Results are
 * therefore bitwise independent of the thread count (OMP_NUM_THREADS).
 */
#include "../include/ibmdem.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LAP 2

typedef struct {
    int src;       /* parcel index it was copied from               */
    int dim;       /* 0,1,2                                         */
    int mirror;    /* 0 periodic shift, 1 wall mirror               */
    double shift;  /* coordShift (periodic) or coordMirror (mirror) */
} image_t;

typedef struct {
    int nx, ny, nz;
    int btype[6];
    double h[3];
    double ihx2, ihy2, ihz2;
    int n;              /* cells */
    double *diag, *inv_diag;   /* per cell */
    double *x, *b, *r;  /* level arrays */
    int rb_valid;
} mg_level;

struct ibmdem_ctx {
    ibmdem_case cs;
    char err[256];
    int nx, ny, nz, px, py, pz;   /* px = nx + 2*LAP */
    int btype[6];
    size_t nbox;
    double dx, dy, dz;
    /* fluid fields on the halo box */
    double *u, *v, *w, *P, *Pc, *velDiv, *uold, *vold, *wold;
    double *fx, *fy, *fz, *fxo, *fyo, *fzo, *RHSul, *RHSvl, *RHSwl;
    double *Px, *Py, *Pz;
    double modP[3];
    /* open sides: inflow normal velocity per dimension; outflow planes
     * (interior transverse cells, lower index fastest) */
    double uin[3];
    double *plane[3];
    /* parcels: locals [0,M), images [M,np) */
    int M, B, np, cap;
    double *x, *vel, *fdrag, *fcol;   /* 3*cap, component-major by cap */
    double *vol, *mass, *radeff;
    int *body, *ptype, *cell;         /* cell = halo-box linear index     */
    image_t *img;                     /* cap entries, valid for >= M       */
    /* bins (CSR over halo cells, insertion order) */
    int *cell_start, *cell_items;
    /* send lists (indices) per side */
    int *send[6]; int nsend[6];
    /* pair list for locals */
    int *npairs, *pairs;              /* pairs[m*max_pairs + q] = parcel idx */
    /* bodies */
    double *bx, *bv, *bvold, *bom, *bL, *bLold, *brot, *brotg, *bI, *bm, *brho;
    int *bfixed;
    double *bfdrag, *bfmom, *bfcol, *bfmomcol, *bfhydro, *bfmomhydro;
    double eta;
    /* multigrid */
    int nlev;
    mg_level lev[16];
    double *cg_p, *cg_Ap, *cg_z, *cg_x, *cg_r;
    double L[3];
    /* per-body probe ring (ibmdem_probe_bodies) */
    int probe_n, probe_cap;
    int *probe_list;
    double *probe_ring;
    long long probe_count;
};

/* ------------------------------------------------------------------------ */
static int set_err(ibmdem_ctx *c, int code, const char *msg) {
    if (c) snprintf(c->err, sizeof c->err, "%s", msg);
    return code;
}
#define ID(c, i, j, k) ((size_t)((k) + LAP - 1) * (size_t)(c)->px * (c)->py + \
                        (size_t)((j) + LAP - 1) * (c)->px + (size_t)((i) + LAP - 1))
#define XC(c, p, d) ((c)->x[(size_t)(d) * (c)->cap + (p)])
#define VC(c, p, d) ((c)->vel[(size_t)(d) * (c)->cap + (p)])
#define FD(c, p, d) ((c)->fdrag[(size_t)(d) * (c)->cap + (p)])
#define FC(c, p, d) ((c)->fcol[(size_t)(d) * (c)->cap + (p)])

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

/* Deterministic reductions: partial sums over fixed chunks of RCHUNK terms
 * (in parallel), then the chunk sums in order. */
#define RCHUNK 4096
static double chunk_sum(const double *part, int nc) {
    double s = 0.0;
    for (int q = 0; q < nc; ++q) s += part[q];
    return s;
}

/* ------------------------------------------------------------------------
 * FillGhostRegionForFluid + DealWithBoundaryFluid + Boundary for one rank
 * (boundary.F90:2-247, fluid.F90:240-250). Side order 1..6; the x passes
 * cover interior j,k, the y passes i over the whole halo row, the z passes
 * i and j over the halo plane, so corners propagate exactly as in the
 * reference. Wall: ghost = reversed copy of the first two interior layers,
 * u,v,w negated (boundary.F90:236-244); periodic: ghost = wrapped copy.
 * ------------------------------------------------------------------------ */
static void fill_ghost_one(ibmdem_ctx *c, double *f, int is_vel) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    /* side order of the reference loop: i=1..6, the receiving side is the
     * opposite one (boundary.F90:127-145) */
    for (int s = 1; s <= 6; ++s) {
        int recv = (s % 2 == 0) ? s - 1 : s + 1;
        int dim = (s - 1) / 2;
        const int bt = c->btype[recv - 1];
        /* every non-periodic side mirrors; velocities negate on walls and
         * inflow sides, scalars on outflow sides */
        int wall = bt != IBMDEM_BC_PERIODIC;
        int neg = is_vel ? (bt == IBMDEM_BC_WALL || bt == IBMDEM_BC_INFLOW) : bt == IBMDEM_BC_OUTFLOW;
        int n = dim == 0 ? nx : dim == 1 ? ny : nz;
        /* index ranges of the other two dims */
        int lo1, hi1, lo2, hi2;
        if (dim == 0) { lo1 = 1; hi1 = ny; lo2 = 1; hi2 = nz; }
        else if (dim == 1) { lo1 = 1 - LAP; hi1 = nx + LAP; lo2 = 1; hi2 = nz; }
        else { lo1 = 1 - LAP; hi1 = nx + LAP; lo2 = 1 - LAP; hi2 = ny + LAP; }
#pragma omp parallel for schedule(static)
        for (int b = lo2; b <= hi2; ++b)
            for (int a = lo1; a <= hi1; ++a)
                for (int l = 0; l < LAP; ++l) {
                    int src, dst;
                    if (!wall) {
                        /* periodic: the neighbour (self) sends its face layers */
                        if (recv % 2 == 0) { src = 1 + l; dst = n + 1 + l; }
                        else { src = n - LAP + 1 + l; dst = 1 - LAP + l; }
                    } else {
                        /* DealWithBoundaryFluid(dim, recv): reversed copy */
                        if (recv % 2 == 0) { src = n - LAP + 1 + l; dst = n + LAP - l; }
                        else { src = 1 + l; dst = 0 - l; }
                    }
                    size_t is, id;
                    if (dim == 0) { is = ID(c, src, a, b); id = ID(c, dst, a, b); }
                    else if (dim == 1) { is = ID(c, a, src, b); id = ID(c, a, dst, b); }
                    else { is = ID(c, a, b, src); id = ID(c, a, b, dst); }
                    double val = f[is];
                    if (wall && neg) val = -val;
                    f[id] = val;
                }
    }
}

/* a transverse halo index of the outflow plane mapped onto the interior,
 * with the velocity mirror sign of its side (outflow sides do not negate) */
static int plane_map(const ibmdem_ctx *c, int a, int d, double *sg) {
    const int n = d == 0 ? c->nx : d == 1 ? c->ny : c->nz;
    const int lo = c->btype[2 * d], hi = c->btype[2 * d + 1];
    if (a < 1) {
        if (lo == IBMDEM_BC_PERIODIC) return a + n;
        *sg = -*sg;
        return 1 - a;
    }
    if (a > n) {
        if (hi == IBMDEM_BC_PERIODIC) return a - n;
        if (hi != IBMDEM_BC_OUTFLOW) *sg = -*sg;
        return 2 * n + 1 - a;
    }
    return a;
}
/* outflow plane value of dimension d seen at halo cell (i,j,k) */
static double plane_at(const ibmdem_ctx *c, int d, int i, int j, int k) {
    double sg = 1.0;
    int a[3] = {i, j, k};
    for (int e = 0; e < 3; ++e)
        if (e != d) a[e] = plane_map(c, a[e], e, &sg);
    size_t off = d == 0 ? (size_t)(a[1] - 1) + (size_t)c->ny * (a[2] - 1)
               : d == 1 ? (size_t)(a[0] - 1) + (size_t)c->nx * (a[2] - 1)
                        : (size_t)(a[0] - 1) + (size_t)c->nx * (a[1] - 1);
    return sg * c->plane[d][off];
}

static void boundary_vel(ibmdem_ctx *c) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
    for (int k = 1 - LAP; k <= nz + LAP; ++k)
        for (int j = 1 - LAP; j <= ny + LAP; ++j)
            for (int i = 1 - LAP; i <= nx + LAP; ++i) {
                size_t id = ID(c, i, j, k);
                if (c->btype[0] == 1 && i <= 1) c->u[id] = 0.0;
                if (c->btype[1] == 1 && i >= nx + 1) c->u[id] = 0.0;
                if (c->btype[2] == 1 && j <= 1) c->v[id] = 0.0;
                if (c->btype[3] == 1 && j >= ny + 1) c->v[id] = 0.0;
                if (c->btype[4] == 1 && k <= 1) c->w[id] = 0.0;
                if (c->btype[5] == 1 && k >= nz + 1) c->w[id] = 0.0;
                /* open sides */
                if (c->btype[0] == IBMDEM_BC_INFLOW && i <= 1) c->u[id] = c->uin[0];
                if (c->btype[2] == IBMDEM_BC_INFLOW && j <= 1) c->v[id] = c->uin[1];
                if (c->btype[4] == IBMDEM_BC_INFLOW && k <= 1) c->w[id] = c->uin[2];
                if (c->btype[1] == IBMDEM_BC_OUTFLOW && i >= nx + 1) c->u[id] = plane_at(c, 0, i, j, k);
                if (c->btype[3] == IBMDEM_BC_OUTFLOW && j >= ny + 1) c->v[id] = plane_at(c, 1, i, j, k);
                if (c->btype[5] == IBMDEM_BC_OUTFLOW && k >= nz + 1) c->w[id] = plane_at(c, 2, i, j, k);
            }
}

/* the normal velocity of the last interior face into the outflow plane of
 * dimension d (zero normal gradient) */
static void plane_from_interior(ibmdem_ctx *c, int d) {
    double *f = d == 0 ? c->u : d == 1 ? c->v : c->w;
    if (d == 0)
        for (int k = 1; k <= c->nz; ++k)
            for (int j = 1; j <= c->ny; ++j) c->plane[0][(size_t)(j - 1) + (size_t)c->ny * (k - 1)] = f[ID(c, c->nx, j, k)];
    else if (d == 1)
        for (int k = 1; k <= c->nz; ++k)
            for (int i = 1; i <= c->nx; ++i) c->plane[1][(size_t)(i - 1) + (size_t)c->nx * (k - 1)] = f[ID(c, i, c->ny, k)];
    else
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) c->plane[2][(size_t)(i - 1) + (size_t)c->nx * (j - 1)] = f[ID(c, i, j, c->nz)];
}

static void fill_ghost_fluid(ibmdem_ctx *c) {
    fill_ghost_one(c, c->u, 1);
    fill_ghost_one(c, c->v, 1);
    fill_ghost_one(c, c->w, 1);
    fill_ghost_one(c, c->P, 0);
    fill_ghost_one(c, c->Pc, 0);
    boundary_vel(c);
}

/* ------------------------------------------------------------------------
 * DeltaFunction3D (interpolation.F90:2-33) and GetWeight
 * (ZJS_INTERPOLATE_DEM.F90:2-17). Roma-Peskin-Berger 3-point kernel.
 * ------------------------------------------------------------------------ */
static double phi1(double r) {
    if (r <= 0.5) return (1.0 + sqrt(1.0 - 3.0 * r * r)) / 3.0;
    if (r <= 1.5) return (5.0 - 3.0 * r - sqrt(1.0 - 3.0 * ((1.0 - r) * (1.0 - r)))) / 6.0;
    return 0.0;
}
static double delta3d(const ibmdem_ctx *c, double x1, double x2, double x3) {
    double w1 = phi1(fabs(x1 / c->dx));
    double w2 = phi1(fabs(x2 / c->dy));
    double w3 = phi1(fabs(x3 / c->dz));
    return w1 * w2 * w3 / (c->dx * c->dy * c->dz);
}
static void get_weight(const ibmdem_ctx *c, const double xp[3], const double xc[3],
                       double wt[3]) {
    double d1 = xp[0] - xc[0], d2 = xp[1] - xc[1], d3 = xp[2] - xc[2];
    wt[0] = delta3d(c, d1 + 0.5 * c->dx, d2, d3);
    wt[1] = delta3d(c, d1, d2 + 0.5 * c->dy, d3);
    wt[2] = delta3d(c, d1, d2, d3 + 0.5 * c->dz);
}

/* cell of a position (ZJS_UPDATE_LIST_DEM.F90:41-43), 1-based halo coords */
static void cell_of(const ibmdem_ctx *c, const double xp[3], int ijk[3]) {
    ijk[0] = (int)floor((xp[0] - 0.0) / c->dx) + 1;
    ijk[1] = (int)floor((xp[1] - 0.0) / c->dy) + 1;
    ijk[2] = (int)floor((xp[2] - 0.0) / c->dz) + 1;
}
static int in_box(const ibmdem_ctx *c, int i, int j, int k) {
    return i >= 1 - LAP && i <= c->nx + LAP && j >= 1 - LAP && j <= c->ny + LAP &&
           k >= 1 - LAP && k <= c->nz + LAP;
}
static void unpack_cell(const ibmdem_ctx *c, int id, int *i, int *j, int *k) {
    *i = id % c->px + 1 - LAP;
    *j = (id / c->px) % c->py + 1 - LAP;
    *k = id / (c->px * c->py) + 1 - LAP;
}

/* ------------------------------------------------------------------------
 * Parcel lists: CompressDEM (periodic wrap in place), PutParcelsIntoCellsDEM,
 * RebuildParcelInCellListForGhostRegionDEM (+ UpdateFullInfoForSolid...),
 * RebuildParcelCollisionListDEM, Collision.
 * ------------------------------------------------------------------------ */
static void send_push(ibmdem_ctx *c, int side, int p) {
    c->send[side][c->nsend[side]++] = p;
}

static void put_into_cells(ibmdem_ctx *c, int p0, int p1, int mode) {
    for (int p = p0; p < p1; ++p) {
        double xp[3] = {XC(c, p, 0), XC(c, p, 1), XC(c, p, 2)};
        int ijk[3];
        cell_of(c, xp, ijk);
        c->cell[p] = (int)ID(c, ijk[0], ijk[1], ijk[2]);
        if (mode <= 0) {
            if (ijk[0] >= 1 && ijk[0] <= 2) send_push(c, 0, p);
            if (ijk[0] >= c->nx - 1 && ijk[0] <= c->nx) send_push(c, 1, p);
        }
        if (mode <= 1) {
            if (ijk[1] >= 1 && ijk[1] <= 2) send_push(c, 2, p);
            if (ijk[1] >= c->ny - 1 && ijk[1] <= c->ny) send_push(c, 3, p);
        }
        if (mode <= 2) {
            if (ijk[2] >= 1 && ijk[2] <= 2) send_push(c, 4, p);
            if (ijk[2] >= c->nz - 1 && ijk[2] <= c->nz) send_push(c, 5, p);
        }
    }
}

/* Refresh one image from its source (ZJS_MPI_DEM.F90:222-269 and the
 * periodic branch of UpdateParcelInfoInGhostRegion). */
static void refresh_image(ibmdem_ctx *c, int p) {
    const image_t *im = &c->img[p];
    int s = im->src;
    for (int d = 0; d < 3; ++d) {
        XC(c, p, d) = XC(c, s, d);
        VC(c, p, d) = VC(c, s, d);
        FD(c, p, d) = FD(c, s, d);
    }
    int d = im->dim;
    if (im->mirror) {
        XC(c, p, d) = 2.0 * im->shift - XC(c, p, d);
        VC(c, p, d) = -VC(c, p, d);
    } else {
        XC(c, p, d) = XC(c, p, d) + im->shift;
    }
}
static void refresh_images(ibmdem_ctx *c) {
    for (int p = c->M; p < c->np; ++p) refresh_image(c, p);
}

static int build_images(ibmdem_ctx *c) {
    const double Lg[3] = {c->L[0], c->L[1], c->L[2]};
    for (int idx = 1; idx <= 6; ++idx) {
        int idxSend = idx, idxRecv = (idx % 2 == 1) ? idx + 1 : idx - 1;
        int dim = (idx - 1) / 2;
        int start = c->np;
        if (c->btype[idxSend - 1] == IBMDEM_BC_PERIODIC &&
            c->btype[idxRecv - 1] == IBMDEM_BC_PERIODIC) {
            /* coordShift(idxSend): +L for odd sides, -L for even sides */
            double shift = (idxSend % 2 == 1) ? Lg[dim] : -Lg[dim];
            for (int q = 0; q < c->nsend[idxSend - 1]; ++q) {
                int s = c->send[idxSend - 1][q];
                int p = c->np++;
                if (p >= c->cap) return set_err(c, IBMDEM_ERR_CAPACITY, "parcel capacity");
                c->img[p].src = s; c->img[p].dim = dim; c->img[p].mirror = 0;
                c->img[p].shift = shift;
                c->vol[p] = c->vol[s]; c->mass[p] = c->mass[s]; c->radeff[p] = c->radeff[s];
                c->body[p] = c->body[s]; c->ptype[p] = c->ptype[s];
                refresh_image(c, p);
            }
        } else if (c->btype[idxRecv - 1] != IBMDEM_BC_PERIODIC) {
            /* UpdateFullInfoForSolidOutletBoundary(dim, idxRecv):
             * coordMirror = 0 on low sides, L on high sides; open sides
             * mirror like the reference's outlet (ZJS_MPI_DEM.F90:203-209) */
            double mirror = (idxRecv % 2 == 1) ? 0.0 : Lg[dim];
            for (int q = 0; q < c->nsend[idxRecv - 1]; ++q) {
                int s = c->send[idxRecv - 1][q];
                int p = c->np++;
                if (p >= c->cap) return set_err(c, IBMDEM_ERR_CAPACITY, "parcel capacity");
                c->img[p].src = s; c->img[p].dim = dim; c->img[p].mirror = 1;
                c->img[p].shift = mirror;
                c->vol[p] = c->vol[s]; c->mass[p] = c->mass[s]; c->radeff[p] = c->radeff[s];
                c->body[p] = c->body[s]; c->ptype[p] = 1;
                refresh_image(c, p);
            }
        }
        put_into_cells(c, start, c->np, dim + 1);
    }
    return IBMDEM_OK;
}

static int collision_flag(const ibmdem_ctx *c, int i, int j) {
    double x1 = XC(c, i, 0) - XC(c, j, 0);
    double x2 = XC(c, i, 1) - XC(c, j, 1);
    double x3 = XC(c, i, 2) - XC(c, j, 2);
    double d2 = x1 * x1 + x2 * x2 + x3 * x3;
    double lim = c->cs.collision_dist * (1.0 + c->cs.collision_buffer);
    return d2 < lim * lim;
}

static int rebuild_lists(ibmdem_ctx *c) {
    /* CompressDEM: markers whose cell left the interior through a periodic
     * face take the position of their ghost copy, x + coordShift */
    int outside = 0, nonfinite = 0;
#pragma omp parallel for schedule(static) reduction(| : outside, nonfinite)
    for (int p = 0; p < c->M; ++p) {
        double xp[3] = {XC(c, p, 0), XC(c, p, 1), XC(c, p, 2)};
        int ijk[3];
        /* a non-finite position (no counterpart: the reference would bin it
         * wherever the conversion lands) is its own error */
        if (!(isfinite(xp[0]) && isfinite(xp[1]) && isfinite(xp[2]))) {
            nonfinite = 1;
            c->ptype[p] = -1;
            continue;
        }
        cell_of(c, xp, ijk);
        int n[3] = {c->nx, c->ny, c->nz};
        for (int d = 0; d < 3; ++d) {
            if (ijk[d] > n[d]) {
                if (c->btype[2 * d + 1] != IBMDEM_BC_PERIODIC) outside = 1;
                else XC(c, p, d) = XC(c, p, d) + (-c->L[d]);
            } else if (ijk[d] < 1) {
                if (c->btype[2 * d] != IBMDEM_BC_PERIODIC) outside = 1;
                else XC(c, p, d) = XC(c, p, d) + c->L[d];
            }
        }
        xp[0] = XC(c, p, 0); xp[1] = XC(c, p, 1); xp[2] = XC(c, p, 2);
        cell_of(c, xp, ijk);
        if (ijk[0] < 1 || ijk[0] > c->nx || ijk[1] < 1 || ijk[1] > c->ny ||
            ijk[2] < 1 || ijk[2] > c->nz)
            outside = 1;
        c->ptype[p] = -1;
    }
    if (nonfinite) return set_err(c, IBMDEM_ERR_NAN, "non-finite marker position");
    if (outside) return set_err(c, IBMDEM_ERR_OUTSIDE, "marker outside the domain");
    c->np = c->M;
    for (int s = 0; s < 6; ++s) c->nsend[s] = 0;
    put_into_cells(c, 0, c->M, 0);
    int rc = build_images(c);
    if (rc) return rc;
    /* bins, insertion order (a stable counting sort of the parcel order) */
    memset(c->cell_start, 0, (c->nbox + 1) * sizeof(int));
    for (int p = 0; p < c->np; ++p) c->cell_start[c->cell[p] + 1]++;
    for (size_t q = 0; q < c->nbox; ++q) c->cell_start[q + 1] += c->cell_start[q];
    int *cur = (int *)xcalloc(c->nbox, sizeof(int));
    for (int p = 0; p < c->np; ++p) {
        int cl = c->cell[p];
        c->cell_items[c->cell_start[cl] + cur[cl]++] = p;
    }
    free(cur);
    /* RebuildParcelCollisionListDEM (ZJS_UPDATE_LIST_DEM.F90:248-286) */
    int over = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : over)
    for (int i = 0; i < c->M; ++i) {
        int ci, cj, ck, cnt = 0;
        unpack_cell(c, c->cell[i], &ci, &cj, &ck);
        for (int kk = -2; kk <= 2; ++kk)
            for (int jj = -2; jj <= 2; ++jj)
                for (int ii = -2; ii <= 2; ++ii) {
                    if (!in_box(c, ci + ii, cj + jj, ck + kk)) continue;
                    size_t cl = ID(c, ci + ii, cj + jj, ck + kk);
                    for (int q = c->cell_start[cl]; q < c->cell_start[cl + 1]; ++q) {
                        int j = c->cell_items[q];
                        if (c->body[i] != c->body[j] && collision_flag(c, i, j) &&
                            c->ptype[j] == -1 && i != j) {
                            if (cnt >= c->cs.max_pairs) { over = 1; continue; }
                            c->pairs[(size_t)i * c->cs.max_pairs + cnt++] = j;
                        }
                    }
                }
        c->npairs[i] = cnt;
    }
    if (over) return set_err(c, IBMDEM_ERR_CAPACITY, "pair list capacity");
    return IBMDEM_OK;
}

/* ------------------------------------------------------------------------
 * IBM interpolation and spreading (ZJS_VOLUMEFRACTION_DEM.F90:2-67)
 * ------------------------------------------------------------------------ */
static void cell_centre(const ibmdem_ctx *c, int i, int j, int k, double xc[3]) {
    xc[0] = (i - 0.5) * c->dx;
    xc[1] = (j - 0.5) * c->dy;
    xc[2] = (k - 0.5) * c->dz;
}

static void ibm_interpolate(ibmdem_ctx *c) {
    const double dV1 = c->dx, dV2 = c->dy, dV3 = c->dz;
#pragma omp parallel for schedule(dynamic, 256)
    for (int p = 0; p < c->M; ++p) {
        int ci, cj, ck;
        unpack_cell(c, c->cell[p], &ci, &cj, &ck);
        double xp[3] = {XC(c, p, 0), XC(c, p, 1), XC(c, p, 2)};
        double ulp = 0.0, vlp = 0.0, wlp = 0.0;
        for (int kk = -2; kk <= 2; ++kk)
            for (int jj = -2; jj <= 2; ++jj)
                for (int ii = -2; ii <= 2; ++ii) {
                    int i = ci + ii, j = cj + jj, k = ck + kk;
                    if (!in_box(c, i, j, k)) continue;
                    double xc[3], wt[3];
                    cell_centre(c, i, j, k, xc);
                    get_weight(c, xp, xc, wt);
                    size_t id = ID(c, i, j, k);
                    ulp = ulp + wt[0] * c->u[id] * dV1 * dV2 * dV3;
                    vlp = vlp + wt[1] * c->v[id] * dV1 * dV2 * dV3;
                    wlp = wlp + wt[2] * c->w[id] * dV1 * dV2 * dV3;
                }
        FD(c, p, 0) = FD(c, p, 0) + 1.0 * (ulp - VC(c, p, 0)) / c->cs.dt;
        FD(c, p, 1) = FD(c, p, 1) + 1.0 * (vlp - VC(c, p, 1)) / c->cs.dt;
        FD(c, p, 2) = FD(c, p, 2) + 1.0 * (wlp - VC(c, p, 2)) / c->cs.dt;
    }
}

static void ibm_spread(ibmdem_ctx *c) {
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->nbox; ++q) {
        c->fxo[q] = c->fx[q]; c->fyo[q] = c->fy[q]; c->fzo[q] = c->fz[q];
        c->fx[q] = 0.0; c->fy[q] = 0.0; c->fz[q] = 0.0;
    }
    /* Thread t owns the halo planes [k0, k1] and walks every parcel in list
     * order, adding only into its own planes: each cell receives its terms
     * in parcel-list order, exactly as the serial loop. */
    const int kmin = 1 - LAP, kmax = c->nz + LAP;
#pragma omp parallel
    {
        int nt = 1, t = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        t = omp_get_thread_num();
#endif
        const int span = kmax - kmin + 1;
        const int k0 = kmin + (int)((long)span * t / nt), k1 = kmin + (int)((long)span * (t + 1) / nt) - 1;
    for (int p = 0; p < c->np; ++p) {
        int ci, cj, ck;
        unpack_cell(c, c->cell[p], &ci, &cj, &ck);
        if (ck + 2 < k0 || ck - 2 > k1) continue;
        double xp[3] = {XC(c, p, 0), XC(c, p, 1), XC(c, p, 2)};
        for (int kk = -2; kk <= 2; ++kk)
            for (int jj = -2; jj <= 2; ++jj)
                for (int ii = -2; ii <= 2; ++ii) {
                    int i = ci + ii, j = cj + jj, k = ck + kk;
                    if (k < k0 || k > k1) continue;
                    if (!in_box(c, i, j, k)) continue;
                    double xc[3], wt[3];
                    cell_centre(c, i, j, k, xc);
                    get_weight(c, xp, xc, wt);
                    size_t id = ID(c, i, j, k);
                    c->fx[id] = c->fx[id] - FD(c, p, 0) * c->vol[p] * wt[0];
                    c->fy[id] = c->fy[id] - FD(c, p, 1) * c->vol[p] * wt[1];
                    c->fz[id] = c->fz[id] - FD(c, p, 2) * c->vol[p] * wt[2];
                }
    }
    }
}

/* VelocityPostGuess (fluid.F90:135-153) */
static void velocity_post_guess(ibmdem_ctx *c) {
    const double dt = c->cs.dt;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t id = ID(c, i, j, k);
                c->u[id] = c->u[id] + dt * (c->fx[id] - c->fxo[id]);
                c->v[id] = c->v[id] + dt * (c->fy[id] - c->fyo[id]);
                c->w[id] = c->w[id] + dt * (c->fz[id] - c->fzo[id]);
            }
}

/* ------------------------------------------------------------------------
 * VelocityGuess(a,b,c) with ConvectionTensor, ViscousTensor and
 * PressureGradient (fluid.F90:2-133). All tensors are evaluated from the
 * velocity before the update; the expression order is the reference's.
 * ------------------------------------------------------------------------ */
static void velocity_guess(ibmdem_ctx *c, double aRK, double bRK, double cRK) {
    const double dx = c->dx, dy = c->dy, dz = c->dz, visc = c->cs.visc, rhof = c->cs.rhof;
    const double dt = c->cs.dt;
    const size_t sx = 1, sy = (size_t)c->px, sz = (size_t)c->px * c->py;
    double *RHSu = (double *)xcalloc(c->nbox, sizeof(double));
    double *RHSv = (double *)xcalloc(c->nbox, sizeof(double));
    double *RHSw = (double *)xcalloc(c->nbox, sizeof(double));
    const double *u = c->u, *v = c->v, *w = c->w, *P = c->P;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                double uuE = 0.5 * (u[o + sx] + u[o]) * 0.5 * (u[o + sx] + u[o]);
                double uuW = 0.5 * (u[o - sx] + u[o]) * 0.5 * (u[o - sx] + u[o]);
                double uvB = 0.5 * (u[o + sy] + u[o]) * 0.5 * (v[o - sx + sy] + v[o + sy]);
                double uvF = 0.5 * (u[o - sy] + u[o]) * 0.5 * (v[o - sx] + v[o]);
                double uwN = 0.5 * (u[o + sz] + u[o]) * 0.5 * (w[o - sx + sz] + w[o + sz]);
                double uwS = 0.5 * (u[o - sz] + u[o]) * 0.5 * (w[o - sx] + w[o]);

                double vuE = 0.5 * (v[o + sx] + v[o]) * 0.5 * (u[o + sx - sy] + u[o + sx]);
                double vuW = 0.5 * (v[o - sx] + v[o]) * 0.5 * (u[o - sy] + u[o]);
                double vvB = 0.5 * (v[o + sy] + v[o]) * 0.5 * (v[o + sy] + v[o]);
                double vvF = 0.5 * (v[o - sy] + v[o]) * 0.5 * (v[o - sy] + v[o]);
                double vwN = 0.5 * (v[o + sz] + v[o]) * 0.5 * (w[o - sy + sz] + w[o + sz]);
                double vwS = 0.5 * (v[o - sz] + v[o]) * 0.5 * (w[o - sy] + w[o]);

                double wuE = 0.5 * (w[o + sx] + w[o]) * 0.5 * (u[o + sx - sz] + u[o + sx]);
                double wuW = 0.5 * (w[o - sx] + w[o]) * 0.5 * (u[o - sz] + u[o]);
                double wvB = 0.5 * (w[o + sy] + w[o]) * 0.5 * (v[o + sy - sz] + v[o + sy]);
                double wvF = 0.5 * (w[o - sy] + w[o]) * 0.5 * (v[o - sz] + v[o]);
                double wwN = 0.5 * (w[o + sz] + w[o]) * 0.5 * (w[o + sz] + w[o]);
                double wwS = 0.5 * (w[o - sz] + w[o]) * 0.5 * (w[o - sz] + w[o]);

                double TIux = (uuE - uuW) / dx, TIuy = (uvB - uvF) / dy, TIuz = (uwN - uwS) / dz;
                double TIvx = (vuE - vuW) / dx, TIvy = (vvB - vvF) / dy, TIvz = (vwN - vwS) / dz;
                double TIwx = (wuE - wuW) / dx, TIwy = (wvB - wvF) / dy, TIwz = (wwN - wwS) / dz;

                double TVux = (u[o + sx] + u[o - sx] - 2 * u[o]) / (dx * dx) * visc;
                double TVuy = (u[o + sy] + u[o - sy] - 2 * u[o]) / (dy * dy) * visc;
                double TVuz = (u[o + sz] + u[o - sz] - 2 * u[o]) / (dz * dz) * visc;
                double TVvx = (v[o + sx] + v[o - sx] - 2 * v[o]) / (dx * dx) * visc;
                double TVvy = (v[o + sy] + v[o - sy] - 2 * v[o]) / (dy * dy) * visc;
                double TVvz = (v[o + sz] + v[o - sz] - 2 * v[o]) / (dz * dz) * visc;
                double TVwx = (w[o + sx] + w[o - sx] - 2 * w[o]) / (dx * dx) * visc;
                double TVwy = (w[o + sy] + w[o - sy] - 2 * w[o]) / (dy * dy) * visc;
                double TVwz = (w[o + sz] + w[o - sz] - 2 * w[o]) / (dz * dz) * visc;

                c->Px[o] = (P[o] - P[o - sx]) / (dx) / rhof;
                c->Py[o] = (P[o] - P[o - sy]) / (dy) / rhof;
                c->Pz[o] = (P[o] - P[o - sz]) / (dz) / rhof;

                /* the pressure gradient on an outflow face, for the
                 * projection of the outflow plane */
                if (i == c->nx && c->btype[1] == IBMDEM_BC_OUTFLOW) c->Px[o + sx] = (P[o + sx] - P[o]) / (dx) / rhof;
                if (j == c->ny && c->btype[3] == IBMDEM_BC_OUTFLOW) c->Py[o + sy] = (P[o + sy] - P[o]) / (dy) / rhof;
                if (k == c->nz && c->btype[5] == IBMDEM_BC_OUTFLOW) c->Pz[o + sz] = (P[o + sz] - P[o]) / (dz) / rhof;

                RHSu[o] = -(TIux + TIuy + TIuz) + (TVux + TVuy + TVuz) + c->cs.cp[0] + c->fx[o];
                RHSv[o] = -(TIvx + TIvy + TIvz) + (TVvx + TVvy + TVvz) + c->cs.cp[1] + c->fy[o];
                RHSw[o] = -(TIwx + TIwy + TIwz) + (TVwx + TVwy + TVwz) + c->cs.cp[2] + c->fz[o];
            }
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                c->u[o] = c->uold[o] + dt * (aRK * RHSu[o] + bRK * c->RHSul[o] - cRK * c->Px[o]);
                c->v[o] = c->vold[o] + dt * (aRK * RHSv[o] + bRK * c->RHSvl[o] - cRK * c->Py[o]);
                c->w[o] = c->wold[o] + dt * (aRK * RHSw[o] + bRK * c->RHSwl[o] - cRK * c->Pz[o]);
                c->RHSul[o] = RHSu[o];
                c->RHSvl[o] = RHSv[o];
                c->RHSwl[o] = RHSw[o];
            }
    free(RHSu); free(RHSv); free(RHSw);
}

/* ModifyBulkVelocity (fluid.F90:155-197); SUM in array element order */
static void modify_bulk(ibmdem_ctx *c) {
    const int *sw = c->cs.bulk_switch;
    if (!(sw[0] == 1 || sw[1] == 1 || sw[2] == 1)) return;
    /* SUM in array element order, as fixed chunks of RCHUNK elements */
    const int nTot0 = c->nx * c->ny * c->nz, nc = (nTot0 + RCHUNK - 1) / RCHUNK;
    double *part = (double *)xcalloc(3 * (size_t)nc, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int ch = 0; ch < nc; ++ch) {
        double a = 0.0, b = 0.0, e = 0.0;
        const int hi = (ch + 1) * RCHUNK < nTot0 ? (ch + 1) * RCHUNK : nTot0;
        for (int q = ch * RCHUNK; q < hi; ++q) {
            const int i = q % c->nx + 1, j = (q / c->nx) % c->ny + 1, k = q / (c->nx * c->ny) + 1;
            size_t o = ID(c, i, j, k);
            a += c->u[o]; b += c->v[o]; e += c->w[o];
        }
        part[ch] = a; part[nc + ch] = b; part[2 * (size_t)nc + ch] = e;
    }
    double us = chunk_sum(part, nc), vs = chunk_sum(part + nc, nc), ws = chunk_sum(part + 2 * (size_t)nc, nc);
    free(part);
    int nTot = c->nx * c->ny * c->nz;
    c->modP[0] = c->modP[1] = c->modP[2] = 0.0;
    const double dt = c->cs.dt;
    if (sw[0] == 1) c->modP[0] = -(us / nTot - c->cs.bulk_target[0]) / dt;
    if (sw[1] == 1) c->modP[1] = -(vs / nTot - c->cs.bulk_target[1]) / dt;
    if (sw[2] == 1) c->modP[2] = -(ws / nTot - c->cs.bulk_target[2]) / dt;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                c->u[o] = c->u[o] + dt * c->modP[0];
                c->v[o] = c->v[o] + dt * c->modP[1];
                c->w[o] = c->w[o] + dt * c->modP[2];
            }
}

/* VelocityDivergence(c) (fluid.F90:89-105); outflow planes first take the
 * last interior face */
static void velocity_divergence(ibmdem_ctx *c, double cRK) {
    const size_t sx = 1, sy = (size_t)c->px, sz = (size_t)c->px * c->py;
    int refill = 0;
    for (int d = 0; d < 3; ++d)
        if (c->btype[2 * d + 1] == IBMDEM_BC_OUTFLOW) { plane_from_interior(c, d); refill = 1; }
    if (refill) boundary_vel(c);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                double d = (c->u[o + sx] - c->u[o]) / c->dx + (c->v[o + sy] - c->v[o]) / c->dy +
                           (c->w[o + sz] - c->w[o]) / c->dz;
                c->velDiv[o] = -d / c->cs.dt / cRK * c->cs.rhof;
            }
}

/* PressureCorrection + VelocityCorrection(c) (fluid.F90:199-238) */
static void pressure_correction(ibmdem_ctx *c) {
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                c->P[o] = c->P[o] + c->Pc[o];
            }
}
static void velocity_correction(ibmdem_ctx *c, double cRK) {
    const size_t sx = 1, sy = (size_t)c->px, sz = (size_t)c->px * c->py;
    const double dt = c->cs.dt, rhof = c->cs.rhof;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) {
                size_t o = ID(c, i, j, k);
                const double *P = c->P;
                c->u[o] = c->u[o] + cRK * c->Px[o] * dt - cRK * (P[o] - P[o - sx]) / (c->dx) * dt / rhof;
                c->v[o] = c->v[o] + cRK * c->Py[o] * dt - cRK * (P[o] - P[o - sy]) / (c->dy) * dt / rhof;
                c->w[o] = c->w[o] + cRK * c->Pz[o] * dt - cRK * (P[o] - P[o - sz]) / (c->dz) * dt / rhof;
                /* the outflow planes, by the same formula on face n+1 */
                if (i == c->nx && c->btype[1] == IBMDEM_BC_OUTFLOW) {
                    double *q = &c->plane[0][(size_t)(j - 1) + (size_t)c->ny * (k - 1)];
                    *q = *q + cRK * c->Px[o + sx] * dt - cRK * (P[o + sx] - P[o]) / (c->dx) * dt / rhof;
                }
                if (j == c->ny && c->btype[3] == IBMDEM_BC_OUTFLOW) {
                    double *q = &c->plane[1][(size_t)(i - 1) + (size_t)c->nx * (k - 1)];
                    *q = *q + cRK * c->Py[o + sy] * dt - cRK * (P[o + sy] - P[o]) / (c->dy) * dt / rhof;
                }
                if (k == c->nz && c->btype[5] == IBMDEM_BC_OUTFLOW) {
                    double *q = &c->plane[2][(size_t)(i - 1) + (size_t)c->nx * (j - 1)];
                    *q = *q + cRK * c->Pz[o + sz] * dt - cRK * (P[o + sz] - P[o]) / (c->dz) * dt / rhof;
                }
            }
}

/* ------------------------------------------------------------------------
 * Poisson: multigrid-preconditioned CG on interior (compact) arrays.
 * Operator: the reference's 7-point -lap_h with wall neighbours omitted and
 * periodic neighbours wrapped (poisson.F90:74-94, initialization.F90:105-173).
 * (Ax)_i = diag_i x_i - nbsum_i; nbsum = (xw+xe)/hx^2 + (xs+xn)/hy^2 +
 * (xb+xt)/hz^2 with missing (wall) neighbours contributing 0.0.
 * Smoother: red-black Gauss-Seidel, pre red->black, post black->red; the
 * coarsest level runs mg_coarse_sweeps symmetric pairs (damped Jacobi if a
 * periodic dimension is odd). Transfers: cell-centred trilinear
 * prolongation (3/4,1/4 per dim) and restriction = prolongation^T / 8, both
 * with the wall mirror / periodic wrap index rules.
 * ------------------------------------------------------------------------ */
static int wrap_or_mirror(int i, int n, int lo_wall, int hi_wall, int *valid) {
    *valid = 1;
    if (i < 0) { if (lo_wall) { *valid = 0; return 0; } return i + n; }
    if (i >= n) { if (hi_wall) { *valid = 0; return n - 1; } return i - n; }
    return i;
}

static inline double nb_val(const mg_level *L, const double *x, int i, int j, int k, int d, int side) {
    int n = d == 0 ? L->nx : d == 1 ? L->ny : L->nz;
    int *t = d == 0 ? &i : d == 1 ? &j : &k;
    int a = *t + (side ? 1 : -1);
    if (a < 0) { if (L->btype[2 * d] != IBMDEM_BC_PERIODIC) return 0.0; a += n; }
    else if (a >= n) { if (L->btype[2 * d + 1] != IBMDEM_BC_PERIODIC) return 0.0; a -= n; }
    *t = a;
    return x[((size_t)k * L->ny + j) * L->nx + i];
}

/* nbsum = (xw+xe)/hx^2 + (xs+xn)/hy^2 + (xb+xt)/hz^2, wall neighbours 0.0 */
static double nbsum_at(const mg_level *L, const double *x, int i, int j, int k) {
    double sx = (nb_val(L, x, i, j, k, 0, 0) + nb_val(L, x, i, j, k, 0, 1)) * L->ihx2;
    double sy = (nb_val(L, x, i, j, k, 1, 0) + nb_val(L, x, i, j, k, 1, 1)) * L->ihy2;
    double sz = (nb_val(L, x, i, j, k, 2, 0) + nb_val(L, x, i, j, k, 2, 1)) * L->ihz2;
    return sx + sy + sz;
}

static void mg_setup_level(mg_level *L) {
    L->n = L->nx * L->ny * L->nz;
    L->ihx2 = 1.0 / (L->h[0] * L->h[0]);
    L->ihy2 = 1.0 / (L->h[1] * L->h[1]);
    L->ihz2 = 1.0 / (L->h[2] * L->h[2]);
    L->inv_diag = (double *)xcalloc(L->n, sizeof(double));
    L->diag = (double *)xcalloc(L->n, sizeof(double));
    L->x = (double *)xcalloc(L->n, sizeof(double));
    L->b = (double *)xcalloc(L->n, sizeof(double));
    L->r = (double *)xcalloc(L->n, sizeof(double));
    int nn[3] = {L->nx, L->ny, L->nz};
    L->rb_valid = 1;
    for (int d = 0; d < 3; ++d)
        if (L->btype[2 * d] == IBMDEM_BC_PERIODIC && (nn[d] % 2))
            L->rb_valid = 0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->nz; ++k)
        for (int j = 0; j < L->ny; ++j)
            for (int i = 0; i < L->nx; ++i) {
                int pos[3] = {i, j, k};
                double ih[3] = {L->ihx2, L->ihy2, L->ihz2};
                double dsum[3];
                for (int d = 0; d < 3; ++d) {
                    int cnt = 0;
                    if (pos[d] > 0 || L->btype[2 * d] == IBMDEM_BC_PERIODIC) cnt++;
                    if (pos[d] < nn[d] - 1 || L->btype[2 * d + 1] == IBMDEM_BC_PERIODIC) cnt++;
                    /* outflow: ghost = -x, (x_g - x)/h^2 = -2x/h^2 */
                    if (pos[d] == nn[d] - 1 && L->btype[2 * d + 1] == IBMDEM_BC_OUTFLOW) cnt += 2;
                    dsum[d] = cnt * ih[d];
                }
                double diag = dsum[0] + dsum[1] + dsum[2];
                L->diag[((size_t)k * L->ny + j) * L->nx + i] = diag;
                L->inv_diag[((size_t)k * L->ny + j) * L->nx + i] = 1.0 / diag;
            }
}

static void mg_apply(const mg_level *L, const double *x, double *ax) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->nz; ++k)
        for (int j = 0; j < L->ny; ++j)
            for (int i = 0; i < L->nx; ++i) {
                size_t id = ((size_t)k * L->ny + j) * L->nx + i;
                ax[id] = L->diag[id] * x[id] - nbsum_at(L, x, i, j, k);
            }
}

static void rb_half(mg_level *L, int color) {
    /* cells of one colour only read the other colour: any order */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->nz; ++k)
        for (int j = 0; j < L->ny; ++j)
            for (int i = 0; i < L->nx; ++i) {
                if (((i + j + k) & 1) != color) continue;
                size_t id = ((size_t)k * L->ny + j) * L->nx + i;
                L->x[id] = (L->b[id] + nbsum_at(L, L->x, i, j, k)) * L->inv_diag[id];
            }
}
static void jacobi_sweep(mg_level *L, double omega) {
    double *t = (double *)xcalloc(L->n, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->nz; ++k)
        for (int j = 0; j < L->ny; ++j)
            for (int i = 0; i < L->nx; ++i) {
                size_t id = ((size_t)k * L->ny + j) * L->nx + i;
                double xn = (L->b[id] + nbsum_at(L, L->x, i, j, k)) * L->inv_diag[id];
                t[id] = L->x[id] + omega * (xn - L->x[id]);
            }
    memcpy(L->x, t, L->n * sizeof(double));
    free(t);
}

static void restrict_to(const mg_level *F, const double *rf, mg_level *C) {
    static const double wt[4] = {0.25, 0.75, 0.75, 0.25};
#pragma omp parallel for schedule(static)
    for (int K = 0; K < C->nz; ++K)
        for (int J = 0; J < C->ny; ++J)
            for (int I = 0; I < C->nx; ++I) {
                double s = 0.0;
                for (int c3 = 0; c3 < 4; ++c3)
                    for (int c2 = 0; c2 < 4; ++c2)
                        for (int c1 = 0; c1 < 4; ++c1) {
                            int ok;
                            int fi = wrap_or_mirror(2 * I - 1 + c1, F->nx, 0, 0, &ok);
                            int fj = wrap_or_mirror(2 * J - 1 + c2, F->ny, 0, 0, &ok);
                            int fk = wrap_or_mirror(2 * K - 1 + c3, F->nz, 0, 0, &ok);
                            /* walls mirror the out-of-range fine index */
                            int ti = 2 * I - 1 + c1, tj = 2 * J - 1 + c2, tk = 2 * K - 1 + c3;
                            if (ti < 0 && F->btype[0] != IBMDEM_BC_PERIODIC) fi = 0;
                            if (ti >= F->nx && F->btype[1] != IBMDEM_BC_PERIODIC) fi = F->nx - 1;
                            if (tj < 0 && F->btype[2] != IBMDEM_BC_PERIODIC) fj = 0;
                            if (tj >= F->ny && F->btype[3] != IBMDEM_BC_PERIODIC) fj = F->ny - 1;
                            if (tk < 0 && F->btype[4] != IBMDEM_BC_PERIODIC) fk = 0;
                            if (tk >= F->nz && F->btype[5] != IBMDEM_BC_PERIODIC) fk = F->nz - 1;
                            double w = wt[c3] * wt[c2] * wt[c1];
                            s = s + w * rf[((size_t)fk * F->ny + fj) * F->nx + fi];
                        }
                C->b[((size_t)K * C->ny + J) * C->nx + I] = 0.125 * s;
            }
}

static void prolong_add(const mg_level *C, mg_level *F) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < F->nz; ++k)
        for (int j = 0; j < F->ny; ++j)
            for (int i = 0; i < F->nx; ++i) {
                int pi[2], pj[2], pk[2];
                int f[3] = {i, j, k};
                int nc[3] = {C->nx, C->ny, C->nz};
                int *pp[3] = {pi, pj, pk};
                for (int d = 0; d < 3; ++d) {
                    int cc = f[d] / 2;
                    int nb = (f[d] & 1) ? cc + 1 : cc - 1;
                    if (nb < 0) nb = (C->btype[2 * d] != IBMDEM_BC_PERIODIC) ? cc : nb + nc[d];
                    else if (nb >= nc[d]) nb = (C->btype[2 * d + 1] != IBMDEM_BC_PERIODIC) ? cc : nb - nc[d];
                    pp[d][0] = cc; pp[d][1] = nb;
                }
                static const double wt[2] = {0.75, 0.25};
                double s = 0.0;
                for (int c3 = 0; c3 < 2; ++c3)
                    for (int c2 = 0; c2 < 2; ++c2)
                        for (int c1 = 0; c1 < 2; ++c1) {
                            double w = wt[c3] * wt[c2] * wt[c1];
                            s = s + w * C->x[((size_t)pk[c3] * C->ny + pj[c2]) * C->nx + pi[c1]];
                        }
                size_t id = ((size_t)k * F->ny + j) * F->nx + i;
                F->x[id] = F->x[id] + s;
            }
}

static void vcycle(ibmdem_ctx *c, int l) {
    mg_level *L = &c->lev[l];
    memset(L->x, 0, L->n * sizeof(double));
    if (l == c->nlev - 1) {
        int ns = c->cs.mg_coarse_sweeps;
        if (L->rb_valid) {
            for (int s = 0; s < ns; ++s) { rb_half(L, 0); rb_half(L, 1); }
            for (int s = 0; s < ns; ++s) { rb_half(L, 1); rb_half(L, 0); }
        } else {
            for (int s = 0; s < 2 * ns; ++s) jacobi_sweep(L, 0.8);
        }
        return;
    }
    for (int s = 0; s < c->cs.mg_pre; ++s) { rb_half(L, 0); rb_half(L, 1); }
    mg_apply(L, L->x, L->r);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < L->n; ++q) L->r[q] = L->b[q] - L->r[q];
    restrict_to(L, L->r, &c->lev[l + 1]);
    vcycle(c, l + 1);
    prolong_add(&c->lev[l + 1], L);
    for (int s = 0; s < c->cs.mg_post; ++s) { rb_half(L, 1); rb_half(L, 0); }
}

static double dot(const double *a, const double *b, int n) {
    const int nc = (n + RCHUNK - 1) / RCHUNK;
    double *part = (double *)xcalloc((size_t)nc, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int ch = 0; ch < nc; ++ch) {
        double s = 0.0;
        const int hi = (ch + 1) * RCHUNK < n ? (ch + 1) * RCHUNK : n;
        for (int q = ch * RCHUNK; q < hi; ++q) s += a[q] * b[q];
        part[ch] = s;
    }
    const double s = chunk_sum(part, nc);
    free(part);
    return s;
}

static int mg_build(ibmdem_ctx *c) {
    mg_level *L = &c->lev[0];
    L->nx = c->nx; L->ny = c->ny; L->nz = c->nz;
    memcpy(L->btype, c->btype, sizeof L->btype);
    L->h[0] = c->dx; L->h[1] = c->dy; L->h[2] = c->dz;
    mg_setup_level(L);
    c->nlev = 1;
    int maxl = c->cs.mg_max_levels > 0 ? c->cs.mg_max_levels : 16;
    while (c->nlev < maxl) {
        mg_level *F = &c->lev[c->nlev - 1];
        if ((F->nx % 2) || (F->ny % 2) || (F->nz % 2)) break;
        if (F->nx < 4 || F->ny < 4 || F->nz < 4) break;
        mg_level *C = &c->lev[c->nlev];
        C->nx = F->nx / 2; C->ny = F->ny / 2; C->nz = F->nz / 2;
        memcpy(C->btype, F->btype, sizeof C->btype);
        C->h[0] = 2 * F->h[0]; C->h[1] = 2 * F->h[1]; C->h[2] = 2 * F->h[2];
        mg_setup_level(C);
        c->nlev++;
    }
    return IBMDEM_OK;
}

/* SetPressureVector + PoissonSolver + GetLocalArray (poisson.F90:107-186) */
static int poisson_solve(ibmdem_ctx *c, ibmdem_solve_info *info) {
    const int n = c->nx * c->ny * c->nz;
    double *b = c->lev[0].b;  /* reused below; keep a copy of the RHS */
    double *rhs = (double *)xcalloc(n, sizeof(double));
    int q = 0;
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) rhs[q++] = c->velDiv[ID(c, i, j, k)];
    /* pinned row 0 of the reference: b[0] = -sum_{i>0} b[i]; none with an
     * outflow side (P' = 0 there makes the operator non-singular) */
    int dirichlet = 0;
    for (int sd = 0; sd < 6; ++sd)
        if (c->btype[sd] == IBMDEM_BC_OUTFLOW) dirichlet = 1;
    /* sum_{i>0} b[i] in fixed chunks (the first chunk starts at element 1) */
    double s;
    {
        const int nc = (n + RCHUNK - 1) / RCHUNK;
        double *part = (double *)xcalloc((size_t)nc, sizeof(double));
#pragma omp parallel for schedule(static)
        for (int ch = 0; ch < nc; ++ch) {
            double a = 0.0;
            const int hi = (ch + 1) * RCHUNK < n ? (ch + 1) * RCHUNK : n;
            for (int e = ch == 0 ? 1 : ch * RCHUNK; e < hi; ++e) a += rhs[e];
            part[ch] = a;
        }
        s = chunk_sum(part, nc);
        free(part);
    }
    if (!dirichlet) rhs[0] = -s;
    (void)b;
    double *x = c->cg_x, *r = c->cg_r, *z = c->cg_z, *p = c->cg_p, *Ap = c->cg_Ap;
    memset(x, 0, n * sizeof(double));
    memcpy(r, rhs, n * sizeof(double));
    double bnorm = sqrt(dot(rhs, rhs, n));
    int it = 0, conv = 0;
    double rr = bnorm * bnorm;
    if (bnorm == 0.0) {
        conv = 1;
    } else {
        memcpy(c->lev[0].b, r, n * sizeof(double));
        vcycle(c, 0);
        memcpy(z, c->lev[0].x, n * sizeof(double));
        memcpy(p, z, n * sizeof(double));
        double rz = dot(r, z, n);
        for (it = 1; it <= c->cs.poisson_max_iter; ++it) {
            mg_apply(&c->lev[0], p, Ap);
            double pAp = dot(p, Ap, n);
            double alpha = rz / pAp;
#pragma omp parallel for schedule(static)
            for (int e = 0; e < n; ++e) { x[e] = x[e] + alpha * p[e]; r[e] = r[e] - alpha * Ap[e]; }
            rr = dot(r, r, n);
            if (!(rr == rr)) break;
            if (sqrt(rr) <= c->cs.poisson_rtol * bnorm) { conv = 1; break; }
            memcpy(c->lev[0].b, r, n * sizeof(double));
            vcycle(c, 0);
            memcpy(z, c->lev[0].x, n * sizeof(double));
            double rzn = dot(r, z, n);
            double beta = rzn / rz;
            rz = rzn;
#pragma omp parallel for schedule(static)
            for (int e = 0; e < n; ++e) p[e] = z[e] + beta * p[e];
        }
        if (it > c->cs.poisson_max_iter) it = c->cs.poisson_max_iter;
    }
    /* gauge P'(1,1,1) = 0 (none with an outflow side) and scatter into Pcorr */
    double x0 = dirichlet ? 0.0 : x[0];
    q = 0;
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) { c->Pc[ID(c, i, j, k)] = x[q] - x0; q++; }
    if (info) {
        info->iterations = it;
        info->converged = conv;
        info->rel_residual = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
        info->rhs_norm = bnorm;
    }
    free(rhs);
    if (!(rr == rr)) return set_err(c, IBMDEM_ERR_NAN, "non-finite residual in the Poisson solve");
    if (!conv) return set_err(c, IBMDEM_ERR_NOT_CONVERGED, "Poisson solve did not converge");
    return IBMDEM_OK;
}

/* ------------------------------------------------------------------------
 * Contact (ZJS_COLLISION_DEM.F90:21-148, probes omitted)
 * ------------------------------------------------------------------------ */
static void contact(ibmdem_ctx *c) {
    const double eta = c->eta, di = c->dx, knC = c->cs.kn_coeff;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->M; ++i) {
        FC(c, i, 0) = 0; FC(c, i, 1) = 0; FC(c, i, 2) = 0;
        for (int q = 0; q < c->npairs[i]; ++q) {
            int j = c->pairs[(size_t)i * c->cs.max_pairs + q];
            double r1 = XC(c, i, 0) - XC(c, j, 0);
            double r2 = XC(c, i, 1) - XC(c, j, 1);
            double r3 = XC(c, i, 2) - XC(c, j, 2);
            double dist = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            double n1 = r1 / dist, n2 = r2 / dist, n3 = r3 / dist;
            double meff = 2 * c->mass[i] * c->mass[j] / (c->mass[i] + c->mass[j]);
            double rij = c->radeff[i] + c->radeff[j];
            double deltan = rij - dist;
            double kn = knC / (0.1 * di) * meff;
            double etan = eta * sqrt(meff * kn);
            double f1 = 0, f2 = 0, f3 = 0;
            double vnij = n1 * (VC(c, i, 0) - VC(c, j, 0)) + n2 * (VC(c, i, 1) - VC(c, j, 1)) +
                          n3 * (VC(c, i, 2) - VC(c, j, 2));
            if (deltan > 0) {
                f1 = n1 * (kn * deltan - etan * vnij);
                f2 = n2 * (kn * deltan - etan * vnij);
                f3 = n3 * (kn * deltan - etan * vnij);
            }
            FC(c, i, 0) = FC(c, i, 0) + f1;
            FC(c, i, 1) = FC(c, i, 1) + f2;
            FC(c, i, 2) = FC(c, i, 2) + f3;
        }
    }
    /* CalculateSolidBoundaryCollisionForce */
    static const double nrm[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->M; ++i) {
        int ci, cj, ck;
        unpack_cell(c, c->cell[i], &ci, &cj, &ck);
        if (!(ci == 1 || ci == c->nx || cj == 1 || cj == c->ny || ck == 1 || ck == c->nz)) continue;
        double x1 = XC(c, i, 0), x2 = XC(c, i, 1), x3 = XC(c, i, 2);
        for (int s = 0; s < 6; ++s) {
            if (c->btype[s] == IBMDEM_BC_PERIODIC) continue;   /* walls, inflow and outflow screens */
            double p1, p2, p3;
            if (s % 2 == 0) { p1 = 0.0; p2 = 0.0; p3 = 0.0; }
            else { p1 = c->L[0]; p2 = c->L[1]; p3 = c->L[2]; }
            double n1 = nrm[s][0], n2 = nrm[s][1], n3 = nrm[s][2];
            double r1 = x1 - p1, r2 = x2 - p2, r3 = x3 - p3;
            double dist = r1 * n1 + r2 * n2 + r3 * n3;
            double meff = c->mass[i];
            double rij = c->radeff[i];
            double deltan = rij - dist;
            double kn = knC / (0.1 * di) * meff;
            double etan = eta * sqrt(meff * kn);
            double vnij = n1 * VC(c, i, 0) + n2 * VC(c, i, 1) + n3 * VC(c, i, 2);
            if (deltan > 0) {
                double f1 = n1 * (kn * deltan - etan * vnij);
                double f2 = n2 * (kn * deltan - etan * vnij);
                double f3 = n3 * (kn * deltan - etan * vnij);
                FC(c, i, 0) = FC(c, i, 0) + f1;
                FC(c, i, 1) = FC(c, i, 1) + f2;
                FC(c, i, 2) = FC(c, i, 2) + f3;
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * Rigid bodies (ZJS_MS_DEM.F90:494-519, couplesolverCylinder.F90:2-217,
 * ZJS_AUXILIARY_DEM.F90:20-72)
 * ------------------------------------------------------------------------ */
static void cross(const double a[3], const double b[3], double d[3]) {
    double c1 = a[1] * b[2] - a[2] * b[1];
    double c2 = -a[0] * b[2] + a[2] * b[0];
    double c3 = a[0] * b[1] - a[1] * b[0];
    d[0] = c1; d[1] = c2; d[2] = c3;
}
static void tmul(const double *A, const double *B, double *D) {
    double C[9];
    C[0] = A[0] * B[0] + A[1] * B[3] + A[2] * B[6];
    C[1] = A[0] * B[1] + A[1] * B[4] + A[2] * B[7];
    C[2] = A[0] * B[2] + A[1] * B[5] + A[2] * B[8];
    C[3] = A[3] * B[0] + A[4] * B[3] + A[5] * B[6];
    C[4] = A[3] * B[1] + A[4] * B[4] + A[5] * B[7];
    C[5] = A[3] * B[2] + A[4] * B[5] + A[5] * B[8];
    C[6] = A[6] * B[0] + A[7] * B[3] + A[8] * B[6];
    C[7] = A[6] * B[1] + A[7] * B[4] + A[8] * B[7];
    C[8] = A[6] * B[2] + A[7] * B[5] + A[8] * B[8];
    memcpy(D, C, sizeof C);
}
static void tvmul(const double *A, const double b[3], double d[3]) {
    double c1 = A[0] * b[0] + A[1] * b[1] + A[2] * b[2];
    double c2 = A[3] * b[0] + A[4] * b[1] + A[5] * b[2];
    double c3 = A[6] * b[0] + A[7] * b[1] + A[8] * b[2];
    d[0] = c1; d[1] = c2; d[2] = c3;
}
static void rotate_matrix(const double th[3], double *mat) {
    double mag = sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
    mag = mag > 1e-10 ? mag : 1e-10;
    double ux = th[0] / mag, uy = th[1] / mag, uz = th[2] / mag;
    double cm = cos(mag), sm = sin(mag);
    mat[0] = (1 - cm) * ux * ux + cm;
    mat[1] = (1 - cm) * ux * uy - sm * uz;
    mat[2] = (1 - cm) * ux * uz + sm * uy;
    mat[3] = (1 - cm) * uy * ux + sm * uz;
    mat[4] = (1 - cm) * uy * uy + cm;
    mat[5] = (1 - cm) * uy * uz - sm * ux;
    mat[6] = (1 - cm) * uz * ux - sm * uy;
    mat[7] = (1 - cm) * uz * uy + sm * ux;
    mat[8] = (1 - cm) * uz * uz + cm;
}

/* nearest periodic image of a body centre seen from position xm */
static void body_image(const ibmdem_ctx *c, int b, const double xm[3], double xb[3]) {
    for (int d = 0; d < 3; ++d) {
        xb[d] = c->bx[3 * 0 + d * c->B + b];
        if (c->btype[2 * d] == IBMDEM_BC_PERIODIC) {
            double dd = xm[d] - xb[d];
            if (dd > 0.5 * c->L[d]) xb[d] = xb[d] + c->L[d];
            else if (dd < -0.5 * c->L[d]) xb[d] = xb[d] + (-c->L[d]);
        }
    }
}

#define B3(arr, b, d) ((arr)[(size_t)(d) * c->B + (b)])

static void body_assemble(ibmdem_ctx *c) {
    for (int b = 0; b < c->B; ++b)
        for (int d = 0; d < 3; ++d) {
            B3(c->bfdrag, b, d) = 0.0; B3(c->bfcol, b, d) = 0.0;
            B3(c->bfmom, b, d) = 0.0; B3(c->bfmomcol, b, d) = 0.0;
        }
    for (int i = 0; i < c->M; ++i) {
        int b = c->body[i];
        double xm[3] = {XC(c, i, 0), XC(c, i, 1), XC(c, i, 2)}, xb[3], rpc[3], fv[3], fc[3], mt[3], mc[3];
        for (int d = 0; d < 3; ++d) {
            B3(c->bfdrag, b, d) = B3(c->bfdrag, b, d) + FD(c, i, d) * c->vol[i] * c->cs.rhof;
            B3(c->bfcol, b, d) = B3(c->bfcol, b, d) + FC(c, i, d);
        }
        body_image(c, b, xm, xb);
        for (int d = 0; d < 3; ++d) {
            rpc[d] = xm[d] - xb[d];
            fv[d] = FD(c, i, d) * c->vol[i];
            fc[d] = FC(c, i, d);
        }
        cross(rpc, fv, mt);
        cross(rpc, fc, mc);
        for (int d = 0; d < 3; ++d) {
            B3(c->bfmom, b, d) = B3(c->bfmom, b, d) + mt[d];
            B3(c->bfmomcol, b, d) = B3(c->bfmomcol, b, d) + mc[d];
        }
    }
}

static void body_hydro(ibmdem_ctx *c) {
    const double dt = c->cs.dt, rhof = c->cs.rhof;
    for (int b = 0; b < c->B; ++b) {
        double ratio = (c->brho[b] - rhof) / c->brho[b];
        for (int d = 0; d < 3; ++d) {
            B3(c->bfhydro, b, d) = (B3(c->bv, b, d) - B3(c->bvold, b, d)) / dt * (1.0 - ratio);
            B3(c->bfmomhydro, b, d) = (B3(c->bL, b, d) - B3(c->bLold, b, d)) / dt * (1.0 - ratio);
            B3(c->bvold, b, d) = B3(c->bv, b, d);
            B3(c->bLold, b, d) = B3(c->bL, b, d);
        }
    }
    body_assemble(c);
    for (int b = 0; b < c->B; ++b)
        for (int d = 0; d < 3; ++d) {
            B3(c->bfhydro, b, d) = B3(c->bfhydro, b, d) + B3(c->bfdrag, b, d) / c->bm[b];
            B3(c->bfmomhydro, b, d) = B3(c->bfmomhydro, b, d) + B3(c->bfmom, b, d);
        }
}

static void body_integrate(ibmdem_ctx *c) {
    const double dtD = c->cs.dt / c->cs.couple_interval, rhof = c->cs.rhof;
    for (int b = 0; b < c->B; ++b) {
        if (c->bfixed[b]) continue;
        double ratio = (c->brho[b] - rhof) / c->brho[b];
        double *rot = c->brot + 9 * b, *rg = c->brotg + 9 * b;
        for (int d = 0; d < 3; ++d)
            B3(c->bv, b, d) = B3(c->bv, b, d) + B3(c->bfhydro, b, d) * dtD +
                              B3(c->bfcol, b, d) / c->bm[b] * dtD + ratio * c->cs.gravity[d] * dtD;
        double th[3] = {B3(c->bom, b, 0) * dtD, B3(c->bom, b, 1) * dtD, B3(c->bom, b, 2) * dtD};
        rotate_matrix(th, rot);
        double ITInv[9];
        memcpy(ITInv, c->bI + 9 * b, sizeof ITInv);
        ITInv[0] = 1.0 / ITInv[0]; ITInv[4] = 1.0 / ITInv[4]; ITInv[8] = 1.0 / ITInv[8];
        tmul(rg, rot, rg);
        double RMInv[9] = {rg[0], rg[3], rg[6], rg[1], rg[4], rg[7], rg[2], rg[5], rg[8]};
        double T1[9], TM[9];
        tmul(rg, ITInv, T1);
        tmul(T1, RMInv, TM);
        double Lv[3];
        for (int d = 0; d < 3; ++d) {
            B3(c->bL, b, d) = B3(c->bL, b, d) + (B3(c->bfmomhydro, b, d) + B3(c->bfmomcol, b, d)) * dtD;
            Lv[d] = B3(c->bL, b, d);
        }
        double om[3];
        tvmul(TM, Lv, om);
        for (int d = 0; d < 3; ++d) B3(c->bom, b, d) = om[d];
        for (int d = 0; d < 3; ++d) B3(c->bx, b, d) = B3(c->bx, b, d) + B3(c->bv, b, d) * dtD;
    }
}

static void markers_follow(ibmdem_ctx *c) {
    const double dtD = c->cs.dt / c->cs.couple_interval;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->M; ++i) {
        int b = c->body[i];
        if (c->bfixed[b]) continue;
        double xm[3] = {XC(c, i, 0), XC(c, i, 1), XC(c, i, 2)}, xb[3];
        body_image(c, b, xm, xb);
        for (int d = 0; d < 3; ++d) xm[d] = xm[d] + B3(c->bv, b, d) * dtD;
        const double *mat = c->brot + 9 * b;
        double t[3] = {xm[0] - xb[0], xm[1] - xb[1], xm[2] - xb[2]};
        xm[0] = mat[0] * t[0] + mat[1] * t[1] + mat[2] * t[2];
        xm[1] = mat[3] * t[0] + mat[4] * t[1] + mat[5] * t[2];
        xm[2] = mat[6] * t[0] + mat[7] * t[1] + mat[8] * t[2];
        for (int d = 0; d < 3; ++d) xm[d] = xm[d] + xb[d];
        double om[3] = {B3(c->bom, b, 0), B3(c->bom, b, 1), B3(c->bom, b, 2)};
        double rr[3] = {xm[0] - xb[0], xm[1] - xb[1], xm[2] - xb[2]}, vv[3];
        cross(om, rr, vv);
        for (int d = 0; d < 3; ++d) {
            XC(c, i, d) = xm[d];
            VC(c, i, d) = vv[d] + B3(c->bv, b, d);
        }
    }
    /* AdjustParcelPosition (ZJS_EVOLUTION_DEM.F90:121-150) */
    double e[3] = {c->dx * 1e-20, c->dy * 1e-20, c->dz * 1e-20};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->M; ++i)
        for (int d = 0; d < 3; ++d) {
            if (fabs(XC(c, i, d) - 0.0) < e[d]) XC(c, i, d) = 0.0 + e[d];
            if (fabs(XC(c, i, d) - c->L[d]) < e[d]) XC(c, i, d) = c->L[d] + e[d];
        }
}

static int dem_substep(ibmdem_ctx *c) {
    contact(c);
    body_assemble(c);
    body_integrate(c);
    markers_follow(c);
    return rebuild_lists(c);
}

/* ------------------------------------------------------------------------
 * ABI
 * ------------------------------------------------------------------------ */
static double eta_of(const ibmdem_case *cs) {
    double le = log(cs->restitution);
    if (cs->eta_textbook) return -2.0 * le / sqrt(3.14159265358979323846 * 3.14159265358979323846 + le * le);
    /* 3.1415926536**2 is a default-REAL (binary32) literal squared in
     * binary32 (ZJS_COLLISION_DEM.F90:30) */
    volatile float pf = 3.1415926536f;
    volatile float pf2 = pf * pf;
    return -2 * le / sqrt((double)pf2 + le);
}

int ibmdem_create(const ibmdem_case *cs, ibmdem_ctx **out) {
    if (!cs || !out) return IBMDEM_ERR_ARG;
    *out = NULL;
    if (cs->nx < 2 || cs->ny < 2 || cs->nz < 2 || cs->dx <= 0 || cs->dy <= 0 || cs->dz <= 0 ||
        cs->dt <= 0 || cs->n_markers < 0 || cs->n_bodies < 0 || cs->couple_interval < 1)
        return IBMDEM_ERR_ARG;
    for (int s = 0; s < 6; ++s) {
        const int b = cs->btype[s];
        if (b != IBMDEM_BC_PERIODIC && b != IBMDEM_BC_WALL && b != IBMDEM_BC_INFLOW && b != IBMDEM_BC_OUTFLOW)
            return IBMDEM_ERR_UNSUPPORTED;
        if ((b == IBMDEM_BC_INFLOW && (s & 1)) || (b == IBMDEM_BC_OUTFLOW && !(s & 1))) return IBMDEM_ERR_ARG;
        if ((b == IBMDEM_BC_INFLOW || b == IBMDEM_BC_OUTFLOW) &&
            (cs->bulk_switch[0] == 1 || cs->bulk_switch[1] == 1 || cs->bulk_switch[2] == 1))
            return IBMDEM_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3; ++d)
        if ((cs->btype[2 * d] == IBMDEM_BC_PERIODIC) != (cs->btype[2 * d + 1] == IBMDEM_BC_PERIODIC))
            return IBMDEM_ERR_ARG;
    /* the same limits as the GPU library: a 5x5x5 marker neighbourhood must
     * not wrap onto itself */
    if (cs->n_markers > 0 && (cs->nx < 5 || cs->ny < 5 || cs->nz < 5)) return IBMDEM_ERR_ARG;
    if (cs->n_markers >= (1 << 23)) return IBMDEM_ERR_UNSUPPORTED;
    if (cs->time_scheme != IBMDEM_SCHEME_EULER && cs->time_scheme != IBMDEM_SCHEME_RK3) return IBMDEM_ERR_ARG;
    if (cs->stencil_tile < 0 || cs->stencil_tile > 3) return IBMDEM_ERR_ARG;   /* a schedule choice only */
    ibmdem_ctx *c = (ibmdem_ctx *)xcalloc(1, sizeof *c);
    c->cs = *cs;
    if (c->cs.n_mdf <= 0) c->cs.n_mdf = 5;
    if (c->cs.max_pairs <= 0) c->cs.max_pairs = 200;
    if (c->cs.mg_pre <= 0) c->cs.mg_pre = 2;
    if (c->cs.mg_post <= 0) c->cs.mg_post = 2;
    if (c->cs.mg_coarse_sweeps <= 0) c->cs.mg_coarse_sweeps = 24;
    if (c->cs.poisson_max_iter <= 0) c->cs.poisson_max_iter = 200;
    if (c->cs.poisson_rtol <= 0) c->cs.poisson_rtol = 1e-10;
    c->nx = cs->nx; c->ny = cs->ny; c->nz = cs->nz;
    memcpy(c->btype, cs->btype, sizeof c->btype);
    c->px = c->nx + 2 * LAP; c->py = c->ny + 2 * LAP; c->pz = c->nz + 2 * LAP;
    c->nbox = (size_t)c->px * c->py * c->pz;
    c->dx = cs->dx; c->dy = cs->dy; c->dz = cs->dz;
    c->L[0] = c->dx * c->nx; c->L[1] = c->dy * c->ny; c->L[2] = c->dz * c->nz;
    double **fields[] = {&c->u, &c->v, &c->w, &c->P, &c->Pc, &c->velDiv, &c->uold, &c->vold, &c->wold,
                         &c->fx, &c->fy, &c->fz, &c->fxo, &c->fyo, &c->fzo, &c->RHSul, &c->RHSvl,
                         &c->RHSwl, &c->Px, &c->Py, &c->Pz};
    for (size_t q = 0; q < sizeof fields / sizeof fields[0]; ++q)
        *fields[q] = (double *)xcalloc(c->nbox, sizeof(double));
    for (int d = 0; d < 3; ++d) {
        c->uin[d] = cs->btype[2 * d] == IBMDEM_BC_INFLOW ? cs->inflow[d] : 0.0;
        const size_t pn = d == 0 ? (size_t)c->ny * c->nz : d == 1 ? (size_t)c->nx * c->nz : (size_t)c->nx * c->ny;
        c->plane[d] = cs->btype[2 * d + 1] == IBMDEM_BC_OUTFLOW ? (double *)xcalloc(pn, sizeof(double)) : NULL;
    }
    fill_ghost_fluid(c);
    c->M = cs->n_markers; c->B = cs->n_bodies;
    c->cap = 8 * c->M + 64;
    c->x = (double *)xcalloc(3 * (size_t)c->cap, sizeof(double));
    c->vel = (double *)xcalloc(3 * (size_t)c->cap, sizeof(double));
    c->fdrag = (double *)xcalloc(3 * (size_t)c->cap, sizeof(double));
    c->fcol = (double *)xcalloc(3 * (size_t)c->cap, sizeof(double));
    c->vol = (double *)xcalloc(c->cap, sizeof(double));
    c->mass = (double *)xcalloc(c->cap, sizeof(double));
    c->radeff = (double *)xcalloc(c->cap, sizeof(double));
    c->body = (int *)xcalloc(c->cap, sizeof(int));
    c->ptype = (int *)xcalloc(c->cap, sizeof(int));
    c->cell = (int *)xcalloc(c->cap, sizeof(int));
    c->img = (image_t *)xcalloc(c->cap, sizeof(image_t));
    c->cell_start = (int *)xcalloc(c->nbox + 1, sizeof(int));
    c->cell_items = (int *)xcalloc(c->cap, sizeof(int));
    for (int s = 0; s < 6; ++s) c->send[s] = (int *)xcalloc(c->cap, sizeof(int));
    c->npairs = (int *)xcalloc(c->M, sizeof(int));
    c->pairs = (int *)xcalloc((size_t)c->M * c->cs.max_pairs, sizeof(int));
    size_t B = (size_t)c->B;
    double **bf[] = {&c->bx, &c->bv, &c->bvold, &c->bom, &c->bL, &c->bLold, &c->bfdrag, &c->bfmom,
                     &c->bfcol, &c->bfmomcol, &c->bfhydro, &c->bfmomhydro};
    for (size_t q = 0; q < sizeof bf / sizeof bf[0]; ++q) *bf[q] = (double *)xcalloc(3 * B, sizeof(double));
    c->brot = (double *)xcalloc(9 * B, sizeof(double));
    c->brotg = (double *)xcalloc(9 * B, sizeof(double));
    c->bI = (double *)xcalloc(9 * B, sizeof(double));
    c->bm = (double *)xcalloc(B, sizeof(double));
    c->brho = (double *)xcalloc(B, sizeof(double));
    c->bfixed = (int *)xcalloc(B, sizeof(int));
    c->eta = eta_of(&c->cs);
    mg_build(c);
    int n = c->nx * c->ny * c->nz;
    c->cg_x = (double *)xcalloc(n, sizeof(double));
    c->cg_r = (double *)xcalloc(n, sizeof(double));
    c->cg_z = (double *)xcalloc(n, sizeof(double));
    c->cg_p = (double *)xcalloc(n, sizeof(double));
    c->cg_Ap = (double *)xcalloc(n, sizeof(double));
    *out = c;
    return IBMDEM_OK;
}

int ibmdem_destroy(ibmdem_ctx *c) {
    if (!c) return IBMDEM_OK;
    double *fl[] = {c->u, c->v, c->w, c->P, c->Pc, c->velDiv, c->uold, c->vold, c->wold, c->fx, c->fy,
                    c->fz, c->fxo, c->fyo, c->fzo, c->RHSul, c->RHSvl, c->RHSwl, c->Px, c->Py, c->Pz,
                    c->x, c->vel, c->fdrag, c->fcol, c->vol, c->mass, c->radeff, c->bx, c->bv, c->bvold,
                    c->bom, c->bL, c->bLold, c->bfdrag, c->bfmom, c->bfcol, c->bfmomcol, c->bfhydro,
                    c->bfmomhydro, c->brot, c->brotg, c->bI, c->bm, c->brho, c->cg_x, c->cg_r, c->cg_z,
                    c->cg_p, c->cg_Ap};
    for (size_t q = 0; q < sizeof fl / sizeof fl[0]; ++q) free(fl[q]);
    for (int d = 0; d < 3; ++d) free(c->plane[d]);
    free(c->body); free(c->ptype); free(c->cell); free(c->img); free(c->cell_start);
    free(c->cell_items); free(c->npairs); free(c->pairs); free(c->bfixed);
    free(c->probe_list); free(c->probe_ring);
    for (int s = 0; s < 6; ++s) free(c->send[s]);
    for (int l = 0; l < c->nlev; ++l) {
        free(c->lev[l].inv_diag); free(c->lev[l].diag); free(c->lev[l].x); free(c->lev[l].b); free(c->lev[l].r);
    }
    free(c);
    return IBMDEM_OK;
}

const char *ibmdem_last_error(const ibmdem_ctx *c) { return c ? c->err : "null context"; }
int ibmdem_sync(ibmdem_ctx *c) { (void)c; return IBMDEM_OK; }

static void copy_in(ibmdem_ctx *c, double *dst, const double *src) {
    size_t q = 0;
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) dst[ID(c, i, j, k)] = src[q++];
}
static void copy_out(const ibmdem_ctx *c, double *dst, const double *src) {
    size_t q = 0;
    for (int k = 1; k <= c->nz; ++k)
        for (int j = 1; j <= c->ny; ++j)
            for (int i = 1; i <= c->nx; ++i) dst[q++] = src[ID(c, i, j, k)];
}

int ibmdem_upload_fields(ibmdem_ctx *c, const double *u, const double *v, const double *w, const double *p) {
    if (!c) return IBMDEM_ERR_ARG;
    if (u) copy_in(c, c->u, u);
    if (v) copy_in(c, c->v, v);
    if (w) copy_in(c, c->w, w);
    if (p) copy_in(c, c->P, p);
    /* an uploaded normal component resets its outflow plane */
    const double *up[3] = {u, v, w};
    for (int d = 0; d < 3; ++d)
        if (up[d] && c->btype[2 * d + 1] == IBMDEM_BC_OUTFLOW) plane_from_interior(c, d);
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
static size_t outflow_size(const ibmdem_ctx *c, int side) {
    if (side != 1 && side != 3 && side != 5) return 0;
    if (c->btype[side] != IBMDEM_BC_OUTFLOW) return 0;
    const int d = side / 2;
    return d == 0 ? (size_t)c->ny * c->nz : d == 1 ? (size_t)c->nx * c->nz : (size_t)c->nx * c->ny;
}
int ibmdem_download_outflow(ibmdem_ctx *c, int32_t side, double *plane) {
    if (!c) return IBMDEM_ERR_ARG;
    const size_t n = outflow_size(c, side);
    if (!n || !plane) return set_err(c, IBMDEM_ERR_ARG, "not an outflow side (1, 3 or 5 of type OUTFLOW)");
    memcpy(plane, c->plane[side / 2], n * sizeof(double));
    return IBMDEM_OK;
}
int ibmdem_upload_outflow(ibmdem_ctx *c, int32_t side, const double *plane) {
    if (!c) return IBMDEM_ERR_ARG;
    const size_t n = outflow_size(c, side);
    if (!n || !plane) return set_err(c, IBMDEM_ERR_ARG, "not an outflow side (1, 3 or 5 of type OUTFLOW)");
    memcpy(c->plane[side / 2], plane, n * sizeof(double));
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_download_fields(ibmdem_ctx *c, double *u, double *v, double *w, double *p) {
    if (!c) return IBMDEM_ERR_ARG;
    if (u) copy_out(c, u, c->u);
    if (v) copy_out(c, v, c->v);
    if (w) copy_out(c, w, c->w);
    if (p) copy_out(c, p, c->P);
    return IBMDEM_OK;
}
int ibmdem_download_force(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    if (!c) return IBMDEM_ERR_ARG;
    if (fx) copy_out(c, fx, c->fx);
    if (fy) copy_out(c, fy, c->fy);
    if (fz) copy_out(c, fz, c->fz);
    return IBMDEM_OK;
}
/* the oracle allocates nothing per call either way */
int ibmdem_download_force_noalloc(ibmdem_ctx *c, double *fx, double *fy, double *fz) {
    return ibmdem_download_force(c, fx, fy, fz);
}

int ibmdem_download_poisson(ibmdem_ctx *c, double *pcorr, double *rhs) {
    if (!c) return IBMDEM_ERR_ARG;
    if (pcorr) copy_out(c, pcorr, c->Pc);
    if (rhs) copy_out(c, rhs, c->velDiv);
    return IBMDEM_OK;
}
int ibmdem_upload_poisson_rhs(ibmdem_ctx *c, const double *rhs) {
    if (!c || !rhs) return IBMDEM_ERR_ARG;
    copy_in(c, c->velDiv, rhs);
    return IBMDEM_OK;
}

int ibmdem_upload_markers(ibmdem_ctx *c, const ibmdem_markers *m) {
    if (!c || !m || !m->x || !m->v || !m->vol || !m->mass || !m->radeff || !m->body) return IBMDEM_ERR_ARG;
    for (int i = 0; i < c->M; ++i) {
        for (int d = 0; d < 3; ++d) {
            XC(c, i, d) = m->x[(size_t)d * c->M + i];
            VC(c, i, d) = m->v[(size_t)d * c->M + i];
            FD(c, i, d) = m->fdrag ? m->fdrag[(size_t)d * c->M + i] : 0.0;
            FC(c, i, d) = m->fcol ? m->fcol[(size_t)d * c->M + i] : 0.0;
        }
        c->vol[i] = m->vol[i]; c->mass[i] = m->mass[i]; c->radeff[i] = m->radeff[i];
        c->body[i] = m->body[i];
        if (m->body[i] < 0 || m->body[i] >= c->B || (i > 0 && m->body[i] < m->body[i - 1]))
            return set_err(c, IBMDEM_ERR_ARG, "markers must be body-contiguous, bodies ascending");
    }
    return rebuild_lists(c);
}
int ibmdem_download_markers(ibmdem_ctx *c, ibmdem_markers *m) {
    if (!c || !m) return IBMDEM_ERR_ARG;
    for (int i = 0; i < c->M; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (m->x) m->x[(size_t)d * c->M + i] = XC(c, i, d);
            if (m->v) m->v[(size_t)d * c->M + i] = VC(c, i, d);
            if (m->fdrag) m->fdrag[(size_t)d * c->M + i] = FD(c, i, d);
            if (m->fcol) m->fcol[(size_t)d * c->M + i] = FC(c, i, d);
        }
        if (m->vol) m->vol[i] = c->vol[i];
        if (m->mass) m->mass[i] = c->mass[i];
        if (m->radeff) m->radeff[i] = c->radeff[i];
        if (m->body) m->body[i] = c->body[i];
    }
    return IBMDEM_OK;
}

int ibmdem_upload_bodies(ibmdem_ctx *c, const ibmdem_bodies *b) {
    if (!c || !b) return IBMDEM_ERR_ARG;
    size_t B3n = 3 * (size_t)c->B, B9 = 9 * (size_t)c->B;
#define CP_IN(dst, src, n) do { if (src) memcpy(dst, src, (n) * sizeof(*(dst))); } while (0)
    CP_IN(c->bx, b->x, B3n); CP_IN(c->bv, b->v, B3n); CP_IN(c->bvold, b->vold, B3n);
    CP_IN(c->bom, b->omega, B3n); CP_IN(c->bL, b->angmom, B3n); CP_IN(c->bLold, b->angmom_old, B3n);
    CP_IN(c->brot, b->rot, B9); CP_IN(c->brotg, b->rot_glo, B9); CP_IN(c->bI, b->inertia, B9);
    CP_IN(c->bm, b->mass, (size_t)c->B); CP_IN(c->brho, b->rho, (size_t)c->B);
    CP_IN(c->bfixed, b->fixed, (size_t)c->B);
    CP_IN(c->bfhydro, b->fhydro, B3n); CP_IN(c->bfmomhydro, b->fmomhydro, B3n);
    return IBMDEM_OK;
}
int ibmdem_download_bodies(ibmdem_ctx *c, ibmdem_bodies *b) {
    if (!c || !b) return IBMDEM_ERR_ARG;
    size_t B3n = 3 * (size_t)c->B, B9 = 9 * (size_t)c->B;
#define CP_OUT(dst, src, n) do { if (dst) memcpy(dst, src, (n) * sizeof(*(dst))); } while (0)
    CP_OUT(b->x, c->bx, B3n); CP_OUT(b->v, c->bv, B3n); CP_OUT(b->vold, c->bvold, B3n);
    CP_OUT(b->omega, c->bom, B3n); CP_OUT(b->angmom, c->bL, B3n); CP_OUT(b->angmom_old, c->bLold, B3n);
    CP_OUT(b->rot, c->brot, B9); CP_OUT(b->rot_glo, c->brotg, B9); CP_OUT(b->inertia, c->bI, B9);
    CP_OUT(b->mass, c->bm, (size_t)c->B); CP_OUT(b->rho, c->brho, (size_t)c->B);
    CP_OUT(b->fixed, c->bfixed, (size_t)c->B);
    CP_OUT(b->fdrag, c->bfdrag, B3n); CP_OUT(b->fmom, c->bfmom, B3n); CP_OUT(b->fcol, c->bfcol, B3n);
    CP_OUT(b->fmomcol, c->bfmomcol, B3n); CP_OUT(b->fhydro, c->bfhydro, B3n);
    CP_OUT(b->fmomhydro, c->bfmomhydro, B3n);
    return IBMDEM_OK;
}

int ibmdem_begin_step(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    memcpy(c->uold, c->u, c->nbox * sizeof(double));
    memcpy(c->vold, c->v, c->nbox * sizeof(double));
    memcpy(c->wold, c->w, c->nbox * sizeof(double));
    for (int i = 0; i < c->M; ++i) { FD(c, i, 0) = 0; FD(c, i, 1) = 0; FD(c, i, 2) = 0; }
    memset(c->fx, 0, c->nbox * sizeof(double));
    memset(c->fy, 0, c->nbox * sizeof(double));
    memset(c->fz, 0, c->nbox * sizeof(double));
    return IBMDEM_OK;
}
int ibmdem_ibm_interpolate(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    ibm_interpolate(c);
    return IBMDEM_OK;
}
int ibmdem_ibm_spread(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    refresh_images(c);
    ibm_spread(c);
    velocity_post_guess(c);
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_fluid_predictor(ibmdem_ctx *c, double a, double b, double cc) {
    if (!c) return IBMDEM_ERR_ARG;
    velocity_guess(c, a, b, cc);
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_bulk_forcing(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    modify_bulk(c);
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_divergence(ibmdem_ctx *c, double cc) {
    if (!c) return IBMDEM_ERR_ARG;
    velocity_divergence(c, cc);
    return IBMDEM_OK;
}
int ibmdem_poisson_solve(ibmdem_ctx *c, ibmdem_solve_info *info) {
    if (!c) return IBMDEM_ERR_ARG;
    int rc = poisson_solve(c, info);
    fill_ghost_fluid(c);
    return rc;
}
int ibmdem_project(ibmdem_ctx *c, double cc) {
    if (!c) return IBMDEM_ERR_ARG;
    pressure_correction(c);
    fill_ghost_fluid(c);
    velocity_correction(c, cc);
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_apply_bc(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    fill_ghost_fluid(c);
    return IBMDEM_OK;
}
int ibmdem_body_hydro(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; body_hydro(c); return IBMDEM_OK; }
int ibmdem_dem_contact(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; contact(c); return IBMDEM_OK; }
int ibmdem_body_assemble(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; body_assemble(c); return IBMDEM_OK; }
int ibmdem_body_integrate(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; body_integrate(c); return IBMDEM_OK; }
int ibmdem_markers_follow(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; markers_follow(c); return IBMDEM_OK; }
int ibmdem_dem_rebuild_lists(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; return rebuild_lists(c); }
int ibmdem_dem_substep(ibmdem_ctx *c) { if (!c) return IBMDEM_ERR_ARG; return dem_substep(c); }
int ibmdem_cloud_evolution(ibmdem_ctx *c) {
    if (!c) return IBMDEM_ERR_ARG;
    body_hydro(c);
    for (int s = 0; s < c->cs.couple_interval; ++s) {
        int rc = dem_substep(c);
        if (rc) return rc;
    }
    return IBMDEM_OK;
}

/* ThirdOrderLowStorageRungeKutta (fluid.F90:329-366): aRK, bRK as the
 * reference sets them, cRK = aRK + bRK; fluid only. Each stage advances from
 * the previous stage's velocity (uold <- u before every VelocityGuess), which
 * is Wray's low-storage scheme; the reference routine never sets uold, so as
 * written its stages would all start from a stale velocity. */
static int step_rk3(ibmdem_ctx *c, ibmdem_solve_info *info) {
    double aRK[3], bRK[3], cRK[3];
    aRK[0] = 32.0 / 60.0; bRK[0] = 0.0;
    aRK[1] = 25.0 / 60.0; bRK[1] = -17.0 / 60.0;
    aRK[2] = 45.0 / 60.0; bRK[2] = -25.0 / 60.0;
    for (int s = 0; s < 3; ++s) cRK[s] = aRK[s] + bRK[s];
    for (int s = 0; s < 3; ++s) {
        memcpy(c->uold, c->u, c->nbox * sizeof(double));
        memcpy(c->vold, c->v, c->nbox * sizeof(double));
        memcpy(c->wold, c->w, c->nbox * sizeof(double));
        ibmdem_fluid_predictor(c, aRK[s], bRK[s], cRK[s]);
        ibmdem_divergence(c, cRK[s]);
        int rc = ibmdem_poisson_solve(c, info);
        if (rc) return rc;
        ibmdem_project(c, cRK[s]);
    }
    return IBMDEM_OK;
}

/* per-body probe record of the step just completed (x, v, omega, fdrag,
 * fmom, fcol of each listed body; PostProcessing, fluid.F90:368-404) */
static void probe_append(ibmdem_ctx *c) {
    if (c->probe_n == 0) return;
    const long long slot = c->probe_count % c->probe_cap;
    const double *src[6] = {c->bx, c->bv, c->bom, c->bfdrag, c->bfmom, c->bfcol};
    for (int i = 0; i < c->probe_n; ++i)
        for (int v = 0; v < 18; ++v)
            c->probe_ring[(slot * c->probe_n + i) * 18 + v] = src[v / 3][(size_t)(v % 3) * c->B + c->probe_list[i]];
    c->probe_count++;
}

int ibmdem_probe_bodies(ibmdem_ctx *c, int32_t n, const int32_t *bodies, int32_t capacity) {
    if (!c || n < 0 || (n > 0 && (!bodies || capacity < 1))) return IBMDEM_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (bodies[i] < 0 || bodies[i] >= c->B) return set_err(c, IBMDEM_ERR_ARG, "probed body index out of range");
    free(c->probe_list); free(c->probe_ring);
    c->probe_list = NULL; c->probe_ring = NULL;
    c->probe_n = 0; c->probe_count = 0;
    if (n == 0) return IBMDEM_OK;
    c->probe_list = (int *)xcalloc((size_t)n, sizeof(int));
    c->probe_ring = (double *)xcalloc((size_t)n * capacity * 18, sizeof(double));
    memcpy(c->probe_list, bodies, (size_t)n * sizeof(int));
    c->probe_n = n; c->probe_cap = capacity;
    return IBMDEM_OK;
}

int ibmdem_probe_read(ibmdem_ctx *c, int32_t *n_steps, int64_t *first_step, double *records) {
    if (!c || !n_steps || !first_step) return IBMDEM_ERR_ARG;
    *n_steps = 0; *first_step = 0;
    if (c->probe_n == 0) return IBMDEM_OK;
    const long long kept = c->probe_count < c->probe_cap ? c->probe_count : c->probe_cap;
    *n_steps = (int32_t)kept;
    *first_step = c->probe_count - kept;
    if (!records) return IBMDEM_OK;
    const size_t rec = (size_t)c->probe_n * 18;
    for (long long q = 0; q < kept; ++q) {
        const long long slot = (c->probe_count - kept + q) % c->probe_cap;
        memcpy(records + q * rec, c->probe_ring + slot * rec, rec * sizeof(double));
    }
    return IBMDEM_OK;
}

int ibmdem_step(ibmdem_ctx *c, ibmdem_solve_info *info) {
    if (!c) return IBMDEM_ERR_ARG;
    if (c->cs.time_scheme == IBMDEM_SCHEME_RK3) return step_rk3(c, info);
    int rc;
    ibmdem_begin_step(c);
    for (int it = 0; it < c->cs.n_mdf; ++it) {
        ibmdem_ibm_interpolate(c);
        ibmdem_ibm_spread(c);
    }
    ibmdem_fluid_predictor(c, 1.0, 0.0, 1.0);
    ibmdem_bulk_forcing(c);
    ibmdem_divergence(c, 1.0);
    rc = ibmdem_poisson_solve(c, info);
    if (rc) return rc;
    ibmdem_project(c, 1.0);
    rc = ibmdem_cloud_evolution(c);
    if (rc) return rc;
    probe_append(c);
    return IBMDEM_OK;
}
int ibmdem_run(ibmdem_ctx *c, int32_t n, ibmdem_solve_info *info) {
    for (int s = 0; s < n; ++s) {
        int rc = ibmdem_step(c, info);
        if (rc) return rc;
    }
    return IBMDEM_OK;
}

int ibmdem_get_cells(ibmdem_ctx *c, int32_t *ijk) {
    if (!c || !ijk) return IBMDEM_ERR_ARG;
    for (int p = 0; p < c->M; ++p) {
        int i, j, k;
        unpack_cell(c, c->cell[p], &i, &j, &k);
        ijk[p] = i; ijk[c->M + p] = j; ijk[2 * (size_t)c->M + p] = k;
    }
    return IBMDEM_OK;
}

/* root source and accumulated periodic shift of a parcel */
static void root_of(const ibmdem_ctx *c, int p, int *root, int sh[3]) {
    sh[0] = sh[1] = sh[2] = 0;
    while (p >= c->M) {
        const image_t *im = &c->img[p];
        if (!im->mirror) sh[im->dim] += im->shift > 0 ? 1 : -1;
        p = im->src;
    }
    *root = p;
}

int ibmdem_get_pairs(ibmdem_ctx *c, int32_t cap, int32_t *count, int32_t *partner, int32_t *shift) {
    if (!c || !count) return IBMDEM_ERR_ARG;
    int rc = IBMDEM_OK;
    for (int i = 0; i < c->M; ++i) {
        count[i] = c->npairs[i];
        for (int q = 0; q < c->npairs[i]; ++q) {
            if (q >= cap) { rc = IBMDEM_ERR_CAPACITY; break; }
            int root, sh[3];
            root_of(c, c->pairs[(size_t)i * c->cs.max_pairs + q], &root, sh);
            if (partner) partner[(size_t)i * cap + q] = root;
            if (shift) shift[(size_t)i * cap + q] = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
        }
    }
    return rc;
}

int ibmdem_profile_kernel(ibmdem_ctx *c, int32_t k) { (void)c; (void)k; return IBMDEM_ERR_UNSUPPORTED; }
int ibmdem_profile_read(ibmdem_ctx *c, double *t, int64_t *n, double *b) {
    (void)c; (void)t; (void)n; (void)b; return IBMDEM_ERR_UNSUPPORTED;
}
int ibmdem_memory_bytes(const ibmdem_ctx *c, int64_t *bytes) {
    if (!c || !bytes) return IBMDEM_ERR_ARG;
    *bytes = 0;
    return IBMDEM_OK;
}

/* ------------------------------------------------------------------------
 * Multigrid cycle parameters and the batched form. The oracle runs a batch's
 * cases one after another; there is nothing to tune on one core, so the
 * autotuner keeps every case's cycle.
 * ------------------------------------------------------------------------ */
/* the predictor schedule is a GPU launch choice: the oracle only keeps the
 * value */
int ibmdem_get_stencil_tile(ibmdem_ctx *c, int32_t *tile) {
    if (!c || !tile) return IBMDEM_ERR_ARG;
    *tile = c->cs.stencil_tile;
    return IBMDEM_OK;
}
int ibmdem_set_stencil_tile(ibmdem_ctx *c, int32_t tile) {
    if (!c) return IBMDEM_ERR_ARG;
    if (tile < 0 || tile > 3) return set_err(c, IBMDEM_ERR_ARG, "stencil_tile must be 0..3");
    c->cs.stencil_tile = tile;
    return IBMDEM_OK;
}

int ibmdem_get_mg_params(ibmdem_ctx *c, int32_t *mg) {
    if (!c || !mg) return IBMDEM_ERR_ARG;
    mg[0] = c->cs.mg_pre;
    mg[1] = c->cs.mg_post;
    mg[2] = c->cs.mg_coarse_sweeps;
    return IBMDEM_OK;
}

int ibmdem_set_mg_params(ibmdem_ctx *c, const int32_t *mg) {
    if (!c || !mg) return IBMDEM_ERR_ARG;
    if (mg[0] < 1 || mg[0] > 8 || mg[1] < 1 || mg[1] > 8 || mg[2] < 1 || mg[2] > 256) return IBMDEM_ERR_ARG;
    c->cs.mg_pre = mg[0];
    c->cs.mg_post = mg[1];
    c->cs.mg_coarse_sweeps = mg[2];
    return IBMDEM_OK;
}

struct ibmdem_batch {
    int n;
    ibmdem_ctx **c;
    int split;   /* kept for ibmdem_batch_get_solve_groups */
};

int ibmdem_batch_destroy(ibmdem_batch *b) {
    if (!b) return IBMDEM_OK;
    for (int q = 0; q < b->n; ++q) ibmdem_destroy(b->c[q]);
    free(b->c);
    free(b);
    return IBMDEM_OK;
}

int ibmdem_batch_create(int32_t n, const ibmdem_case *cases, ibmdem_batch **out) {
    if (!out || !cases || n < 1) return IBMDEM_ERR_ARG;
    *out = NULL;
    ibmdem_batch *b = (ibmdem_batch *)xcalloc(1, sizeof *b);
    b->n = n;
    b->c = (ibmdem_ctx **)xcalloc((size_t)n, sizeof *b->c);
    for (int q = 0; q < n; ++q) {
        int rc = ibmdem_create(&cases[q], &b->c[q]);
        if (rc) {
            ibmdem_batch_destroy(b);
            return rc;
        }
    }
    *out = b;
    return IBMDEM_OK;
}

int ibmdem_batch_context(ibmdem_batch *b, int32_t q, ibmdem_ctx **ctx) {
    if (!b || !ctx || q < 0 || q >= b->n) return IBMDEM_ERR_ARG;
    *ctx = b->c[q];
    return IBMDEM_OK;
}

const char *ibmdem_batch_last_error(const ibmdem_batch *b) { return b ? "" : "no batch"; }

/* Step by step over the cases: a failing case stops, the others complete
 * the step in progress, then the first failure (lowest case) is returned. */
int ibmdem_batch_run(ibmdem_batch *b, int32_t n_steps, ibmdem_solve_info *info) {
    if (!b || n_steps < 0) return IBMDEM_ERR_ARG;
    int first = IBMDEM_OK;
    for (int s = 0; s < n_steps && first == IBMDEM_OK; ++s)
        for (int q = 0; q < b->n; ++q) {
            int rc = ibmdem_step(b->c[q], info ? &info[q] : NULL);
            if (rc && first == IBMDEM_OK) first = rc;
        }
    return first;
}

int ibmdem_batch_autotune(ibmdem_batch *b, int32_t trial_steps) {
    (void)trial_steps;
    return b ? IBMDEM_OK : IBMDEM_ERR_ARG;
}

/* the oracle has no batched solve: one case after another either way */
int ibmdem_batch_set_fused(ibmdem_batch *b, int32_t mode) {
    if (!b || mode < 0 || mode > 4) return IBMDEM_ERR_ARG;
    return IBMDEM_OK;
}
int ibmdem_batch_set_solve_groups(ibmdem_batch *b, int32_t groups) {
    if (!b || groups < 1 || groups > 64) return IBMDEM_ERR_ARG;
    b->split = groups;
    return IBMDEM_OK;
}
int ibmdem_batch_get_solve_groups(const ibmdem_batch *b, int32_t *groups) {
    if (!b || !groups) return IBMDEM_ERR_ARG;
    *groups = b->split > 0 ? b->split : 1;
    return IBMDEM_OK;
}
/* no batched solves to profile */
int ibmdem_batch_profile(ibmdem_batch *b, int32_t on) {
    (void)on;
    return b ? IBMDEM_OK : IBMDEM_ERR_ARG;
}
int ibmdem_batch_profile_read(const ibmdem_batch *b, int32_t kind, double *total_ms, int64_t *launches,
                              double *bytes) {
    if (!b || kind < 1 || kind > 8 || !total_ms || !launches || !bytes) return IBMDEM_ERR_ARG;
    *total_ms = 0.0;
    *launches = 0;
    *bytes = 0.0;
    return IBMDEM_OK;
}
int ibmdem_batch_stats(const ibmdem_batch *b, int64_t *fused_solves, double *ms) {
    if (!b) return IBMDEM_ERR_ARG;
    if (fused_solves) *fused_solves = 0;
    if (ms)
        for (int i = 0; i < 5; ++i) ms[i] = 0.0;
    return IBMDEM_OK;
}

/* ------------------------------------------------------------------------
 * Test hooks (oracle only): the scalar building blocks the closed-form tests
 * pin directly.
 * ------------------------------------------------------------------------ */
double ibmdem_oracle_phi(double r) { return phi1(r); }
/* OpenMP threads of the following calls (the CPU baseline times the oracle at
 * several counts in one process); returns the count in effect */
int ibmdem_oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
double ibmdem_oracle_eta(double restitution, int textbook) {
    ibmdem_case cs;
    memset(&cs, 0, sizeof cs);
    cs.restitution = restitution;
    cs.eta_textbook = textbook;
    return eta_of(&cs);
}
void ibmdem_oracle_rotation(const double theta[3], double mat[9]) { rotate_matrix(theta, mat); }
