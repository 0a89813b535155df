"""Seeded case generator: grids, spheres, markers and bodies.

Lattice units (h = 1, rho_f = 1 unless stated). Every array is produced on
the host from a named seed, so the GPU library and the CPU oracle receive
byte-identical initial states.

Formulas follow the reference's set-up code (ZJS_INIT_DEM.F90):
  * markers sit on a sphere retracted by 0.3h: rs = R - 0.3h (:250);
  * marker volume vol = pi_f*h/3/N*(12 rs^2 + h^2) with pi_f the binary32
    literal 3.1415926536 the reference multiplies by (:312);
  * marker contact radius radeff = 0.75h and collisionDist = (0.75h)*2
    (:311, :413); marker mass = rho_p * vol (:314);
  * body radius R, volume 4/3 pi R^3, mass rho_p V, inertia 0.4 m R^2 on the
    diagonal of the body-frame tensor (:327-331);
  * knCoeff = 4e4 and e = 1e-4 are the reference defaults (:73-74).
The Fibonacci lattice replaces the reference's missing SPHERE_LP600.dat.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .abi import (BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, BODY_MAT, FP64, MIXED, BodyArrays,
                  Case, MarkerArrays)

PI_F32 = float(np.float32(3.1415926536))
SEED_BASE = 20260000

# The reference's contact stiffness is in SI units at its own resolution:
# dt = 5e-5 s (initialization.F90:259), dtDEM = dt/10 (ZJS_INIT_DEM.F90:76),
# dx = 9e-3 m / 240 (main.F90:20). kn = knCoeff/(0.1 di) * meff
# (ZJS_COLLISION_DEM.F90:46), so kn dtDEM^2 / meff = knCoeff dtDEM^2/(0.1 di)
# is the dimensionless stiffness that sets the explicit integrator's
# stability. lattice_kn keeps it equal at the lattice case's dtDEM and h.
REF_DT_DEM = 5e-5 / 10
REF_DX = 9e-3 / 240


def lattice_kn(kn_ref: float, dt_dem: float, h: float = 1.0) -> float:
    """knCoeff in lattice units with the reference's kn dtDEM^2 / meff."""
    return kn_ref * (REF_DT_DEM / dt_dem) ** 2 * (h / REF_DX)


def fibonacci_sphere(n: int) -> np.ndarray:
    """n points on the unit sphere, (n, 3), golden-angle spiral."""
    k = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = k * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def markers_for_diameter(d: float, h: float = 1.0) -> int:
    """About pi D^2 / h^2 markers (one per h^2 of surface)."""
    return int(round(math.pi * d * d / (h * h)))


def stable_dt(h: float, nu: float, u_max: float) -> float:
    """Forward-Euler/central-difference (FTCS) limits: CFL 0.5, diffusion
    nu dt/h^2 <= 0.9/6, advective dt <= 0.9 * 2 nu / U^2."""
    lims = [0.9 * h * h / (6.0 * nu)]
    if u_max > 0:
        lims += [0.5 * h / u_max, 0.9 * 2.0 * nu / (u_max * u_max)]
    return min(lims)


@dataclass
class Sphere:
    centre: tuple
    diameter: float
    rho: float
    n_markers: int = 0
    fixed: bool = False


@dataclass
class Setup:
    name: str
    case: Case
    markers: MarkerArrays
    bodies: BodyArrays
    fields: tuple          # u, v, w, p  (nz, ny, nx)
    meta: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return int(self.case.nx) * int(self.case.ny) * int(self.case.nz)


def build_case(nx, ny, nz, btype, *, h=1.0, visc=0.04, rhof=1.0, dt=None,
               gravity=(0.0, 0.0, 0.0), kn_coeff=4e4, restitution=1e-4,
               bulk_switch=(0, 0, 0), bulk_target=(0.0, 0.0, 0.0),
               precision=FP64, rtol=1e-10, max_iter=200, n_mdf=5,
               couple_interval=10, u_max=0.2, n_markers=0, n_bodies=0,
               max_pairs=200, mg_pre=2, mg_post=2, mg_coarse=24, time_scheme=0,
               inflow=(0.0, 0.0, 0.0)) -> Case:
    c = Case()
    c.nx, c.ny, c.nz = nx, ny, nz
    for s in range(6):
        c.btype[s] = int(btype[s])
    c.dx = c.dy = c.dz = float(h)
    c.visc, c.rhof = float(visc), float(rhof)
    c.dt = float(dt if dt is not None else stable_dt(h, visc, u_max))
    for d in range(3):
        c.cp[d] = 0.0
        c.bulk_switch[d] = int(bulk_switch[d])
        c.bulk_target[d] = float(bulk_target[d])
        c.gravity[d] = float(gravity[d])
    c.n_mdf = n_mdf
    c.couple_interval = couple_interval
    # kn_coeff is given in the reference's units (4e4 is its default)
    c.kn_coeff = float(lattice_kn(kn_coeff, c.dt / couple_interval, h))
    c.restitution = float(restitution)
    c.collision_dist = (h * 0.75) * 2
    c.collision_buffer = 0.1
    c.eta_textbook = 0
    c.n_markers = n_markers
    c.n_bodies = n_bodies
    c.max_pairs = max_pairs
    c.precision = precision
    c.poisson_rtol = rtol
    c.poisson_max_iter = max_iter
    c.mg_pre, c.mg_post, c.mg_coarse_sweeps = mg_pre, mg_post, mg_coarse
    c.mg_max_levels = 0
    c.stencil_tile = 0
    c.time_scheme = time_scheme
    for d in range(3):
        c.inflow[d] = float(inflow[d])
    return c


def place_spheres(case: Case, spheres: list[Sphere], h: float = 1.0):
    """Markers (body-contiguous) and bodies for a list of spheres."""
    nb = len(spheres)
    counts = [s.n_markers or markers_for_diameter(s.diameter, h) for s in spheres]
    m = int(sum(counts))
    mk = MarkerArrays.empty(m)
    bd = BodyArrays.empty(nb)
    L = np.array([case.nx * case.dx, case.ny * case.dy, case.nz * case.dz])
    periodic = [case.btype[2 * d] == BC_PERIODIC for d in range(3)]
    o = 0
    for b, (s, n) in enumerate(zip(spheres, counts)):
        R = 0.5 * s.diameter
        rs = R - 0.3 * h
        pts = fibonacci_sphere(n) * rs + np.asarray(s.centre, dtype=np.float64)
        for d in range(3):
            if periodic[d]:
                pts[:, d] = np.mod(pts[:, d], L[d])
        vol = PI_F32 * h / 3 / n * (12 * rs ** 2 + h ** 2)
        mk.x[:, o:o + n] = pts.T
        mk.vol[o:o + n] = vol
        mk.mass[o:o + n] = s.rho * vol
        mk.radeff[o:o + n] = h * 0.75
        mk.body[o:o + n] = b
        bvol = 4.0 / 3.0 * math.pi * R ** 3
        mass = bvol * s.rho
        bd.x[:, b] = s.centre
        bd.mass[b] = mass
        bd.rho[b] = s.rho
        inert = 0.4 * mass * R * R
        bd.inertia[b] = [inert, 0, 0, 0, inert, 0, 0, 0, inert]
        bd.rot[b] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        bd.rot_glo[b] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        bd.fixed[b] = 1 if s.fixed else 0
        o += n
    case.n_markers = m
    case.n_bodies = nb
    return mk, bd


def rsa_centres(rng, n, diameter, box, min_gap, periodic, margin=0.0,
                max_tries=200000):
    """Random sequential addition of n equal spheres (seeded)."""
    box = np.asarray(box, dtype=np.float64)
    lo = np.array([0.0 if p else margin for p in periodic])
    hi = np.array([b if p else b - margin for b, p in zip(box, periodic)])
    dmin = diameter + min_gap
    out = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"RSA placed only {k} of {n} spheres")
        c = lo + rng.random(3) * (hi - lo)
        if k:
            dd = out[:k] - c
            for d in range(3):
                if periodic[d]:
                    dd[:, d] -= box[d] * np.round(dd[:, d] / box[d])
            if np.min(np.einsum("ij,ij->i", dd, dd)) < dmin * dmin:
                continue
        out[k] = c
        k += 1
    return out


def zero_fields(case: Case):
    shp = (case.nz, case.ny, case.nx)
    return tuple(np.zeros(shp) for _ in range(4))


# ---------------------------------------------------------------------------
# The benchmark configurations
# ---------------------------------------------------------------------------
def config_settling(precision=FP64, rtol=1e-10) -> Setup:
    """Plain FP64 single sphere settling in a closed box: 64x64x256, D=16h,
    800 markers, rho_p/rho_f = 1.5, Re_g = 40 (nu = 0.04, g = 1.25e-3),
    six no-slip walls."""
    nu = 0.04
    D, rr = 16.0, 1.5
    g = (0.1 ** 2) / ((rr - 1.0) * D)   # U_ref = sqrt((rr-1) g D) = 0.1
    case = build_case(64, 64, 256, [BC_WALL] * 6, visc=nu,
                      gravity=(0.0, 0.0, -g), u_max=0.2, precision=precision,
                      rtol=rtol)
    mk, bd = place_spheres(case, [Sphere((32.0, 32.0, 192.0), D, rr, 800)])
    return Setup("sphere-settling-64x64x256", case, mk, bd, zero_fields(case),
                 {"steps": 1000, "re_g": 40.0, "g": g})


def config_suspension(n_spheres=2000, n=256, precision=FP64, rtol=1e-10,
                      seed=SEED_BASE + 2) -> Setup:
    """Plain FP64 triply periodic suspension: n^3 grid, n_spheres spheres of
    D = 12h (452 markers) placed by seeded RSA (gap >= 2h), rho_p/rho_f = 2,
    knCoeff = 1e5, e = 0.9, gravity by the Re_g = 40 recipe
    (U_ref = 0.1, nu = U_ref D / 40), zero-net-flux bulk forcing."""
    D, rr = 12.0, 2.0
    u_ref = 0.1
    nu = u_ref * D / 40.0
    g = u_ref ** 2 / ((rr - 1.0) * D)
    case = build_case(n, n, n, [BC_PERIODIC] * 6, visc=nu,
                      gravity=(0.0, 0.0, -g), kn_coeff=1e5, restitution=0.9,
                      bulk_switch=(1, 1, 1), u_max=2 * u_ref,
                      precision=precision, rtol=rtol)
    rng = np.random.Generator(np.random.PCG64(seed))
    cen = rsa_centres(rng, n_spheres, D, (n, n, n), 2.0, [True] * 3)
    mk, bd = place_spheres(case, [Sphere(tuple(c), D, rr, 452) for c in cen])
    phi = n_spheres * math.pi * D ** 3 / 6 / n ** 3
    return Setup(f"suspension-{n}^3-{n_spheres}x{int(D)}h", case, mk, bd,
                 zero_fields(case), {"steps": 500, "solid_fraction": phi,
                                     "seed": seed})


def fill_upward(rng, diameters, box, gap=1.0, wall_gap=1.0, slab=24.0, batch=256, rise=1.0):
    """Seeded bottom-up random sequential addition in a walled box: each
    sphere (in the given order) takes the lowest of `batch` random candidate
    positions in the slab [zcap - slab, zcap] that keeps a surface gap
    >= `gap` to every placed sphere and `wall_gap` to the walls; the slab
    rises by `rise` whenever no candidate fits. Returns (centres (n, 3),
    fill height = highest sphere top)."""
    L = np.asarray(box, dtype=np.float64)
    n = len(diameters)
    cen = np.zeros((n, 3))
    rad = 0.5 * np.asarray(diameters, dtype=np.float64)
    rmax = rad.max()
    zcap = 2.0 * rmax + wall_gap + slab
    for q in range(n):
        R = rad[q]
        lo = np.array([R + wall_gap, R + wall_gap, R + wall_gap])
        hi_xy = L[:2] - R - wall_gap
        while True:
            zlo = max(lo[2], zcap - slab)
            if zlo + R + wall_gap > L[2]:
                raise RuntimeError(f"fill_upward placed only {q} of {n} spheres")
            c = np.empty((batch, 3))
            c[:, :2] = lo[:2] + rng.random((batch, 2)) * (hi_xy - lo[:2])
            c[:, 2] = zlo + rng.random(batch) * (min(zcap, L[2] - R - wall_gap) - zlo)
            ok = c[:, 2] + R + wall_gap <= L[2]
            if q:
                near = np.nonzero(cen[:q, 2] >= zlo - R - rmax - gap)[0]
                if near.size:
                    d = c[:, None, :] - cen[near][None, :, :]
                    reach = rad[near][None, :] + R + gap
                    ok &= np.all(np.einsum("ijk,ijk->ij", d, d) >= reach * reach, axis=1)
            if ok.any():
                k = np.nonzero(ok)[0]
                cen[q] = c[k[np.argmin(c[k, 2])]]
                break
            zcap += rise
    return cen, float(np.max(cen[:, 2] + rad))


def config_fluidized_bed(precision=MIXED, rtol=1e-10, seed=SEED_BASE + 3, n_spheres=10000,
                         n=(192, 192, 768)) -> Setup:
    """Polydisperse fluidized-bed channel: 192x192x768, inflow w = U_in at
    z = 0, outflow at the top, no-slip side walls; 10,000 spheres with
    ln D ~ N(ln 10, 0.25^2) clipped to [8h, 16h], placed bottom-up
    (fill_upward; the fill height is recorded: random addition cannot put
    them in the lower 40 %; surface gaps >= 1h, so marker shells start
    outside the contact distance), rho_p/rho_f = 2.5, about one marker per
    h^2.
    Re_g = 40 at D = 10h with nu = 0.04 fixes g; U_in = U_ref =
    sqrt((rho_r - 1) g D) = 0.16; knCoeff = 1e5, e = 0.9. The fluid starts as
    the uniform upflow w = U_in."""
    nx, ny, nz = n
    rng = np.random.Generator(np.random.PCG64(seed))
    nu, rr, Dref = 0.04, 2.5, 10.0
    g = (40.0 * nu) ** 2 / ((rr - 1.0) * Dref ** 3)
    u_ref = math.sqrt((rr - 1.0) * g * Dref)
    diam = np.clip(np.exp(rng.normal(math.log(10.0), 0.25, size=n_spheres)), 8.0, 16.0)
    cen, height = fill_upward(rng, diam, (nx, ny, nz))
    bt = [BC_WALL, BC_WALL, BC_WALL, BC_WALL, BC_INFLOW, BC_OUTFLOW]
    case = build_case(nx, ny, nz, bt, visc=nu, gravity=(0.0, 0.0, -g), kn_coeff=1e5, restitution=0.9,
                      u_max=2 * u_ref, precision=precision, rtol=rtol, inflow=(0.0, 0.0, u_ref))
    mk, bd = place_spheres(case, [Sphere(tuple(c), float(D), rr) for c, D in zip(cen, diam)])
    u, v, w, p = zero_fields(case)
    w[:] = u_ref
    phi = float(np.sum(np.pi * diam ** 3 / 6) / (nx * ny * height))
    return Setup(f"fluidized-bed-{nx}x{ny}x{nz}-{n_spheres}", case, mk, bd, (u, v, w, p),
                 {"steps": 300, "seed": seed, "fill_height": height, "fill_fraction": phi,
                  "u_in": u_ref, "g": g})


def cylinder_markers(centre, D, span, n_around, n_span, h=1.0):
    """Structured circumferential x spanwise lattice on a cylinder (axis
    along z) retracted by 0.3h like the sphere template; each marker carries
    the area it represents times h as its volume."""
    rs = 0.5 * D - 0.3 * h
    th = 2.0 * np.pi * (np.arange(n_around) + 0.5) / n_around
    zz = (np.arange(n_span) + 0.5) * span / n_span
    T, Z = np.meshgrid(th, zz, indexing="ij")
    x = np.stack([centre[0] + rs * np.cos(T).ravel(), centre[1] + rs * np.sin(T).ravel(), Z.ravel()])
    vol = 2.0 * np.pi * rs * span / (n_around * n_span) * h
    return x, vol


def config_cylinder(precision=MIXED, rtol=1e-10, seed=SEED_BASE + 4, n_free=200) -> Setup:
    """Fixed spanwise-extruded cylinder in crossflow: 512x256x64, uniform
    inflow u = 1 at x = 0, outflow at x = 512, periodic in y and in the span
    z. Cylinder D = 32h on the z axis through (128, 128), fixed, 2,295
    markers on a 51 x 45 circumferential x spanwise lattice (1.9h x 1.4h).
    Re = U D / nu = 200 -> nu = 0.16; forward Euler needs dt = 0.072
    (advective FTCS limit at U_max = 2). 200 free spheres of D = 8h
    (201 markers, rho_p/rho_f = 2) by seeded random addition with centres in
    x in [16, 96]; knCoeff = 1e5, e = 0.9, no gravity."""
    nx, ny, nz = 512, 256, 64
    rng = np.random.Generator(np.random.PCG64(seed))
    D, U, Re = 32.0, 1.0, 200.0
    nu = U * D / Re
    bt = [BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC]
    case = build_case(nx, ny, nz, bt, visc=nu, kn_coeff=1e5, restitution=0.9, u_max=2 * U,
                      precision=precision, rtol=rtol, inflow=(U, 0.0, 0.0))
    free = rsa_centres(rng, n_free, 8.0, (112.0, ny, nz), 2.0, [False, True, True], margin=16.0)
    mk, bd = place_spheres(case, [Sphere(tuple(c), 8.0, 2.0) for c in free])
    # the cylinder is body n_free, after the free spheres' markers
    cx, cvol = cylinder_markers((128.0, 128.0), D, float(nz), 51, 45)
    nc = cx.shape[1]
    m0, b0 = mk.x.shape[1], bd.mass.size
    mk2, bd2 = MarkerArrays.empty(m0 + nc), BodyArrays.empty(b0 + 1)
    for k in mk.__dataclass_fields__:
        getattr(mk2, k)[..., :m0] = getattr(mk, k)          # (3, M) or (M,)
    for k in bd.__dataclass_fields__:
        a, b = getattr(bd, k), getattr(bd2, k)
        if k in BODY_MAT:
            b[:b0] = a                                        # (B, 9)
        else:
            b[..., :b0] = a                                   # (3, B) or (B,)
    rho_c = 2.0
    mass = rho_c * math.pi * (0.5 * D) ** 2 * nz
    mk2.x[:, m0:] = cx
    mk2.vol[m0:] = cvol
    mk2.mass[m0:] = rho_c * cvol
    mk2.radeff[m0:] = 0.75
    mk2.body[m0:] = b0
    bd2.x[:, b0] = (128.0, 128.0, 0.5 * nz)
    bd2.mass[b0] = mass
    bd2.rho[b0] = rho_c
    Ixx = mass * (3 * (0.5 * D) ** 2 + nz ** 2) / 12.0
    Izz = 0.5 * mass * (0.5 * D) ** 2
    bd2.inertia[b0] = [Ixx, 0, 0, 0, Ixx, 0, 0, 0, Izz]
    bd2.rot[b0] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    bd2.rot_glo[b0] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    bd2.fixed[b0] = 1
    case.n_markers = m0 + nc
    case.n_bodies = b0 + 1
    u, v, w, p = zero_fields(case)
    u[:] = U
    return Setup(f"cylinder-crossflow-{nx}x{ny}x{nz}", case, mk2, bd2, (u, v, w, p),
                 {"steps": 2000, "seed": seed, "Re": Re, "D": D, "U": U, "span": float(nz),
                  "cylinder_body": b0, "cylinder_markers": nc})


def config_batch(n_cases=64, precision=MIXED, seed=SEED_BASE + 5, rtol=1e-10,
                 max_spheres=40) -> list[Setup]:
    """Independent settling cases of uneven size for the batched form: grids
    with nx, ny drawn from 32..96 and nz from 64..192 (multiples of 16), six
    no-slip walls, 1..max_spheres spheres of D uniform in [8h, 16h]
    (rho_p/rho_f = 1.5, the single-sphere settling physics: nu = 0.04,
    g from Re_g = 40 at D = 12h) placed by seeded random sequential addition
    inside the walls (gap >= 2h; a case keeps the spheres that fit)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nu, rr, u_ref = 0.04, 1.5, 0.1
    g = u_ref ** 2 / ((rr - 1.0) * 12.0)
    out = []
    for q in range(n_cases):
        nx, ny = (int(16 * rng.integers(2, 7)) for _ in range(2))
        nz = int(16 * rng.integers(4, 13))
        want = int(rng.integers(1, max_spheres + 1))
        diam = np.sort(rng.uniform(8.0, 16.0, size=want))[::-1]
        L = np.array([nx, ny, nz], dtype=np.float64)
        cen, dia = [], []
        for D in diam:
            R = 0.5 * D
            lo, hi = R + 1.0, L - R - 1.0
            if np.any(hi <= lo):
                continue
            for _ in range(400):
                c = lo + rng.random(3) * (hi - lo)
                if all(np.linalg.norm(c - c2) >= R + 0.5 * d2 + 2.0 for c2, d2 in zip(cen, dia)):
                    cen.append(c)
                    dia.append(D)
                    break
        case = build_case(nx, ny, nz, [BC_WALL] * 6, visc=nu, gravity=(0.0, 0.0, -g), u_max=2 * u_ref,
                          precision=precision, rtol=rtol)
        mk, bd = place_spheres(case, [Sphere(tuple(c), D, rr) for c, D in zip(cen, dia)])
        out.append(Setup(f"batch-{q}-{nx}x{ny}x{nz}-{len(cen)}", case, mk, bd, zero_fields(case),
                         {"seed": seed, "spheres": len(cen)}))
    return out


# ---------------------------------------------------------------------------
# The reference's own case: 86 tilted cylinders (InitCylinderParcelsDEM)
# ---------------------------------------------------------------------------
def read_marker_template(path) -> np.ndarray:
    """A marker template file of the reference (CYLINDER_LP*.dat: one marker
    per line, three columns x y z, read list-directed at
    ZJS_INIT_DEM.F90:371-373) as an (n, 3) float64 array."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                rows.append([float(v) for v in parts[:3]])
    return np.asarray(rows, dtype=np.float64)


def rotation_matrix(theta) -> np.ndarray:
    """CalculateRotateMatrix (couplesolverCylinder.F90:76-97): Rodrigues'
    formula for the rotation by |theta| about theta/|theta| (|theta| clamped
    to >= 1e-10), as the 9 row-major entries, in the reference's operation
    order."""
    t0, t1, t2 = (float(v) for v in theta)
    mag = math.sqrt(t0 * t0 + t1 * t1 + t2 * t2)
    mag = mag if mag > 1e-10 else 1e-10
    ux, uy, uz = t0 / mag, t1 / mag, t2 / mag
    cm, sm = math.cos(mag), math.sin(mag)
    return np.array([(1 - cm) * ux * ux + cm, (1 - cm) * ux * uy - sm * uz, (1 - cm) * ux * uz + sm * uy,
                     (1 - cm) * uy * ux + sm * uz, (1 - cm) * uy * uy + cm, (1 - cm) * uy * uz - sm * ux,
                     (1 - cm) * uz * ux - sm * uy, (1 - cm) * uz * uy + sm * ux, (1 - cm) * uz * uz + cm])


def reference_cylinder_centres(n_bodies=86):
    """Body centres of InitCylinderParcelsDEM (ZJS_INIT_DEM.F90:382-394), in
    metres: a 9 x 2 x 9 lattice of pitch 1e-3 m visited with stride 7, the y
    positions mapped onto the 1.2e-3 m span."""
    delta = 3e-3 / 3
    ntx, nty, ntz = 9, 2, 9
    nt = ntx * nty * ntz
    out = np.zeros((n_bodies, 3))
    for i in range(1, n_bodies + 1):
        idx = (7 * i - 1) % nt + 1
        iz = (idx - 1) // (ntx * nty)
        iy = (idx - 1 - iz * ntx * nty) // ntx
        ix = (idx - 1) % ntx
        x = ix * delta + delta * 0.5 + 1e-10
        y = iy * delta + delta * 0.5 + 1e-10
        y = y / delta * 1.2e-3 / 2 + 1e-10
        z = iz * delta + delta * 0.5 + 1e-10
        out[i - 1] = (x, y, z)
    return out


def config_reference_cylinders(template: np.ndarray, seed=SEED_BASE + 6, n_bodies=86, precision=FP64,
                               rtol=1e-10) -> Setup:
    """The reference's shipped case in its own SI units (main.F90:20,
    initialization.F90:241-270,341-377, ZJS_INIT_DEM.F90:67-78,350-485): a
    240 x 32 x 240 grid of 9e-3 x 1.2e-3 x 9e-3 m, periodic in x and z,
    no-slip walls in y, water (nu = 1e-6, rho_f = 1e3), dt = 5e-5 s, bulk
    forcing to a mean u of 0.1 m/s (and w = 0), gravity -9.81 m/s^2 along z,
    knCoeff 4e4, e = 1e-4; 86 cylinders (rho = 2650) from the marker template
    (template x, y scaled by (rad - 0.3 dx) height / rad, z by height - 0.6 dx,
    rad = the binary32 literal 2.08e-4, height = 4 rad), each tilted about y
    by pi/2 times a seeded uniform number (the reference draws it unseeded).
    Marker volume pi dx / n (rad^2 + 2 rad height + dx^2 / 4); body volume
    pi rad^2 height, inertia m (3 rad^2 + height^2) / 12 about x and y and
    m rad^2 / 2 about the axis (ZJS_INIT_DEM.F90:422-431,477)."""
    nx, ny, nz = 240, 32, 240
    dx = 9e-3 / nx
    case = build_case(nx, ny, nz, [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC],
                      h=dx, visc=1e-6, rhof=1e3, dt=1e-4 / 2, gravity=(0.0, 0.0, -9.81), kn_coeff=4e4,
                      restitution=0.0001, bulk_switch=(1, 0, 1), bulk_target=(0.1, 0.0, 0.0),
                      precision=precision, rtol=rtol)
    case.kn_coeff = 1e4 * 4          # SI: the reference's own value
    rng = np.random.Generator(np.random.PCG64(seed))
    rnd = rng.random(n_bodies)
    pi = math.acos(-1.0)
    rad = float(np.float32(2.08e-4))
    height = rad * 4
    radratio = (rad - 0.3 * dx) * height / rad
    heightratio = height - 0.3 * 2 * dx
    rho = 2.65e3
    n = template.shape[0]
    centres = reference_cylinder_centres(n_bodies)
    mk = MarkerArrays.empty(n * n_bodies)
    bd = BodyArrays.empty(n_bodies)
    vol_m = pi * dx / n * (rad ** 2 + 2 * rad * height + 0.25 * dx ** 2)
    for j in range(n_bodies):
        c = centres[j]
        R = rotation_matrix((0.0, pi / 2.0 * rnd[j], 0.0))
        xt = np.empty((n, 3))
        xt[:, 0] = radratio * template[:, 0] + c[0]
        xt[:, 1] = radratio * template[:, 1] + c[1]
        xt[:, 2] = heightratio * template[:, 2] + c[2]
        # RotateLagrangianPoints (couplesolverCylinder.F90:60-74)
        t = xt - c
        xr = np.empty_like(xt)
        xr[:, 0] = R[0] * t[:, 0] + R[1] * t[:, 1] + R[2] * t[:, 2]
        xr[:, 1] = R[3] * t[:, 0] + R[4] * t[:, 1] + R[5] * t[:, 2]
        xr[:, 2] = R[6] * t[:, 0] + R[7] * t[:, 1] + R[8] * t[:, 2]
        xr = xr + c
        o = j * n
        mk.x[:, o:o + n] = xr.T
        mk.vol[o:o + n] = vol_m
        mk.mass[o:o + n] = rho * vol_m
        mk.radeff[o:o + n] = dx * 0.75
        mk.body[o:o + n] = j
        vol_b = (pi * rad ** 2) * height
        m = vol_b * rho
        bd.x[:, j] = c
        bd.mass[j] = m
        bd.rho[j] = rho
        i1 = m * (3.0 * rad ** 2 + height ** 2) / 12.0
        bd.inertia[j] = [i1, 0, 0, 0, i1, 0, 0, 0, 0.5 * m * rad ** 2]
        bd.rot[j] = R
        bd.rot_glo[j] = R
    case.n_markers = n * n_bodies
    case.n_bodies = n_bodies
    return Setup(f"reference-cylinders-{nx}x{ny}x{nz}-{n_bodies}", case, mk, bd, zero_fields(case),
                 {"seed": seed, "rad": rad, "height": height, "dx": dx, "tilt": rnd})


# ---------------------------------------------------------------------------
# Small cases for parity tests
# ---------------------------------------------------------------------------
def small_periodic(n=(32, 24, 28), n_spheres=6, D=8.0, rho=2.0, seed=SEED_BASE + 11,
                   precision=FP64, rtol=1e-12, contact=True, gap=0.9,
                   velocities=True) -> Setup:
    """Triply periodic box with a cluster of spheres in near contact (marker
    shells closer than the 1.65h candidate distance), some straddling the
    periodic faces, random initial body and fluid velocities."""
    nx, ny, nz = n
    rng = np.random.Generator(np.random.PCG64(seed))
    case = build_case(nx, ny, nz, [BC_PERIODIC] * 6, visc=0.05,
                      gravity=(0.0, 0.0, -1e-3), kn_coeff=1e5, restitution=0.9,
                      bulk_switch=(1, 0, 1), bulk_target=(0.01, 0.0, 0.0),
                      u_max=0.3, precision=precision, rtol=rtol)
    spheres = []
    # a 3 x 2 grid at pitch D + gap (marker shells 2*0.3h + gap apart, inside
    # the 1.65h candidate distance), straddling the x and y periodic faces
    pitch = D + (gap if contact else 4.0)
    start = np.array([nx - 0.5 * D + 0.37, 0.5 * ny + 0.21, 0.5 * nz - 0.13])
    for k in range(n_spheres):
        off = np.array([(k % 3) * pitch, (k // 3) * pitch, 0.35 * math.sin(1.3 * k)])
        c = np.mod(start + off, [nx, ny, nz])
        spheres.append(Sphere(tuple(c), D, rho, markers_for_diameter(D)))
    mk, bd = place_spheres(case, spheres)
    if velocities:
        bd.v[:] = rng.normal(0.0, 0.02, size=bd.v.shape)
        bd.vold[:] = bd.v
        bd.omega[:] = rng.normal(0.0, 0.01, size=bd.omega.shape)
        for b in range(bd.mass.size):
            inert = bd.inertia[b].reshape(3, 3)
            bd.angmom[:, b] = inert @ bd.omega[:, b]
        bd.angmom_old[:] = bd.angmom
        # marker velocities consistent with the rigid motion
        for i in range(mk.body.size):
            b = mk.body[i]
            r = mk.x[:, i] - bd.x[:, b]
            L = np.array([nx, ny, nz], dtype=np.float64)
            r -= L * np.round(r / L)
            mk.v[:, i] = bd.v[:, b] + np.cross(bd.omega[:, b], r)
    u, v, w, p = zero_fields(case)
    if velocities:
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        u[:] = 0.02 * np.sin(2 * np.pi * y / ny) * np.cos(2 * np.pi * z / nz)
        v[:] = 0.015 * np.cos(2 * np.pi * x / nx) * np.sin(4 * np.pi * z / nz)
        w[:] = 0.01 * np.sin(2 * np.pi * x / nx + 0.3) * np.cos(2 * np.pi * y / ny)
        p[:] = 0.001 * rng.normal(size=p.shape)
    return Setup(f"periodic-{nx}x{ny}x{nz}-{n_spheres}", case, mk, bd, (u, v, w, p),
                 {"seed": seed})


def small_walls(n=(24, 20, 32), D=8.0, rho=1.5, seed=SEED_BASE + 12,
                precision=FP64, rtol=1e-12, periodic_x=False) -> Setup:
    """Closed box (or x-periodic channel) with two spheres: one touching the
    bottom wall (wall contact, wall-mirror markers), one above it in near
    contact, both off-centre so every wall sees mirror images."""
    nx, ny, nz = n
    rng = np.random.Generator(np.random.PCG64(seed))
    bt = [BC_WALL] * 6
    if periodic_x:
        bt[0] = bt[1] = BC_PERIODIC
    case = build_case(nx, ny, nz, bt, visc=0.04, gravity=(0.0, 0.0, -2e-3),
                      u_max=0.3, precision=precision, rtol=rtol)
    R = 0.5 * D
    c1 = (R + 0.45, R + 0.25, R - 0.3 * 1.0 + 0.05)
    c2 = (R + 1.2, R + 0.9, c1[2] + D + 0.6)
    mk, bd = place_spheres(case, [Sphere(c1, D, rho), Sphere(c2, D, rho)])
    bd.v[:] = rng.normal(0.0, 0.02, size=bd.v.shape)
    bd.v[2, 0] = -0.03
    bd.vold[:] = bd.v
    for i in range(mk.body.size):
        mk.v[:, i] = bd.v[:, mk.body[i]]
    u, v, w, p = zero_fields(case)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u[:] = 0.01 * np.sin(np.pi * (y + 0.5) / ny) * np.sin(np.pi * (z + 0.5) / nz)
    v[:] = 0.01 * np.sin(np.pi * (x + 0.5) / nx) * np.sin(np.pi * (z + 0.5) / nz)
    w[:] = 0.01 * np.sin(np.pi * (x + 0.5) / nx) * np.sin(np.pi * (y + 0.5) / ny)
    p[:] = 0.001 * rng.normal(size=p.shape)
    return Setup(f"walls-{nx}x{ny}x{nz}", case, mk, bd, (u, v, w, p), {"seed": seed})


def small_open(n=(32, 20, 24), axis=0, U=0.05, D=8.0, rho=1.5, seed=SEED_BASE + 13,
               precision=FP64, rtol=1e-12, periodic=(), time_scheme=0, gap=0.75,
               spheres=True) -> Setup:
    """Channel with an inflow side (low `axis`, normal velocity U) and an
    outflow side (high `axis`); the other sides walls, or periodic for the
    dimensions in `periodic`. Two spheres in near contact sit against the
    inflow face (their markers reach the inflow ghost layers and their mirror
    images spread into the domain), a third near the outflow face; random
    smooth initial fields with the inflow face at U (spheres=False: the
    fluid alone)."""
    nx, ny, nz = n
    rng = np.random.Generator(np.random.PCG64(seed))
    bt = [BC_WALL] * 6
    for d in periodic:
        bt[2 * d] = bt[2 * d + 1] = BC_PERIODIC
    bt[2 * axis], bt[2 * axis + 1] = BC_INFLOW, BC_OUTFLOW
    inflow = [0.0, 0.0, 0.0]
    inflow[axis] = U
    case = build_case(nx, ny, nz, bt, visc=0.04, gravity=(0.0, 0.0, -1e-3), u_max=0.3,
                      precision=precision, rtol=rtol, inflow=inflow, time_scheme=time_scheme)
    L = np.array([nx, ny, nz], dtype=np.float64)
    R = 0.5 * D
    o = (axis + 1) % 3
    c1 = 0.5 * L + np.array([0.31, -0.27, 0.19])
    c1[axis] = R + 0.4
    c1[o] = R + 0.6
    c2 = c1.copy()
    c2[o] += D + gap            # marker shells inside the 1.65h candidate distance
    c2[axis] += 0.35
    c3 = 0.5 * L + np.array([-0.23, 0.41, -0.17])
    c3[axis] = L[axis] - R - 0.7
    sph = [Sphere(tuple(c1), D, rho), Sphere(tuple(c2), D, rho), Sphere(tuple(c3), D, rho)]
    mk, bd = place_spheres(case, sph if spheres else [])
    bd.v[:] = rng.normal(0.0, 0.01, size=bd.v.shape)
    bd.vold[:] = bd.v
    for i in range(mk.body.size):
        mk.v[:, i] = bd.v[:, mk.body[i]]
    u, v, w, p = zero_fields(case)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = [u, v, w]
    f[axis][:] = U
    u += 0.01 * np.sin(np.pi * (y + 0.5) / ny + 0.2) * np.cos(np.pi * (z + 0.5) / nz)
    v += 0.01 * np.sin(np.pi * (x + 0.5) / nx) * np.cos(np.pi * (z + 0.5) / nz + 0.1)
    w += 0.01 * np.cos(np.pi * (x + 0.5) / nx) * np.sin(np.pi * (y + 0.5) / ny)
    p[:] = 0.001 * rng.normal(size=p.shape)
    return Setup(f"open{'xyz'[axis]}-{nx}x{ny}x{nz}", case, mk, bd, (u, v, w, p), {"seed": seed})


def load(sim, setup: Setup) -> None:
    """Upload a setup's initial state into a Simulation."""
    sim.upload_bodies(setup.bodies)
    sim.upload_markers(setup.markers)
    sim.upload_fields(*setup.fields)
