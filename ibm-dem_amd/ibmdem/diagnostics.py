"""Diagnostics and output of a run: force coefficients, per-body probe lines
and Tecplot-readable field dumps.

The formats follow the reference's post-processing:
  - field dumps: OutputFluidTecplot (io.F90:43-77) writes
    ``variables=x,y,z,u,v,w,vel,P`` and ``zone i=..,j=..,k=..`` followed by one
    line per cell (cell centre, the staggered u, v, w as stored, |u|, P);
    here the interior cells are written (the reference includes its halo);
  - probes: PostProcessing (fluid.F90:368-404) appends per body
    ``step, x(3), v(3), omega(3), fdrag/m*rhof (3), fmom*rhof (3), fcol/m (3)``;
    the reference divides the torque by I, which is 0 for its cylinders
    (fluid.F90:396, ZJS_INIT_DEM.F90:427); the torque itself is written here;
  - force coefficients: C = 2 F / (rho_f U^2 D L) from the assembled
    hydrodynamic force of a body (AssembleLocalForceMS, ZJS_MS_DEM.F90:494-519);
  - the discrete divergence of a velocity state (VelocityDivergence's
    stencil, fluid.F90:97-100), outflow planes included.
"""
from __future__ import annotations

import numpy as np

from .abi import BC_OUTFLOW, BC_PERIODIC, BodyArrays, Case


def force_coefficients(bodies: BodyArrays, body: int, rho: float, U: float, D: float, L: float,
                       drag_axis: int = 0, lift_axis: int = 1):
    """(C_D, C_L) of one body: C = 2 F / (rho U^2 D L), F the body's
    assembled hydrodynamic force (bodies.fdrag, component-major)."""
    q = 0.5 * rho * U * U * D * L
    if not q > 0:
        raise ValueError("rho, U, D and L must be positive")
    F = bodies.fdrag[:, body]
    return float(F[drag_axis] / q), float(F[lift_axis] / q)


def divergence(case: Case, fields, planes=None):
    """Cell divergence of the staggered velocity (u, v, w of shape
    (nz, ny, nx)) with the boundary rules: the high face of a dimension is
    the wrapped first face (periodic), zero (wall) or the outflow plane
    (planes[side] for side 1, 3, 5; the array Simulation.download_outflow
    returns)."""
    h = (case.dx, case.dy, case.dz)
    d = 0.0
    for c in range(3):
        a = np.asarray(fields[c], dtype=np.float64)
        na = 2 - c
        side = 2 * c + 1
        if case.btype[side] == BC_OUTFLOW:
            hi = np.expand_dims(np.asarray(planes[side], dtype=np.float64), na)
        elif case.btype[side] == BC_PERIODIC:
            hi = np.take(a, [0], axis=na)
        else:
            hi = np.zeros_like(np.take(a, [0], axis=na))
        d = d + (np.concatenate([np.take(a, range(1, a.shape[na]), axis=na), hi], axis=na) - a) / h[c]
    return d


def shedding_statistics(t, cd, cl, D: float, U: float, start_fraction: float = 0.5):
    """Mean drag coefficient, rms lift coefficient and Strouhal number
    St = f D / U of a force-coefficient history, over the samples from
    start_fraction of the run on (past the start-up transient). f comes
    from the upward zero crossings of C_L - mean(C_L), located by linear
    interpolation between samples: (crossings - 1) periods between the
    first and the last. St is None when fewer than two crossings fall in
    the window."""
    t, cd, cl = (np.asarray(a, dtype=np.float64) for a in (t, cd, cl))
    k0 = int(len(t) * start_fraction)
    t, cd, cl = t[k0:], cd[k0:], cl[k0:]
    y = cl - cl.mean()
    up = np.nonzero((y[:-1] < 0) & (y[1:] >= 0))[0]
    times = t[up] + (t[up + 1] - t[up]) * (-y[up]) / (y[up + 1] - y[up])
    st = None
    if len(times) >= 2:
        f = (len(times) - 1) / (times[-1] - times[0])
        st = float(f * D / U)
    return {"mean_cd": float(cd.mean()), "rms_cl": float(np.sqrt(np.mean(y * y))), "strouhal": st,
            "periods": max(len(times) - 1, 0), "window": [float(t[0]), float(t[-1])]}


def probe_line(step: int, bodies: BodyArrays, body: int, rhof: float) -> str:
    """One line of the reference's per-body probe file (fluid.F90:398-401)."""
    b = body
    m = bodies.mass[b]
    vals = [*bodies.x[:, b], *bodies.v[:, b], *bodies.omega[:, b],
            *(bodies.fdrag[:, b] / m * rhof), *(bodies.fmom[:, b] * rhof), *(bodies.fcol[:, b] / m)]
    return f"{step:d} " + " ".join(f"{x:.16E}" for x in vals)


def write_tecplot(path, case: Case, fields) -> None:
    """Field dump in the reference's Tecplot layout (io.F90:55-77) over the
    interior cells, i fastest."""
    u, v, w, p = (np.asarray(a, dtype=np.float64) for a in fields)
    nz, ny, nx = u.shape
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    xc = (i - 0.5) * case.dx
    yc = (j - 0.5) * case.dy
    zc = (k - 0.5) * case.dz
    vel = np.sqrt(u * u + v * v + w * w)
    cols = np.stack([a.reshape(-1) for a in (xc, yc, zc, u, v, w, vel, p)], axis=1)
    with open(path, "w") as f:
        f.write(" variables=x,y,z,u,v,w,vel,P\n")
        f.write(f" zone i= {nx} j= {ny} k= {nz}\n")
        np.savetxt(f, cols, fmt="%.16E")


def read_tecplot(path):
    """-> (nx, ny, nz, array of shape (nx*ny*nz, 8)) of a dump written above."""
    with open(path) as f:
        head = f.readline()
        if "variables=x,y,z,u,v,w,vel,P" not in head:
            raise ValueError(f"{path}: not a field dump")
        zone = f.readline().replace("=", " ").split()
        nx, ny, nz = int(zone[2]), int(zone[4]), int(zone[6])
        data = np.loadtxt(f, ndmin=2)
    if data.shape != (nx * ny * nz, 8):
        raise ValueError(f"{path}: {data.shape[0]} rows for a {nx}x{ny}x{nz} zone")
    return nx, ny, nz, data
