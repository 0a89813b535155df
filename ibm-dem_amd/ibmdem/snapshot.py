"""Snapshot files: the whole state of one case in one little-endian stream.

The Fortran host side (``fortran/ibmdem_host.f90``, ReadSnapshot and
WriteSnapshot) reads and writes the same layout, so a state prepared here can
be advanced by the Fortran time loop and read back, and a run can be stopped
and resumed bit for bit.

Layout::

    b"IBMDEMS1", int32 sizeof(ibmdem_case), the ibmdem_case bytes,
    markers  x v fdrag fcol (3M each, component-major) vol mass radeff (M),
             body (M int32),
    bodies   x v vold omega angmom angmom_old (3B each, component-major),
             rot rot_glo inertia (9B each, row-major per body),
             mass rho (B), fixed (B int32),
             fdrag fmom fcol fmomcol fhydro fmomhydro (3B each),
    fields   u v w p (nx*ny*nz each, i fastest),
    planes   the outflow plane of each outflow side 1, 3, 5 in that order
             (ny*nz, nx*nz, nx*ny; absent for other sides),
    solve    int32 iterations, int32 converged, f64 rel_residual, f64 rhs_norm
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import (BC_OUTFLOW, BODY_DIAG, BODY_MAT, BODY_VEC3, BodyArrays, Case,
                  MarkerArrays, SolveInfo)

MAGIC = b"IBMDEMS1"
_MARKER_F64 = ("x", "v", "fdrag", "fcol", "vol", "mass", "radeff")


def outflow_sides(case: Case):
    return [s for s in (1, 3, 5) if case.btype[s] == BC_OUTFLOW]


def _state(sim):
    fields = tuple(sim.download_fields())
    planes = {s: sim.download_outflow(s) for s in outflow_sides(sim.case)}
    return sim.download_markers(), sim.download_bodies(), fields, planes


def write(path, case: Case, markers: MarkerArrays, bodies: BodyArrays,
          fields, info: SolveInfo | None = None, planes: dict | None = None) -> None:
    """planes: {side: outflow plane}; a missing outflow side gets the last
    interior face of its normal component (what a field upload sets)."""
    info = info or SolveInfo()
    planes = dict(planes or {})
    for s in outflow_sides(case):
        if s not in planes:
            d = s // 2
            planes[s] = np.take(np.asarray(fields[d]), -1, axis=2 - d)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.int32(C.sizeof(Case)).tobytes())
        f.write(bytes(case))
        for k in _MARKER_F64:
            f.write(np.ascontiguousarray(getattr(markers, k), "<f8").tobytes())
        f.write(np.ascontiguousarray(markers.body, "<i4").tobytes())
        for k in BODY_VEC3 + BODY_MAT + ("mass", "rho"):
            f.write(np.ascontiguousarray(getattr(bodies, k), "<f8").tobytes())
        f.write(np.ascontiguousarray(bodies.fixed, "<i4").tobytes())
        for k in BODY_DIAG:
            f.write(np.ascontiguousarray(getattr(bodies, k), "<f8").tobytes())
        for a in fields:
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
        for s in outflow_sides(case):
            f.write(np.ascontiguousarray(planes[s], "<f8").tobytes())
        f.write(np.array([info.iterations, info.converged], "<i4").tobytes())
        f.write(np.array([info.rel_residual, info.rhs_norm], "<f8").tobytes())


def write_simulation(path, sim, info: SolveInfo | None = None) -> None:
    """Snapshot of a live Simulation (downloads its state)."""
    markers, bodies, fields, planes = _state(sim)
    write(path, sim.case, markers, bodies, fields, info, planes)


def read(path, planes: bool = False):
    """-> (case, markers, bodies, fields, info), and the outflow planes
    ({side: array}) as a sixth item when `planes`."""
    raw = open(path, "rb").read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not an ibmdem snapshot")
    size = int(np.frombuffer(raw, "<i4", 1, 8)[0])
    if size != C.sizeof(Case):
        raise ValueError(f"{path}: case record of {size} bytes, expected {C.sizeof(Case)}")
    off = 12
    case = Case.from_buffer_copy(raw[off:off + size])
    off += size

    def take(dtype, shape):
        nonlocal off
        n = int(np.prod(shape))
        a = np.frombuffer(raw, dtype, n, off).reshape(shape).copy()
        off += n * np.dtype(dtype).itemsize
        return a

    m, b = case.n_markers, case.n_bodies
    mk = MarkerArrays.empty(m)
    for k in _MARKER_F64:
        setattr(mk, k, take("<f8", getattr(mk, k).shape).astype(np.float64))
    mk.body = take("<i4", (m,)).astype(np.int32)
    bd = BodyArrays.empty(b)
    for k in BODY_VEC3 + BODY_MAT + ("mass", "rho"):
        setattr(bd, k, take("<f8", getattr(bd, k).shape).astype(np.float64))
    bd.fixed = take("<i4", (b,)).astype(np.int32)
    for k in BODY_DIAG:
        setattr(bd, k, take("<f8", (3, b)).astype(np.float64))
    shape = (case.nz, case.ny, case.nx)
    fields = tuple(take("<f8", shape).astype(np.float64) for _ in range(4))
    pshape = {1: (case.nz, case.ny), 3: (case.nz, case.nx), 5: (case.ny, case.nx)}
    pl = {s: take("<f8", pshape[s]).astype(np.float64) for s in outflow_sides(case)}
    it = take("<i4", (2,))
    res = take("<f8", (2,))
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    info = SolveInfo(int(it[0]), int(it[1]), float(res[0]), float(res[1]))
    if planes:
        return case, mk, bd, fields, info, pl
    return case, mk, bd, fields, info


def load(sim, path) -> SolveInfo:
    """Restore a snapshot into a Simulation of the same case (fields, then
    outflow planes, since a field upload resets them)."""
    case, mk, bd, fields, info, pl = read(path, planes=True)
    if bytes(case) != bytes(sim.case):
        raise ValueError(f"{path}: snapshot of another case")
    if case.n_bodies:
        sim.upload_bodies(bd)
    if case.n_markers:
        sim.upload_markers(mk)
    sim.upload_fields(*fields)
    for s, a in pl.items():
        sim.upload_outflow(s, a)
    return info
