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
    solve    int32 iterations, int32 converged, f64 rel_residual, f64 rhs_norm
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import (BODY_DIAG, BODY_MAT, BODY_VEC3, BodyArrays, Case,
                  MarkerArrays, SolveInfo)

MAGIC = b"IBMDEMS1"
_MARKER_F64 = ("x", "v", "fdrag", "fcol", "vol", "mass", "radeff")


def _state(sim):
    fields = tuple(sim.download_fields())
    return sim.download_markers(), sim.download_bodies(), fields


def write(path, case: Case, markers: MarkerArrays, bodies: BodyArrays,
          fields, info: SolveInfo | None = None) -> None:
    info = info or SolveInfo()
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
        f.write(np.array([info.iterations, info.converged], "<i4").tobytes())
        f.write(np.array([info.rel_residual, info.rhs_norm], "<f8").tobytes())


def write_simulation(path, sim, info: SolveInfo | None = None) -> None:
    """Snapshot of a live Simulation (downloads its state)."""
    markers, bodies, fields = _state(sim)
    write(path, sim.case, markers, bodies, fields, info)


def read(path):
    """-> (case, markers, bodies, fields, info)."""
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
    it = take("<i4", (2,))
    res = take("<f8", (2,))
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    info = SolveInfo(int(it[0]), int(it[1]), float(res[0]), float(res[1]))
    return case, mk, bd, fields, info
