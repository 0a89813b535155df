"""ctypes mirror of include/ibmdem.h and a thin wrapper around one library.

The wrapper takes the path of a shared library exporting the ibmdem C ABI.
The product path (``ibmdem.gpu_library``) only ever opens the HIP library
built in this package; the test suite opens the CPU oracle the same way, as
the checker. Arrays cross the boundary as contiguous numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

IBMDEM_OK = 0
ERRORS = {
    1: "HIP runtime error",
    2: "invalid argument",
    3: "capacity exceeded",
    4: "marker outside the domain",
    5: "Poisson solve did not converge",
    6: "non-finite value",
    7: "device allocation failed",
    8: "unsupported",
}
(ERR_HIP, ERR_ARG, ERR_CAPACITY, ERR_OUTSIDE, ERR_NOT_CONVERGED, ERR_NAN, ERR_ALLOC,
 ERR_UNSUPPORTED) = range(1, 9)
BC_PERIODIC = 0
BC_WALL = 1
BC_INFLOW = 2
BC_OUTFLOW = 3
FP64 = 0
MIXED = 1
SCHEME_EULER = 0
SCHEME_RK3 = 1

K_PREDICTOR, K_MG_SMOOTH, K_MATVEC, K_SPREAD, K_INTERP, K_CONTACT, K_PROJECT, \
    K_DIVERGENCE, K_CG_UPDATE = range(1, 10)

_i32 = C.c_int32
_f64 = C.c_double
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)


class Case(C.Structure):
    _fields_ = [
        ("nx", _i32), ("ny", _i32), ("nz", _i32),
        ("btype", _i32 * 6),
        ("dx", _f64), ("dy", _f64), ("dz", _f64),
        ("visc", _f64), ("rhof", _f64), ("dt", _f64),
        ("cp", _f64 * 3),
        ("bulk_switch", _i32 * 3),
        ("bulk_target", _f64 * 3),
        ("gravity", _f64 * 3),
        ("n_mdf", _i32),
        ("couple_interval", _i32),
        ("kn_coeff", _f64),
        ("restitution", _f64),
        ("collision_dist", _f64),
        ("collision_buffer", _f64),
        ("eta_textbook", _i32),
        ("n_markers", _i32),
        ("n_bodies", _i32),
        ("max_pairs", _i32),
        ("precision", _i32),
        ("poisson_rtol", _f64),
        ("poisson_max_iter", _i32),
        ("mg_pre", _i32), ("mg_post", _i32),
        ("mg_coarse_sweeps", _i32),
        ("mg_max_levels", _i32),
        ("stencil_tile", _i32),
        ("time_scheme", _i32),
        ("inflow", _f64 * 3),
        ("reserved", _i32 * 6),
    ]


class Markers(C.Structure):
    _fields_ = [("x", _pd), ("v", _pd), ("fdrag", _pd), ("fcol", _pd),
                ("vol", _pd), ("mass", _pd), ("radeff", _pd), ("body", _pi)]


class Bodies(C.Structure):
    _fields_ = [("x", _pd), ("v", _pd), ("vold", _pd),
                ("omega", _pd), ("angmom", _pd), ("angmom_old", _pd),
                ("rot", _pd), ("rot_glo", _pd), ("inertia", _pd),
                ("mass", _pd), ("rho", _pd), ("fixed", _pi),
                ("fdrag", _pd), ("fmom", _pd), ("fcol", _pd),
                ("fmomcol", _pd), ("fhydro", _pd), ("fmomhydro", _pd)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32),
                ("rel_residual", _f64), ("rhs_norm", _f64)]


class IbmdemError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)}"
                         + (f" ({detail})" if detail else ""))


_FUNCS = {
    "ibmdem_create": [C.POINTER(Case), C.POINTER(C.c_void_p)],
    "ibmdem_destroy": [C.c_void_p],
    "ibmdem_sync": [C.c_void_p],
    "ibmdem_upload_fields": [C.c_void_p, _pd, _pd, _pd, _pd],
    "ibmdem_download_fields": [C.c_void_p, _pd, _pd, _pd, _pd],
    "ibmdem_download_outflow": [C.c_void_p, _i32, _pd],
    "ibmdem_upload_outflow": [C.c_void_p, _i32, _pd],
    "ibmdem_download_force": [C.c_void_p, _pd, _pd, _pd],
    "ibmdem_download_force_noalloc": [C.c_void_p, _pd, _pd, _pd],
    "ibmdem_probe_bodies": [C.c_void_p, _i32, _pi, _i32],
    "ibmdem_get_stencil_tile": [C.c_void_p, _pi],
    "ibmdem_set_stencil_tile": [C.c_void_p, _i32],
    "ibmdem_probe_read": [C.c_void_p, C.POINTER(_i32), C.POINTER(C.c_int64), _pd],
    "ibmdem_download_poisson": [C.c_void_p, _pd, _pd],
    "ibmdem_upload_poisson_rhs": [C.c_void_p, _pd],
    "ibmdem_upload_markers": [C.c_void_p, C.POINTER(Markers)],
    "ibmdem_download_markers": [C.c_void_p, C.POINTER(Markers)],
    "ibmdem_upload_bodies": [C.c_void_p, C.POINTER(Bodies)],
    "ibmdem_download_bodies": [C.c_void_p, C.POINTER(Bodies)],
    "ibmdem_begin_step": [C.c_void_p],
    "ibmdem_ibm_interpolate": [C.c_void_p],
    "ibmdem_ibm_spread": [C.c_void_p],
    "ibmdem_fluid_predictor": [C.c_void_p, _f64, _f64, _f64],
    "ibmdem_bulk_forcing": [C.c_void_p],
    "ibmdem_divergence": [C.c_void_p, _f64],
    "ibmdem_poisson_solve": [C.c_void_p, C.POINTER(SolveInfo)],
    "ibmdem_project": [C.c_void_p, _f64],
    "ibmdem_apply_bc": [C.c_void_p],
    "ibmdem_body_hydro": [C.c_void_p],
    "ibmdem_dem_contact": [C.c_void_p],
    "ibmdem_body_assemble": [C.c_void_p],
    "ibmdem_body_integrate": [C.c_void_p],
    "ibmdem_markers_follow": [C.c_void_p],
    "ibmdem_dem_rebuild_lists": [C.c_void_p],
    "ibmdem_dem_substep": [C.c_void_p],
    "ibmdem_cloud_evolution": [C.c_void_p],
    "ibmdem_step": [C.c_void_p, C.POINTER(SolveInfo)],
    "ibmdem_run": [C.c_void_p, _i32, C.POINTER(SolveInfo)],
    "ibmdem_get_cells": [C.c_void_p, _pi],
    "ibmdem_get_pairs": [C.c_void_p, _i32, _pi, _pi, _pi],
    "ibmdem_profile_kernel": [C.c_void_p, _i32],
    "ibmdem_profile_read": [C.c_void_p, _pd, C.POINTER(C.c_int64), _pd],
    "ibmdem_memory_bytes": [C.c_void_p, C.POINTER(C.c_int64)],
    "ibmdem_get_mg_params": [C.c_void_p, _pi],
    "ibmdem_set_mg_params": [C.c_void_p, _pi],
    "ibmdem_batch_create": [_i32, C.POINTER(Case), C.POINTER(C.c_void_p)],
    "ibmdem_batch_destroy": [C.c_void_p],
    "ibmdem_batch_context": [C.c_void_p, _i32, C.POINTER(C.c_void_p)],
    "ibmdem_batch_run": [C.c_void_p, _i32, C.POINTER(SolveInfo)],
    "ibmdem_batch_autotune": [C.c_void_p, _i32],
    "ibmdem_batch_set_fused": [C.c_void_p, _i32],
    "ibmdem_batch_set_solve_groups": [C.c_void_p, _i32],
    "ibmdem_batch_get_solve_groups": [C.c_void_p, _pi],
    "ibmdem_batch_stats": [C.c_void_p, C.POINTER(C.c_int64), _pd],
    "ibmdem_batch_profile": [C.c_void_p, _i32],
    "ibmdem_batch_profile_read": [C.c_void_p, _i32, _pd, C.POINTER(C.c_int64), _pd],
}
EXPORTED = tuple(_FUNCS) + ("ibmdem_last_error", "ibmdem_batch_last_error")


def load_library(path: str) -> C.CDLL:
    lib = C.CDLL(path)
    for name, args in _FUNCS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.ibmdem_last_error.argtypes = [C.c_void_p]
    lib.ibmdem_last_error.restype = C.c_char_p
    lib.ibmdem_batch_last_error.argtypes = [C.c_void_p]
    lib.ibmdem_batch_last_error.restype = C.c_char_p
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags.c_contiguous, "arrays must be contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(_pd)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_pi)
    raise TypeError(a.dtype)


@dataclass
class MarkerArrays:
    x: np.ndarray        # (3, M)
    v: np.ndarray        # (3, M)
    fdrag: np.ndarray    # (3, M)
    fcol: np.ndarray     # (3, M)
    vol: np.ndarray      # (M,)
    mass: np.ndarray     # (M,)
    radeff: np.ndarray   # (M,)
    body: np.ndarray     # (M,) int32

    @classmethod
    def empty(cls, m: int) -> "MarkerArrays":
        z3 = lambda: np.zeros((3, m))
        return cls(z3(), z3(), z3(), z3(), np.zeros(m), np.zeros(m),
                   np.zeros(m), np.zeros(m, dtype=np.int32))

    def copy(self) -> "MarkerArrays":
        return MarkerArrays(**{k: getattr(self, k).copy()
                               for k in self.__dataclass_fields__})

    def struct(self) -> Markers:
        return Markers(_p(self.x), _p(self.v), _p(self.fdrag), _p(self.fcol),
                       _p(self.vol), _p(self.mass), _p(self.radeff),
                       _p(self.body))


BODY_VEC3 = ("x", "v", "vold", "omega", "angmom", "angmom_old")
BODY_MAT = ("rot", "rot_glo", "inertia")
BODY_DIAG = ("fdrag", "fmom", "fcol", "fmomcol", "fhydro", "fmomhydro")


@dataclass
class BodyArrays:
    x: np.ndarray
    v: np.ndarray
    vold: np.ndarray
    omega: np.ndarray
    angmom: np.ndarray
    angmom_old: np.ndarray
    rot: np.ndarray       # (B, 9)
    rot_glo: np.ndarray   # (B, 9)
    inertia: np.ndarray   # (B, 9)
    mass: np.ndarray
    rho: np.ndarray
    fixed: np.ndarray     # int32
    fdrag: np.ndarray
    fmom: np.ndarray
    fcol: np.ndarray
    fmomcol: np.ndarray
    fhydro: np.ndarray
    fmomhydro: np.ndarray

    @classmethod
    def empty(cls, b: int) -> "BodyArrays":
        kw = {k: np.zeros((3, b)) for k in BODY_VEC3 + BODY_DIAG}
        kw.update({k: np.zeros((b, 9)) for k in BODY_MAT})
        kw.update(mass=np.zeros(b), rho=np.zeros(b),
                  fixed=np.zeros(b, dtype=np.int32))
        return cls(**kw)

    def copy(self) -> "BodyArrays":
        return BodyArrays(**{k: getattr(self, k).copy()
                             for k in self.__dataclass_fields__})

    def struct(self) -> Bodies:
        return Bodies(*[_p(getattr(self, k)) for k in
                        ("x", "v", "vold", "omega", "angmom", "angmom_old",
                         "rot", "rot_glo", "inertia", "mass", "rho", "fixed",
                         "fdrag", "fmom", "fcol", "fmomcol", "fhydro",
                         "fmomhydro")])


class Simulation:
    """One ibmdem context on one library (GPU library or the CPU oracle)."""

    def __init__(self, lib: C.CDLL, case: Case):
        self.lib = lib
        self.case = case
        self.shape = (case.nz, case.ny, case.nx)
        self.n_markers = case.n_markers
        self.n_bodies = case.n_bodies
        h = C.c_void_p()
        rc = lib.ibmdem_create(C.byref(case), C.byref(h))
        if rc != IBMDEM_OK:
            raise IbmdemError(rc, "ibmdem_create")
        self.h = h

    # -- plumbing ----------------------------------------------------------
    def _check(self, rc: int, where: str) -> None:
        if rc != IBMDEM_OK:
            detail = self.lib.ibmdem_last_error(self.h)
            raise IbmdemError(rc, where, detail.decode() if detail else "")

    def call(self, name: str, *args) -> None:
        self._check(getattr(self.lib, name)(self.h, *args), name)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.ibmdem_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------
    def upload_fields(self, u=None, v=None, w=None, p=None) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (u, v, w, p)]
        for a in arrs:
            if a is not None and a.shape != self.shape:
                raise ValueError(f"field shape {a.shape} != {self.shape}")
        self.call("ibmdem_upload_fields", *[_p(a) for a in arrs])

    def download_fields(self):
        out = [np.zeros(self.shape) for _ in range(4)]
        self.call("ibmdem_download_fields", *[_p(a) for a in out])
        return out

    def outflow_shape(self, side: int):
        """Array shape of the outflow plane of side 1, 3 or 5 (slowest
        transverse dimension first)."""
        nz, ny, nx = self.shape
        shapes = {1: (nz, ny), 3: (nz, nx), 5: (ny, nx)}
        if side not in shapes:
            raise ValueError("outflow planes exist on sides 1, 3 and 5")
        return shapes[side]

    def download_outflow(self, side: int):
        """Normal velocity on the outflow face of `side` (the outflow plane)."""
        out = np.zeros(self.outflow_shape(side))
        self.call("ibmdem_download_outflow", side, _p(out))
        return out

    def upload_outflow(self, side: int, plane) -> None:
        a = np.ascontiguousarray(plane, dtype=np.float64)
        if a.shape != self.outflow_shape(side):
            raise ValueError(f"plane shape {a.shape} != {self.outflow_shape(side)}")
        self.call("ibmdem_upload_outflow", side, _p(a))

    def download_force(self):
        out = [np.zeros(self.shape) for _ in range(3)]
        self.call("ibmdem_download_force", *[_p(a) for a in out])
        return out

    PROBE_FIELDS = ("x", "v", "omega", "fdrag", "fmom", "fcol")

    def probe_bodies(self, bodies, capacity: int) -> None:
        """Record, after every step, the listed bodies' x, v, omega, fdrag,
        fmom, fcol (18 values) in a device ring of `capacity` steps."""
        b = np.ascontiguousarray(bodies, dtype=np.int32)
        self._probe_n = b.size
        self.call("ibmdem_probe_bodies", int(b.size), _p(b) if b.size else None, int(capacity))

    def probe_read(self):
        """(first_step, records) with records of shape (steps, bodies, 18),
        oldest step first."""
        n = _i32(0)
        first = C.c_int64(0)
        self.call("ibmdem_probe_read", C.byref(n), C.byref(first), None)
        out = np.zeros((n.value, getattr(self, "_probe_n", 0), 18))
        if n.value:
            self.call("ibmdem_probe_read", C.byref(n), C.byref(first), _p(out))
        return int(first.value), out

    def download_force_noalloc(self):
        """download_force through the library's fixed device scratch (no
        device allocation)."""
        out = [np.zeros(self.shape) for _ in range(3)]
        self.call("ibmdem_download_force_noalloc", *[_p(a) for a in out])
        return out

    def download_poisson(self):
        out = [np.zeros(self.shape) for _ in range(2)]
        self.call("ibmdem_download_poisson", *[_p(a) for a in out])
        return out

    def upload_poisson_rhs(self, rhs) -> None:
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self.call("ibmdem_upload_poisson_rhs", _p(rhs))

    def upload_markers(self, m: MarkerArrays) -> None:
        s = m.struct()
        self.call("ibmdem_upload_markers", C.byref(s))

    def download_markers(self) -> MarkerArrays:
        m = MarkerArrays.empty(self.n_markers)
        s = m.struct()
        self.call("ibmdem_download_markers", C.byref(s))
        return m

    def upload_bodies(self, b: BodyArrays) -> None:
        s = b.struct()
        self.call("ibmdem_upload_bodies", C.byref(s))

    def download_bodies(self) -> BodyArrays:
        b = BodyArrays.empty(self.n_bodies)
        s = b.struct()
        self.call("ibmdem_download_bodies", C.byref(s))
        return b

    def cells(self) -> np.ndarray:
        out = np.zeros((3, self.n_markers), dtype=np.int32)
        self.call("ibmdem_get_cells", _p(out))
        return out

    def pairs(self, cap: int | None = None):
        cap = cap or int(self.case.max_pairs or 200)
        m = self.n_markers
        count = np.zeros(m, dtype=np.int32)
        partner = np.full((m, cap), -1, dtype=np.int32)
        shift = np.full((m, cap), -1, dtype=np.int32)
        self.call("ibmdem_get_pairs", cap, _p(count), _p(partner), _p(shift))
        return count, partner, shift

    # -- operators ---------------------------------------------------------
    def begin_step(self): self.call("ibmdem_begin_step")
    def interpolate(self): self.call("ibmdem_ibm_interpolate")
    def spread(self): self.call("ibmdem_ibm_spread")
    def predictor(self, a=1.0, b=0.0, c=1.0): self.call("ibmdem_fluid_predictor", a, b, c)
    def bulk(self): self.call("ibmdem_bulk_forcing")
    def divergence(self, c=1.0): self.call("ibmdem_divergence", c)
    def project(self, c=1.0): self.call("ibmdem_project", c)
    def apply_bc(self): self.call("ibmdem_apply_bc")
    def body_hydro(self): self.call("ibmdem_body_hydro")
    def contact(self): self.call("ibmdem_dem_contact")
    def body_assemble(self): self.call("ibmdem_body_assemble")
    def body_integrate(self): self.call("ibmdem_body_integrate")
    def markers_follow(self): self.call("ibmdem_markers_follow")
    def rebuild_lists(self): self.call("ibmdem_dem_rebuild_lists")
    def dem_substep(self): self.call("ibmdem_dem_substep")
    def cloud_evolution(self): self.call("ibmdem_cloud_evolution")
    def sync(self): self.call("ibmdem_sync")

    def poisson(self) -> SolveInfo:
        info = SolveInfo()
        self.call("ibmdem_poisson_solve", C.byref(info))
        return info

    def step(self) -> SolveInfo:
        info = SolveInfo()
        self.call("ibmdem_step", C.byref(info))
        return info

    def run(self, n: int) -> SolveInfo:
        info = SolveInfo()
        self.call("ibmdem_run", int(n), C.byref(info))
        return info

    # -- measurement -------------------------------------------------------
    def profile_kernel(self, kid: int) -> None:
        self.call("ibmdem_profile_kernel", int(kid))

    def profile_read(self):
        t = C.c_double()
        n = C.c_int64()
        b = C.c_double()
        self.call("ibmdem_profile_read", C.byref(t), C.byref(n), C.byref(b))
        return t.value, n.value, b.value

    def memory_bytes(self) -> int:
        n = C.c_int64()
        self.call("ibmdem_memory_bytes", C.byref(n))
        return n.value

    def mg_params(self):
        mg = np.zeros(3, dtype=np.int32)
        self.call("ibmdem_get_mg_params", _p(mg))
        return tuple(int(x) for x in mg)

    def set_mg_params(self, pre: int, post: int, coarse: int) -> None:
        mg = np.array([pre, post, coarse], dtype=np.int32)
        self.call("ibmdem_set_mg_params", _p(mg))

    def stencil_tile(self) -> int:
        t = np.zeros(1, dtype=np.int32)
        self.call("ibmdem_get_stencil_tile", _p(t))
        return int(t[0])

    def set_stencil_tile(self, tile: int) -> None:
        self.call("ibmdem_set_stencil_tile", int(tile))


class _BorrowedSimulation(Simulation):
    """A context owned by a Batch (closing it is the batch's job)."""

    def __init__(self, lib, case, handle):
        self.lib = lib
        self.case = case
        self.shape = (case.nz, case.ny, case.nx)
        self.n_markers = case.n_markers
        self.n_bodies = case.n_bodies
        self.h = handle

    def close(self) -> None:
        self.h = None


class Batch:
    """Independent cases of uneven size advanced together (one context and
    HIP stream each). ``batch[q]`` is case q's context for uploads,
    downloads and single operators."""

    def __init__(self, lib: C.CDLL, cases):
        self.lib = lib
        self.cases = list(cases)
        arr = (Case * len(self.cases))(*self.cases)
        h = C.c_void_p()
        rc = lib.ibmdem_batch_create(len(self.cases), arr, C.byref(h))
        if rc != IBMDEM_OK:
            raise IbmdemError(rc, "ibmdem_batch_create")
        self.h = h
        self.sims = []
        for q, cs in enumerate(self.cases):
            ch = C.c_void_p()
            self._check(lib.ibmdem_batch_context(self.h, q, C.byref(ch)), "ibmdem_batch_context")
            self.sims.append(_BorrowedSimulation(lib, cs, ch))

    def _check(self, rc, where):
        if rc != IBMDEM_OK:
            detail = self.lib.ibmdem_batch_last_error(self.h)
            raise IbmdemError(rc, where, detail.decode() if detail else "")

    def __len__(self):
        return len(self.sims)

    def __getitem__(self, q) -> Simulation:
        return self.sims[q]

    def run(self, n: int):
        infos = (SolveInfo * len(self.sims))()
        self._check(self.lib.ibmdem_batch_run(self.h, int(n), infos), "ibmdem_batch_run")
        return list(infos)

    def autotune(self, trial_steps: int = 3) -> None:
        self._check(self.lib.ibmdem_batch_autotune(self.h, int(trial_steps)), "ibmdem_batch_autotune")

    def set_fused(self, mode) -> None:
        """0/False: each case's step on its own stream; 1/True (the default):
        the whole step batched; 2: only the solve; 3: the solve and the
        post-solve part; 4: the forcing loop, the solve and the post-solve
        part."""
        self._check(self.lib.ibmdem_batch_set_fused(self.h, int(mode)), "ibmdem_batch_set_fused")

    def fused_solves(self) -> int:
        """Group solves that have run batched so far."""
        return self.stats()[0]

    def stats(self):
        """(batched group solves, {solve GPU ms, host ms per phase})"""
        n = C.c_int64()
        ms = (C.c_double * 5)()
        self._check(self.lib.ibmdem_batch_stats(self.h, C.byref(n), ms), "ibmdem_batch_stats")
        keys = ("solve_gpu", "queue_first", "wait_solve", "queue_second", "wait_end")
        return int(n.value), dict(zip(keys, list(ms)))

    BATCH_KINDS = {"smooth0": 1, "smooth": 2, "restrict": 3, "prolong": 4, "bottom": 5, "matvec": 6,
                   "vector": 7, "scalar": 8}

    def profile(self, on: bool = True) -> None:
        """Time every launch of the batched solves (by kind) from now on."""
        self._check(self.lib.ibmdem_batch_profile(self.h, 1 if on else 0), "ibmdem_batch_profile")

    def profile_read(self):
        """{kind: (total ms, launches, total algorithmic bytes)}"""
        out = {}
        for name, k in self.BATCH_KINDS.items():
            ms, n, by = C.c_double(), C.c_int64(), C.c_double()
            self._check(self.lib.ibmdem_batch_profile_read(self.h, k, C.byref(ms), C.byref(n), C.byref(by)),
                        "ibmdem_batch_profile_read")
            out[name] = (ms.value, int(n.value), by.value)
        return out

    def set_solve_groups(self, groups: int) -> None:
        self._check(self.lib.ibmdem_batch_set_solve_groups(self.h, int(groups)), "ibmdem_batch_set_solve_groups")

    def solve_groups(self) -> int:
        g = (C.c_int32 * 1)()
        self._check(self.lib.ibmdem_batch_get_solve_groups(self.h, g), "ibmdem_batch_get_solve_groups")
        return int(g[0])

    def close(self) -> None:
        if getattr(self, "h", None):
            for s in self.sims:
                s.close()
            self.lib.ibmdem_batch_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
