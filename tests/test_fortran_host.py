"""The Fortran host side (ibm-dem_amd/fortran): the ISO_C_BINDING layer, the
driver module with the reference's subroutine names and the time-loop
program.

CPU tests link the same Fortran sources against the CPU oracle library
(which exports the identical C ABI) and check that the reference-order
composition FirstOrderForwardEuler() (fluid.F90:252-327), issued subroutine
by subroutine from Fortran, reproduces the library's whole-step entry point
bit for bit; the bind(C) types have the C layout; snapshots round-trip.
The GPU test runs the product program (linked against the HIP library)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from ibmdem import abi, cases, snapshot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "ibm-dem_amd", "fortran")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
PRODUCT = os.path.join(FDIR, "_build", "ibmdem_run")

needs_flang = pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")


@pytest.fixture(scope="module")
def oracle_program(tmp_path_factory):
    """ibmdem_run linked against the CPU oracle (test build, in a tmp dir)."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not present")
    libdir = os.path.dirname(oracle.build())
    out = tmp_path_factory.mktemp("fortran")
    binp = str(out / "ibmdem_run_oracle")
    subprocess.run(["make", "-s", "-C", FDIR, f"OUT={out}/mod", f"BIN={binp}",
                    f"LIBDIR={libdir}", "LIBNAME=ibmdem_oracle", f"RPATH={libdir}"],
                   check=True)
    return binp


def _run(program, setup, tmp_path, nsteps, mode):
    src = str(tmp_path / "in.snap")
    dst = str(tmp_path / f"out_{mode}.snap")
    snapshot.write(src, setup.case, setup.markers, setup.bodies, setup.fields)
    r = subprocess.run([program, src, dst, str(nsteps), mode],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert "ibmdem_run:" in r.stdout
    return snapshot.read(dst)


def _python_run(lib, setup, nsteps):
    sim = abi.Simulation(lib, setup.case)
    cases.load(sim, setup)
    info = sim.run(nsteps)
    out = (sim.download_markers(), sim.download_bodies(), sim.download_fields(), info)
    sim.close()
    return out


def _assert_same_state(got, want):
    _, mk2, bd2, f2, info2 = got
    mk, bd, f, info = want
    for a, b in zip(f2, f):
        np.testing.assert_array_equal(a, b)
    for k in ("x", "v", "fdrag", "fcol"):
        np.testing.assert_array_equal(getattr(mk2, k), getattr(mk, k))
    for k in abi.BODY_VEC3 + abi.BODY_MAT:
        np.testing.assert_array_equal(getattr(bd2, k), getattr(bd, k))
    assert info2.iterations == info.iterations
    assert info2.rel_residual == info.rel_residual


@needs_flang
def test_bind_c_sizes_match_c(oracle_program):
    import ctypes as C
    r = subprocess.run([oracle_program, "--sizes"], capture_output=True, text=True, check=True)
    sizes = [int(x) for x in r.stdout.split()[:4]]
    assert sizes == [C.sizeof(abi.Case), C.sizeof(abi.Markers),
                     C.sizeof(abi.Bodies), C.sizeof(abi.SolveInfo)]


def _named(name):
    return {"periodic": lambda: cases.small_periodic(gap=0.75),
            "walls": lambda: cases.small_walls(),
            "open": lambda: cases.small_open(axis=2)}[name]()


@needs_flang
@pytest.mark.parametrize("name", ["periodic", "walls", "open"])
@pytest.mark.parametrize("mode", ["ops", "step"])
def test_fortran_driver_matches_whole_step(oracle_program, tmp_path, name, mode):
    setup = _named(name)
    got = _run(oracle_program, setup, tmp_path, 2, mode)
    want = _python_run(oracle.library(), setup, 2)
    _assert_same_state(got, want)


@needs_flang
def test_fortran_rejects_foreign_file(oracle_program, tmp_path):
    bad = tmp_path / "bad.snap"
    bad.write_bytes(b"NOTASNAPxxxx")
    r = subprocess.run([oracle_program, str(bad), str(tmp_path / "o.snap"), "1"],
                       capture_output=True, text=True)
    assert r.returncode != 0
    assert "not an ibmdem snapshot" in r.stderr


@needs_flang
def test_fortran_reports_library_errors(oracle_program, tmp_path):
    """A case the library rejects stops the program with the library's status
    and message (the reference would run on with garbage)."""
    setup = cases.small_periodic(gap=0.75)
    setup.case.btype[1] = abi.BC_WALL     # x- periodic, x+ wall: invalid
    src = str(tmp_path / "in.snap")
    snapshot.write(src, setup.case, setup.markers, setup.bodies, setup.fields)
    r = subprocess.run([oracle_program, src, str(tmp_path / "o.snap"), "1"],
                       capture_output=True, text=True)
    assert r.returncode == 2          # IBMDEM_ERR_ARG
    assert "ibmdem_create" in r.stderr


def test_snapshot_round_trip(tmp_path):
    setup = cases.small_walls()
    p = str(tmp_path / "s.snap")
    info = abi.SolveInfo(7, 1, 1.5e-11, 2.0)
    snapshot.write(p, setup.case, setup.markers, setup.bodies, setup.fields, info)
    case, mk, bd, f, info2 = snapshot.read(p)
    assert bytes(case) == bytes(setup.case)
    for a, b in zip(f, setup.fields):
        np.testing.assert_array_equal(a, b)
    for k in mk.__dataclass_fields__:
        np.testing.assert_array_equal(getattr(mk, k), getattr(setup.markers, k))
    for k in bd.__dataclass_fields__:
        np.testing.assert_array_equal(getattr(bd, k), getattr(setup.bodies, k))
    assert (info2.iterations, info2.converged, info2.rel_residual, info2.rhs_norm) == (7, 1, 1.5e-11, 2.0)


@pytest.mark.parametrize("name", ["periodic", "open"])
def test_snapshot_resume_is_bitwise(tmp_path, oracle_lib, name):
    """2 steps == 1 step, snapshot, restore in a fresh context, 1 step (an
    open case needs its outflow plane from the snapshot)."""
    setup = _named(name)
    want = _python_run(oracle_lib, setup, 2)
    sim = abi.Simulation(oracle_lib, setup.case)
    cases.load(sim, setup)
    info = sim.step()
    p = str(tmp_path / "mid.snap")
    snapshot.write_simulation(p, sim, info)
    sim.close()
    case = snapshot.read(p)[0]
    sim = abi.Simulation(oracle_lib, case)
    snapshot.load(sim, p)
    info = sim.step()
    got = (None, sim.download_markers(), sim.download_bodies(), sim.download_fields(), info)
    sim.close()
    _assert_same_state(got, want)


@pytest.mark.gpu
def test_fortran_program_on_gpu(tmp_path, gpu_lib, oracle_lib):
    """The product program (Fortran time loop over the HIP library): same
    state as the Python host side on the GPU, bit for bit, and within the
    FP64 tolerance of the oracle."""
    if not os.path.exists(PRODUCT):
        pytest.skip("Fortran program not built (flang absent at build time)")
    assert shutil.which(PRODUCT) or os.access(PRODUCT, os.X_OK)
    setup = cases.small_walls()
    got = _run(PRODUCT, setup, tmp_path, 3, "ops")
    _assert_same_state(got, _python_run(gpu_lib, setup, 3))
    mk, bd, f, _ = _python_run(oracle_lib, setup, 3)
    for a, b in zip(got[3], f):
        assert np.abs(a - b).max() <= 1e-10 * max(np.abs(b).max(), 1e-300)
    assert np.abs(got[2].x - bd.x).max() <= 1e-10 * np.abs(bd.x).max()
