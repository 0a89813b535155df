"""Inflow and outflow sides on the GPU (through the C ABI): the closed forms
the oracle is pinned by hold on the GPU too, at sizes beyond the oracle's,
and the RK3 and mixed-precision forms agree with their references.

Operator-by-operator and whole-step parity with the oracle for open cases
runs in test_gpu_parity.py (its "open-x" and "open-z" setups)."""
import numpy as np
import pytest

from ibmdem import cases, snapshot
from ibmdem.abi import (BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, FP64, MIXED, SCHEME_RK3, Simulation)

pytestmark = pytest.mark.gpu


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


def _div(fields, plane, axis, bt):
    d = 0.0
    for c in range(3):
        a = fields[c]
        na = 2 - c
        if bt[2 * c + 1] == BC_OUTFLOW:
            hi = np.expand_dims(plane, na)
        elif bt[2 * c + 1] == BC_PERIODIC:
            hi = np.take(a, [0], axis=na)
        else:
            hi = np.zeros_like(np.take(a, [0], axis=na))
        d = d + (np.concatenate([np.take(a, range(1, a.shape[na]), axis=na), hi], axis=na) - a)
    return d


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_stream_is_steady(gpu_lib, axis, precision):
    n = (64, 48, 40)
    bt = [BC_PERIODIC] * 6
    bt[2 * axis], bt[2 * axis + 1] = BC_INFLOW, BC_OUTFLOW
    U = 0.07
    inflow = [0.0] * 3
    inflow[axis] = U
    case = cases.build_case(*n, bt, visc=0.05, dt=0.2, inflow=inflow, precision=precision)
    u, v, w, p = cases.zero_fields(case)
    [u, v, w][axis][:] = U
    with Simulation(gpu_lib, case) as s:
        s.upload_fields(u, v, w, p)
        for _ in range(4):
            assert s.step().iterations == 0
        got = s.download_fields()
        plane = s.download_outflow(2 * axis + 1)
    for c in range(3):
        np.testing.assert_array_equal(got[c], U if c == axis else 0.0)
    np.testing.assert_array_equal(got[3], 0.0)
    np.testing.assert_array_equal(plane, U)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_divergence_free_and_mass_balance_at_scale(gpu_lib, precision):
    """A 128x96x192 bed-like channel (inflow at the bottom, outflow at the
    top, walled sides) with three spheres: after each step every cell,
    the last layer included, is divergence-free to the solver tolerance and
    the outflow flux equals the inflow flux."""
    st = cases.small_open(n=(128, 96, 192), axis=2, precision=precision, rtol=1e-10)
    bt = list(st.case.btype)
    U = st.case.inflow[2]
    with Simulation(gpu_lib, st.case) as s:
        cases.load(s, st)
        for _ in range(3):
            info = s.step()
            assert info.converged
            u, v, w, p = s.download_fields()
            plane = s.download_outflow(5)
            div = _div((u, v, w), plane, 2, bt)
            assert np.abs(div).max() < 1e-8 * U
            np.testing.assert_array_equal(w[0], U)
            assert abs(plane.sum() - U * plane.size) < 1e-9 * U * plane.size


def test_open_rk3_matches_oracle(gpu_lib, oracle_lib):
    st = cases.small_open(axis=0, periodic=(2,), time_scheme=SCHEME_RK3, spheres=False, rtol=1e-13)
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, st.case) as s:
            cases.load(s, st)
            s.run(3)
            out.append((s.download_fields(), s.download_outflow(1)))
    (g, pg), (r, pr) = out
    vscale = max(np.abs(q).max() for q in r[:3])
    for a, b in zip(g[:3], r[:3]):
        assert np.abs(a - b).max() < 1e-11 * vscale
    assert rel(g[3], r[3]) < 1e-10
    assert np.abs(pg - pr).max() < 1e-11 * vscale


@pytest.mark.parametrize("name", ["walls", "open-z"])
def test_mixed_precision_error_against_fp64(gpu_lib, name):
    """The mixed form (FP32 predictor and V-cycle inside FP64 CG) against the
    FP64 form after 5 steps: trajectory error <= 1e-6 relative and the
    projected field divergence-free to <= 1e-8 of the velocity scale per
    cell; velocity and pressure errors printed for the record."""
    make = {"walls": lambda p: cases.small_walls(precision=p, rtol=1e-10),
            "open-z": lambda p: cases.small_open(axis=2, precision=p, rtol=1e-10)}[name]
    out = {}
    for prec in (FP64, MIXED):
        st = make(prec)
        with Simulation(gpu_lib, st.case) as s:
            cases.load(s, st)
            s.run(5)
            planes = [s.download_outflow(sd) for sd in (1, 3, 5) if st.case.btype[sd] == BC_OUTFLOW]
            out[prec] = (s.download_fields(), s.download_bodies(), planes, list(st.case.btype))
    (f64, b64, p64, bt), (f32, b32, p32, _) = out[FP64], out[MIXED]
    vscale = max(np.abs(q).max() for q in f64[:3])
    uerr = max(np.abs(a - b).max() for a, b in zip(f32[:3], f64[:3])) / vscale
    perr = rel(f32[3], f64[3])
    xerr = np.abs(b32.x - b64.x).max() / np.abs(b64.x).max()
    div = _div(f32[:3], p32[0] if p32 else None, 2, bt)
    print(f"{name}: velocity {uerr:.2e} pressure {perr:.2e} position {xerr:.2e} "
          f"divergence {np.abs(div).max() / vscale:.2e}")
    assert xerr <= 1e-6
    assert np.abs(div).max() <= 1e-8 * vscale
    assert uerr <= 1e-5


def test_snapshot_resume_carries_the_outflow_plane(gpu_lib, tmp_path):
    st = cases.small_open(axis=2)
    with Simulation(gpu_lib, st.case) as s:
        cases.load(s, st)
        s.run(2)
        want = (s.download_fields(), s.download_outflow(5), s.download_bodies().x)
    with Simulation(gpu_lib, st.case) as s:
        cases.load(s, st)
        info = s.step()
        p = str(tmp_path / "mid.snap")
        snapshot.write_simulation(p, s, info)
    with Simulation(gpu_lib, st.case) as s:
        snapshot.load(s, p)
        s.step()
        got = (s.download_fields(), s.download_outflow(5), s.download_bodies().x)
    for a, b in zip(got[0], want[0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[2], want[2])
