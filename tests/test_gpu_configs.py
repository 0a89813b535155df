"""The five benchmark configurations at full size on the GPU.

C1  sphere settling, 64x64x256, walls, FP64: GPU against the oracle for 10
    steps (cells bit-exact; fields and body state within 1e-10), and the
    1000-step body trajectory against the oracle's committed run
    (tests/golden/settling_trajectory.npz).
C2  periodic suspension, 256^3, 2000 spheres, 904 k markers, FP64: two
    steps GPU against the oracle (the first step moves only the bodies, the
    second spreads their forcing into the fluid); the spread's partition of
    unity, run-to-run identity and the projected divergence at full size.
C3  fluidized bed, 192x192x768, 10,000 spheres, inflow/outflow: the
    projected field is divergence-free in every cell, the outflow flux
    equals the inflow flux, and the mixed form stays within 1e-6 of FP64.
C4  cylinder in crossflow, 512x256x64: the same properties, and the
    cylinder's drag and lift coefficients in the mixed form against FP64.
C5  64 uneven settling cases: the batch equals the cases stepped alone,
    bit for bit.
Tolerances are written at each check; the FP64 target is 1e-10 relative.
"""
import os

import numpy as np
import pytest

import oracle
from ibmdem import abi, cases, diagnostics
from ibmdem.abi import BC_OUTFLOW, FP64, MIXED, Simulation

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "settling_trajectory.npz")


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


def divergence(s):
    """Cell divergence of the state of s (ibmdem.diagnostics.divergence, the
    outflow planes included)."""
    planes = {sd: s.download_outflow(sd) for sd in (1, 3, 5) if s.case.btype[sd] == BC_OUTFLOW}
    return diagnostics.divergence(s.case, s.download_fields()[:3], planes)


# ---------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------
def test_c1_settling_against_the_oracle(gpu_lib, oracle_lib):
    setup = cases.config_settling()
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, setup.case) as s:
            cases.load(s, setup)
            its = [s.step().iterations for _ in range(10)]
            out.append((s.cells(), s.download_fields(), s.download_bodies(), s.download_markers(), its))
    (cg, fg, bg, mg, ig), (cr, fr, br, mr, ir) = out
    np.testing.assert_array_equal(cg, cr)
    assert max(abs(a - b) for a, b in zip(ig, ir)) <= 1
    for a, b, name in zip(fg, fr, "uvwp"):
        assert rel(a, b) < 1e-10, name
    assert np.abs(fr[2]).max() > 1e-4   # the sphere drives a flow
    for k in ("x", "v", "omega"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-10, k
    assert rel(bg.x - setup.bodies.x, br.x - setup.bodies.x) < 1e-10   # the displacement itself
    assert rel(mg.x, mr.x) < 1e-12


def test_c1_settling_1000_step_trajectory(gpu_lib):
    """The GPU run of the full 1000 steps against the oracle's trajectory:
    body position (displacement from the start), velocity and the fluid's
    kinetic energy at every step."""
    gold = np.load(GOLDEN)
    steps = int(gold["steps"])
    setup = cases.config_settling()
    x0 = setup.bodies.x[:, 0].copy()
    xs, vs, ke = [], [], []
    with Simulation(gpu_lib, setup.case) as s:
        cases.load(s, setup)
        for _ in range(steps):
            s.step()
            b = s.download_bodies()
            u, v, w, _ = s.download_fields()
            xs.append(b.x[:, 0].copy())
            vs.append(b.v[:, 0].copy())
            ke.append(float((u * u).sum() + (v * v).sum() + (w * w).sum()))
    xs, vs, ke = np.array(xs), np.array(vs), np.array(ke)
    dx_err = np.abs((xs - x0) - (gold["x"] - x0)).max() / np.abs(gold["x"] - x0).max()
    v_err = np.abs(vs - gold["v"]).max() / np.abs(gold["v"]).max()
    ke_err = np.abs(ke - gold["ke"]).max() / np.abs(gold["ke"]).max()
    print(f"C1 1000 steps: displacement {dx_err:.3e}, velocity {v_err:.3e}, kinetic energy {ke_err:.3e}")
    assert np.abs(gold["x"][-1, 2] - x0[2]) > 50.0   # it settled a long way
    # the same algorithm and CG iteration counts: only summation orders
    # differ, so the trajectories agree far inside the 1e-10 FP64 target
    assert dx_err < 1e-12
    assert v_err < 1e-12
    assert ke_err < 1e-12


# ---------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2_setup():
    return cases.config_suspension()


def test_c2_suspension_two_steps_against_the_oracle(gpu_lib, oracle_lib, c2_setup):
    setup = c2_setup
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, setup.case) as s:
            cases.load(s, setup)
            its = [s.step().iterations for _ in range(2)]
            cnt, _, _ = s.pairs(cap=8)
            out.append((s.cells(), cnt, s.download_fields(), s.download_bodies(), s.download_markers(), its))
    (cg, ng, fg, bg, mg, ig), (cr, nr, fr, br, mr, ir) = out
    np.testing.assert_array_equal(cg, cr)
    np.testing.assert_array_equal(ng, nr)
    assert abs(ig[1] - ir[1]) <= 1 and ir[1] > 0
    for a, b, name in zip(fg, fr, "uvwp"):
        assert np.abs(b).max() > 0, name
        assert rel(a, b) < 1e-10, name
    for k in ("x", "v", "omega"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-10, k
    assert rel(mg.x, mr.x) < 1e-12
    assert rel(mg.fdrag, mr.fdrag) < 1e-10


def test_c2_suspension_full_size_properties(gpu_lib, c2_setup):
    """Partition of unity of the spread (sum f dV = -sum fdrag vol per
    component), run-to-run identity, and the projected divergence: the
    projection leaves div u = (dt / rho_f) r, so ||div u|| / ||div u*|| is
    the solver's relative residual (<= 1e-10, 1e-9 allowed for rounding)."""
    setup = c2_setup
    runs = []
    for _ in range(2):
        with Simulation(gpu_lib, setup.case) as s:
            cases.load(s, setup)
            s.run(2)
            # one more step by operators, stopping after the spread
            s.begin_step()
            s.interpolate()
            s.spread()
            f = s.download_force()
            m = s.download_markers()
            for _ in range(setup.case.n_mdf - 1):
                s.interpolate()
                s.spread()
            s.predictor()
            s.bulk()
            s.divergence()
            _, rhs = s.download_poisson()
            s.poisson()
            s.project()
            div = divergence(s)
            runs.append((s.download_fields(), f, m.fdrag, rhs, div))
    (fa, f, fdrag, rhs, div), (fb, *_) = runs
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)
    dV = setup.case.dx * setup.case.dy * setup.case.dz
    fv = fdrag * m.vol[None, :]
    for d in range(3):
        total = np.abs(fv[d]).sum()
        assert total > 0
        assert abs(f[d].sum() * dV + fv[d].sum()) <= 1e-12 * total, d
    div_star = -rhs * setup.case.dt / setup.case.rhof
    ratio = np.linalg.norm(div) / np.linalg.norm(div_star)
    print(f"C2 projected divergence ratio {ratio:.3e}")
    assert ratio < 1e-9


# ---------------------------------------------------------------------------
# C3, C4
# ---------------------------------------------------------------------------
def _open_case_properties(gpu_lib, make, steps):
    """FP64 and mixed runs from the same state: per-cell divergence, flux
    balance, and the mixed form's error against FP64."""
    out = {}
    base = make()
    for prec in (FP64, MIXED):
        setup = cases.Setup(base.name, abi.Case.from_buffer_copy(base.case), base.markers, base.bodies, base.fields,
                            base.meta)
        setup.case.precision = prec
        with Simulation(gpu_lib, setup.case) as s:
            cases.load(s, setup)
            s.run(steps)
            div = divergence(s)
            u, v, w, p = s.download_fields()
            out[prec] = dict(setup=setup, div=div, fields=(u, v, w, p), bodies=s.download_bodies(),
                             planes={sd: s.download_outflow(sd) for sd in (1, 3, 5)
                                     if setup.case.btype[sd] == BC_OUTFLOW})
    return out


def _check_open(out, inflow_axis, U):
    for prec, r in out.items():
        setup = r["setup"]
        vel = max(np.abs(f).max() for f in r["fields"][:3])
        h = setup.case.dx
        # the projected field is divergence-free in every cell
        assert np.abs(r["div"]).max() * h <= 1e-8 * vel, prec
        # outflow flux = inflow flux
        plane = r["planes"][2 * inflow_axis + 1]
        inflow_face = [slice(None)] * 3
        inflow_face[2 - inflow_axis] = 0
        q_in = r["fields"][inflow_axis][tuple(inflow_face)].sum()
        assert q_in == pytest.approx(U * plane.size, rel=1e-14)
        assert abs(plane.sum() - q_in) <= 1e-10 * abs(q_in), prec
    a, b = out[MIXED], out[FP64]
    errs = {name: rel(x, y) for name, x, y in zip("uvwp", a["fields"], b["fields"])}
    x0 = b["setup"].bodies.x
    disp = np.abs(b["bodies"].x - x0).max()
    errs["body_x"] = rel(a["bodies"].x, b["bodies"].x)
    errs["body_disp"] = np.abs((a["bodies"].x - x0) - (b["bodies"].x - x0)).max() / max(disp, 1e-300)
    print("mixed vs FP64:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k in "uvw":
        assert errs[k] < 1e-6, k
    assert errs["p"] < 1e-6
    assert errs["body_x"] < 1e-6
    assert errs["body_disp"] < 1e-6
    return errs


def test_c3_fluidized_bed_full_size(gpu_lib):
    out = _open_case_properties(gpu_lib, cases.config_fluidized_bed, steps=3)
    _check_open(out, 2, out[FP64]["setup"].meta["u_in"])


def test_c4_cylinder_full_size(gpu_lib):
    out = _open_case_properties(gpu_lib, cases.config_cylinder, steps=5)
    _check_open(out, 0, 1.0)
    coef = {}
    for prec, r in out.items():
        m = r["setup"].meta
        coef[prec] = diagnostics.force_coefficients(r["bodies"], m["cylinder_body"], rho=r["setup"].case.rhof,
                                                    U=m["U"], D=m["D"], L=m["span"])
    cd64, cl64 = coef[FP64]
    cd32, cl32 = coef[MIXED]
    assert cd64 > 0
    assert abs(cd32 - cd64) <= 1e-6 * abs(cd64)
    assert abs(cl32 - cl64) <= 1e-6 * max(abs(cd64), 1e-12)


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
def test_c5_batch_of_64_equals_single_cases(gpu_lib):
    setups = cases.config_batch()
    assert len(setups) == 64
    steps = 3
    with abi.Batch(gpu_lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
        b.run(steps)
        got = [(b[q].download_fields(), b[q].download_bodies().x) for q in range(len(b))]
    for s, g in zip(setups, got):
        with Simulation(gpu_lib, s.case) as one:
            cases.load(one, s)
            one.run(steps)
            want = (one.download_fields(), one.download_bodies().x)
        for a, c in zip(g[0], want[0]):
            np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(g[1], want[1])


# ---------------------------------------------------------------------------
# The reference's own case: 86 tilted cylinders, 240 x 32 x 240, SI units
# ---------------------------------------------------------------------------
def test_reference_cylinders_against_the_oracle(gpu_lib, oracle_lib):
    """The case the reference ships (InitCylinderParcelsDEM with
    CYLINDER_LP1125.dat; walls in y, periodic in x and z, bulk forcing to
    0.1 m/s), in its SI units, so every kernel takes its non-power-of-two
    spacing path: three steps GPU against the oracle."""
    t = np.load(os.path.join(os.path.dirname(GOLDEN), "cylinder_lp1125.npy"))
    setup = cases.config_reference_cylinders(t)
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, setup.case) as s:
            cases.load(s, setup)
            its = [s.step().iterations for _ in range(3)]
            out.append((s.cells(), s.download_fields(), s.download_bodies(), s.download_markers(), its))
    (cg, fg, bg, mg, ig), (cr, fr, br, mr, ir) = out
    np.testing.assert_array_equal(cg, cr)
    assert ig == ir
    for a, b, name in zip(fg, fr, "uvwp"):
        assert np.abs(b).max() > 0, name
        assert rel(a, b) < 1e-10, name
    for k in ("x", "v", "omega", "fdrag"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-10, k
    assert rel(mg.x, mr.x) < 1e-12
    assert rel(mg.fdrag, mr.fdrag) < 1e-10
