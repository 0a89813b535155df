"""GPU (HIP library, through the C ABI) vs CPU oracle on identical seeded
inputs, operator by operator and for whole coupled steps.

Tolerances: integer outputs (cell ids, contact candidate lists) and the
operators whose arithmetic is order-identical (interpolation, predictor
without IBM force, divergence, contact force, marker update) must match
bit for bit; sums in another fixed order (spreading, body reductions,
inner products) within 1e-14 relative to the field's scale (a few ulps of
each cell's largest term); composite steps within 1e-10 relative (the FP64
target), marker positions within 1e-12.
"""
import numpy as np
import pytest

from ibmdem import cases
from ibmdem.abi import (BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, ERR_NOT_CONVERGED, IbmdemError,
                        Simulation)

pytestmark = pytest.mark.gpu


def _setups():
    return {
        "periodic": lambda: cases.small_periodic(gap=0.75),
        "walls": lambda: cases.small_walls(),
        "channel": lambda: cases.small_walls(periodic_x=True),
        # inflow/outflow along x (spanwise-periodic z) and along z (walled)
        "open-x": lambda: cases.small_open(axis=0, periodic=(2,)),
        "open-z": lambda: cases.small_open(axis=2),
        # odd level-0 extents: an odd periodic x (no colour-split storage,
        # level 0 is the coarsest level, swept by damped Jacobi) and an odd
        # walled y (the matrix-vector product's and the half sweep's last
        # unpaired row)
        "odd-periodic": lambda: cases.small_periodic(n=(25, 21, 20), n_spheres=3, D=6.0, gap=0.75),
        "odd-walls": lambda: cases.small_walls(n=(24, 21, 20)),
    }


def _planes(s):
    return [s.download_outflow(side) for side in (1, 3, 5) if s.case.btype[side] == BC_OUTFLOW]


@pytest.fixture(params=list(_setups()))
def pair(request, gpu_lib, oracle_lib):
    from ibmdem.abi import Simulation
    setup = _setups()[request.param]()
    g = Simulation(gpu_lib, setup.case)
    r = Simulation(oracle_lib, setup.case)
    for s in (g, r):
        cases.load(s, setup)
    yield setup, g, r
    g.close()
    r.close()


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


def test_cells_and_pairs_bit_exact(pair):
    setup, g, r = pair
    np.testing.assert_array_equal(g.cells(), r.cells())
    cg, pg, sg = g.pairs()
    cr, pr, sr = r.pairs()
    np.testing.assert_array_equal(cg, cr)
    assert cr.sum() > 0, "case has no contact candidates"
    for m in range(cg.size):
        n = cg[m]
        np.testing.assert_array_equal(pg[m, :n], pr[m, :n])
        np.testing.assert_array_equal(sg[m, :n], sr[m, :n])


def test_interpolation_bit_exact(pair):
    setup, g, r = pair
    for s in (g, r):
        s.begin_step()
        s.interpolate()
    mg, mr = g.download_markers(), r.download_markers()
    assert np.abs(mr.fdrag).max() > 0
    np.testing.assert_array_equal(mg.fdrag, mr.fdrag)


def test_spread_and_mdf_update(pair):
    setup, g, r = pair
    for s in (g, r):
        s.begin_step()
        for _ in range(2):
            s.interpolate()
            s.spread()
    fg, fr = g.download_force(), r.download_force()
    for a, b in zip(fg, fr):
        assert np.abs(b).max() > 0
        assert rel(a, b) < 1e-14
    ug, ur = g.download_fields(), r.download_fields()
    for a, b in zip(ug[:3], ur[:3]):
        assert rel(a, b) < 1e-14
    # fdrag after the second interpolation saw the spread velocity
    assert rel(g.download_markers().fdrag, r.download_markers().fdrag) < 1e-12


def test_predictor_bit_exact_without_ibm_force(pair):
    setup, g, r = pair
    for s in (g, r):
        s.begin_step()
        s.predictor(1.0, 0.0, 1.0)
    ug, ur = g.download_fields(), r.download_fields()
    for a, b in zip(ug[:3], ur[:3]):
        np.testing.assert_array_equal(a, b)


def test_bulk_divergence_projection(pair):
    setup, g, r = pair
    for s in (g, r):
        s.begin_step()
        s.predictor(1.0, 0.0, 1.0)
        s.bulk()
    ug, ur = g.download_fields(), r.download_fields()
    for a, b in zip(ug[:3], ur[:3]):
        assert rel(a, b) < 1e-14
    for s in (g, r):
        s.divergence(1.0)
    _, rhs_g = g.download_poisson()
    _, rhs_r = r.download_poisson()
    assert rel(rhs_g, rhs_r) < 1e-12
    ig, ir = g.poisson(), r.poisson()
    assert ig.converged and ir.converged
    assert abs(ig.iterations - ir.iterations) <= 1
    pg, _ = g.download_poisson()
    pr, _ = r.download_poisson()
    if not _planes(r):
        # the reference's gauge; an outflow side fixes P' = 0 on its face
        assert pg.flat[0] == 0.0 and pr.flat[0] == 0.0
    assert rel(pg, pr) < 1e-10
    for s in (g, r):
        s.project(1.0)
    ug, ur = g.download_fields(), r.download_fields()
    for a, b in zip(ug, ur):
        assert rel(a, b) < 1e-10
    for a, b in zip(_planes(g), _planes(r)):
        assert rel(a, b) < 1e-10


def test_poisson_same_rhs(pair):
    setup, g, r = pair
    rng = np.random.default_rng(7)
    rhs = rng.normal(size=(setup.case.nz, setup.case.ny, setup.case.nx))
    rhs -= rhs.mean()
    for s in (g, r):
        s.upload_poisson_rhs(rhs)
    ig, ir = g.poisson(), r.poisson()
    assert ig.converged and ir.converged
    assert ig.rel_residual <= setup.case.poisson_rtol
    assert abs(ig.iterations - ir.iterations) <= 1
    pg, _ = g.download_poisson()
    pr, _ = r.download_poisson()
    assert rel(pg, pr) < 1e-9


_P, _W, _I, _O = BC_PERIODIC, BC_WALL, BC_INFLOW, BC_OUTFLOW


@pytest.mark.parametrize("n,bt", [
    ((48, 40, 36), [_P] * 6),
    ((40, 24, 20), [_W] * 6),
    ((64, 16, 48), [_P, _P, _W, _W, _W, _W]),
    ((36, 28, 40), [_W, _W, _W, _W, _I, _O]),
    ((40, 32, 24), [_I, _O, _W, _W, _P, _P]),
], ids=["periodic", "walls", "channel", "open-z", "open-x"])
def test_one_preconditioned_iteration(gpu_lib, oracle_lib, n, bt):
    """One CG iteration from x = 0 is x = (r.z / p.Ap) z with z = one
    V-cycle applied to the right-hand side, so the iterate exposes the
    V-cycle (smoother, residual, restriction, prolongation, coarsest solve)
    up to the rounding of two inner products summed in different orders.
    Level sizes here leave partial coarse tiles in every dimension."""
    case = cases.build_case(*n, bt, rtol=1e-30, max_iter=1)
    rng = np.random.default_rng(5)
    rhs = rng.normal(size=(n[2], n[1], n[0]))
    rhs -= rhs.mean()
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, case) as s:
            s.upload_poisson_rhs(rhs)
            with pytest.raises(IbmdemError) as e:
                s.poisson()
            assert e.value.code == ERR_NOT_CONVERGED
            out.append(s.download_poisson()[0])
    assert np.abs(out[1]).max() > 0
    assert rel(out[0], out[1]) < 1e-12


def test_contact_and_bodies(pair):
    setup, g, r = pair
    for s in (g, r):
        s.contact()
    mg, mr = g.download_markers(), r.download_markers()
    assert np.abs(mr.fcol).max() > 0, "no marker is in contact"
    np.testing.assert_array_equal(mg.fcol, mr.fcol)
    for s in (g, r):
        s.body_assemble()
    bg, br = g.download_bodies(), r.download_bodies()
    for k in ("fcol", "fmomcol", "fdrag", "fmom"):
        a, b = getattr(bg, k), getattr(br, k)
        assert rel(a, b) < 1e-12 or np.abs(a - b).max() < 1e-14, k
    for s in (g, r):
        s.body_integrate()
    bg, br = g.download_bodies(), r.download_bodies()
    for k in ("x", "v", "omega", "angmom", "rot", "rot_glo"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-12, k
    # marker update from identical body states is bit-exact
    for s in (g, r):
        s.upload_bodies(br)
        s.markers_follow()
    mg, mr = g.download_markers(), r.download_markers()
    np.testing.assert_array_equal(mg.x, mr.x)
    np.testing.assert_array_equal(mg.v, mr.v)
    for s in (g, r):
        s.rebuild_lists()
    np.testing.assert_array_equal(g.cells(), r.cells())


def test_full_steps(pair):
    setup, g, r = pair
    for step in range(3):
        ig, ir = g.step(), r.step()
        assert ig.converged and ir.converged
    ug, ur = g.download_fields(), r.download_fields()
    for a, b, name in zip(ug, ur, "uvwp"):
        assert rel(a, b) < 1e-10, name
    for a, b in zip(_planes(g), _planes(r)):
        assert rel(a, b) < 1e-10
    mg, mr = g.download_markers(), r.download_markers()
    assert rel(mg.x, mr.x) < 1e-12
    assert rel(mg.v, mr.v) < 1e-10
    bg, br = g.download_bodies(), r.download_bodies()
    for k in ("x", "v", "omega"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-10, k


def test_gpu_run_to_run_bitwise(gpu_lib):
    from ibmdem.abi import Simulation
    setup = cases.small_periodic(gap=0.75)
    outs = []
    for _ in range(2):
        with Simulation(gpu_lib, setup.case) as g:
            cases.load(g, setup)
            g.run(2)
            outs.append((g.download_fields(), g.download_markers().x))
    for a, b in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def _ops_step(s):
    """One step issued operator by operator (the reference's call order)."""
    s.begin_step()
    for _ in range(s.case.n_mdf):
        s.interpolate()
        s.spread()
    s.predictor(1.0, 0.0, 1.0)
    s.bulk()
    s.divergence(1.0)
    s.poisson()
    s.project(1.0)
    s.cloud_evolution()


@pytest.mark.parametrize("name", ["periodic", "walls", "open-z"])
def test_graph_steps_match_eager_steps(gpu_lib, monkeypatch, name):
    """The whole-step entry point replays captured HIP graphs once the CG
    iteration count is known; every replay (two parities, several times
    each, with host waits between) must equal the eager launches bit for
    bit, and operator-by-operator calls afterwards must see the same state."""
    from ibmdem.abi import Simulation
    setup = _setups()[name]()
    monkeypatch.setenv("IBMDEM_NO_GRAPH", "1")
    e = Simulation(gpu_lib, setup.case)
    monkeypatch.delenv("IBMDEM_NO_GRAPH")
    g = Simulation(gpu_lib, setup.case)
    try:
        for s in (e, g):
            cases.load(s, setup)
        for step in range(6):
            ie, ig = e.step(), g.step()
            assert (ie.iterations, ie.converged, ie.rel_residual) == (ig.iterations, ig.converged, ig.rel_residual)
            for a, b in zip(e.download_fields(), g.download_fields()):
                np.testing.assert_array_equal(a, b)
            me, mg = e.download_markers(), g.download_markers()
            np.testing.assert_array_equal(me.x, mg.x)
            np.testing.assert_array_equal(me.fdrag, mg.fdrag)
            np.testing.assert_array_equal(e.download_bodies().omega, g.download_bodies().omega)
            for a, b in zip(_planes(e), _planes(g)):
                np.testing.assert_array_equal(a, b)
        for s in (e, g):
            _ops_step(s)
        for a, b in zip(e.download_fields(), g.download_fields()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(e.download_force(), g.download_force()):
            np.testing.assert_array_equal(a, b)
    finally:
        e.close()
        g.close()


def test_rk3_steps_match_oracle(gpu_lib, oracle_lib):
    """Contexts for the RK3 scheme: three stages of the reference's own
    VelocityGuess(a,b,c) form (uold, RHSl kept), each followed by divergence,
    solve and projection; GPU against the oracle on a Taylor-Green vortex
    with the IBM force of a marker-free case (zero)."""
    from ibmdem.abi import BC_PERIODIC, SCHEME_RK3, Simulation
    n = 16
    h = 2 * np.pi / n
    case = cases.build_case(n, n, 4, [BC_PERIODIC] * 6, h=h, visc=0.1, dt=0.05, rtol=1e-13, max_iter=300,
                            time_scheme=SCHEME_RK3)
    z, y, x = np.meshgrid(np.arange(4), np.arange(n), np.arange(n), indexing="ij")
    u = np.cos(x * h) * np.sin((y + 0.5) * h)
    v = -np.sin((x + 0.5) * h) * np.cos(y * h)
    p = -0.25 * (np.cos(2 * (x + 0.5) * h) + np.cos(2 * (y + 0.5) * h))
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, case) as s:
            s.upload_fields(u, v, np.zeros_like(u), p)
            # one predictor stage alone is the reference's arithmetic: bitwise
            s.begin_step()
            s.predictor(25.0 / 60.0, -17.0 / 60.0, 8.0 / 60.0)
            first = s.download_fields()
            s.upload_fields(u, v, np.zeros_like(u), p)
            s.run(3)
            out.append((first, s.download_fields()))
    (fg, g), (fr, r) = out
    for a, b in zip(fg[:3], fr[:3]):
        np.testing.assert_array_equal(a, b)
    # w stays at round-off level in this 2-D flow: compare on the velocity scale
    vscale = max(np.abs(q).max() for q in r[:3])
    for a, b in zip(g[:3], r[:3]):
        assert np.abs(a - b).max() < 1e-11 * vscale
    assert rel(g[3], r[3]) < 1e-10


def test_general_coefficients_need_an_rk3_context(gpu_lib):
    from ibmdem.abi import IbmdemError, Simulation
    setup = cases.small_walls()
    with Simulation(gpu_lib, setup.case) as s:
        cases.load(s, setup)
        s.begin_step()
        with pytest.raises(IbmdemError):
            s.predictor(25.0 / 60.0, -17.0 / 60.0, 8.0 / 60.0)


def test_fixed_body_stays_and_matches_oracle(gpu_lib, oracle_lib):
    """A fixed body (the obstacle of the cylinder-in-crossflow case) keeps
    its position and velocity; the free body beside it moves as in the
    oracle, and the fixed body's assembled force is still reported."""
    from ibmdem.abi import Simulation
    setup = cases.small_walls()
    setup.bodies.fixed[0] = 1
    setup.bodies.v[:, 0] = 0.0
    setup.bodies.vold[:, 0] = 0.0
    setup.markers.v[:, setup.markers.body == 0] = 0.0
    out = []
    for lib in (gpu_lib, oracle_lib):
        with Simulation(lib, setup.case) as s:
            cases.load(s, setup)
            s.run(3)
            out.append((s.download_fields(), s.download_bodies(), s.download_markers()))
    (fg, bg, mg), (fr, br, mr) = out
    np.testing.assert_array_equal(bg.x[:, 0], setup.bodies.x[:, 0])
    np.testing.assert_array_equal(bg.v[:, 0], 0.0)
    np.testing.assert_array_equal(mg.x[:, mg.body == 0], setup.markers.x[:, setup.markers.body == 0])
    assert np.abs(bg.fdrag[:, 0]).max() > 0
    for a, b in zip(fg, fr):
        assert rel(a, b) < 1e-10
    for k in ("x", "v", "fdrag"):
        assert rel(getattr(bg, k), getattr(br, k)) < 1e-10, k


@pytest.mark.parametrize("name", ["periodic", "walls", "open-x"])
def test_predictor_schedules_agree_bitwise(gpu_lib, name):
    """Every predictor schedule (case.stencil_tile: LDS z-marching tiles of
    64x4 over 32 or 16 planes, 64x8 over 32 planes, the direct kernel) gives
    the same fields bit for bit, bulk forcing and open sides included."""
    from ibmdem.abi import Simulation
    out = []
    for tile in (3, 0, 1, 2):
        setup = _setups()[name]()
        setup.case.stencil_tile = tile
        with Simulation(gpu_lib, setup.case) as s:
            cases.load(s, setup)
            s.begin_step()
            s.predictor(1.0, 0.0, 1.0)
            first = s.download_fields()
            s.run(2)
            out.append((first, s.download_fields()))
    for first, after in out[1:]:
        for a, b in zip(first, out[0][0]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(after, out[0][1]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name,precision", [("periodic", 0), ("periodic32", 0), ("periodic32", 1), ("walls", 0),
                                            ("open-z", 0), ("walls", 1), ("odd", 0)])
def test_small_level_work_group_matches_launches(gpu_lib, monkeypatch, name, precision):
    """The small multigrid levels run as one work-group in LDS (one launch
    per V-cycle); the result equals the one-launch-per-operation path bit
    for bit, in both precisions; a hierarchy whose coarsest level has an odd
    periodic dimension keeps the launches."""
    def setup():
        if name == "odd":
            return cases.small_periodic(n=(36, 24, 20), n_spheres=3)
        if name == "periodic32":   # levels 32^3 ... 2^3: the work-group takes 16^3 and below
            return cases.small_periodic(n=(32, 32, 32), n_spheres=4)
        return _setups()[name]()
    outs = []
    for no_tail in ("1", "0"):
        monkeypatch.setenv("IBMDEM_NO_TAIL", no_tail)
        st = setup()
        st.case.precision = precision
        with Simulation(gpu_lib, st.case) as s:
            cases.load(s, st)
            infos = [s.step() for _ in range(3)]
            outs.append((s.download_fields(), [(i.iterations, i.rel_residual) for i in infos]))
    monkeypatch.delenv("IBMDEM_NO_TAIL")
    assert outs[0][1] == outs[1][1]
    for a, b in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, b)
