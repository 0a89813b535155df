"""The cell bins are rebuilt between DEM sub-steps only when the next
sub-step's contact scan reads them (some body has another within its
culling distance); the last sub-step of a call always bins. These tests
start two spheres outside that distance and drive them together, so the
culling switches on in the middle of a call, and check that the gated
library equals one that bins after every sub-step (IBMDEM_NO_BIN_GATE=1,
read at creation) bit for bit, and the oracle (which always bins) to the
FP64 tolerances of the other parity tests."""
import numpy as np
import pytest

from ibmdem import cases
from ibmdem.abi import Simulation

pytestmark = pytest.mark.gpu


def _approaching(distance=9.4, speed=0.15, fluid=False):
    """Two D = 8 spheres (marker shells of radius 3.7h, so the bodies come
    within culling reach at 3.7 + 3.7 + 1.65 = 9.05h) `distance` apart
    along x across the periodic face, closing at 2 * speed."""
    setup = cases.small_periodic(n_spheres=2, contact=False, velocities=False)
    mk, bd = setup.markers, setup.bodies
    L = np.array([setup.case.nx, setup.case.ny, setup.case.nz], dtype=np.float64)
    shift = bd.x[0, 0] + distance - L[0] - bd.x[0, 1]
    bd.x[0, 1] += shift
    sel = mk.body == 1
    mk.x[0, sel] = np.mod(mk.x[0, sel] + shift, L[0])
    bd.v[:] = 0.0
    bd.v[0, 0], bd.v[0, 1] = speed, -speed
    bd.vold[:] = bd.v
    mk.v[:] = bd.v[:, mk.body]
    if not fluid:
        for f in setup.fields:
            f[:] = 0.0
    return setup


def _approaching_walls(distance=9.4, speed=0.15):
    """The closed box's two D = 8 spheres, the lower one at rest on the
    bottom wall (its lowest markers in wall contact throughout), the upper
    one moved up to `distance` above it and falling at `speed`."""
    setup = cases.small_walls()
    mk, bd = setup.markers, setup.bodies
    d = bd.x[:, 1] - bd.x[:, 0]
    lift = np.sqrt(distance ** 2 - d[0] ** 2 - d[1] ** 2) - d[2]
    bd.x[2, 1] += lift
    mk.x[2, mk.body == 1] += lift
    bd.v[:] = 0.0
    bd.v[2, 1] = -speed
    bd.vold[:] = bd.v
    bd.omega[:] = 0.0
    bd.angmom[:] = 0.0
    bd.angmom_old[:] = 0.0
    mk.v[:] = bd.v[:, mk.body]
    for f in setup.fields:
        f[:] = 0.0
    return setup


def _pair(gpu_lib, monkeypatch, setup):
    monkeypatch.setenv("IBMDEM_NO_BIN_GATE", "1")
    ref = Simulation(gpu_lib, setup.case)
    monkeypatch.delenv("IBMDEM_NO_BIN_GATE")
    gated = Simulation(gpu_lib, setup.case)
    for s in (ref, gated):
        cases.load(s, setup)
    return gated, ref


def _same(a, b):
    ma, mb = a.download_markers(), b.download_markers()
    for k in ("x", "v", "fcol", "fdrag"):
        np.testing.assert_array_equal(getattr(ma, k), getattr(mb, k), err_msg=k)
    ba, bb = a.download_bodies(), b.download_bodies()
    for k in ("x", "v", "omega", "fcol"):
        np.testing.assert_array_equal(getattr(ba, k), getattr(bb, k), err_msg=k)
    np.testing.assert_array_equal(a.cells(), b.cells())
    return ma


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


@pytest.mark.parametrize("walls", [False, True])
def test_sub_steps_gated_equal_always_binned(gpu_lib, oracle_lib, monkeypatch, walls):
    setup = _approaching_walls() if walls else _approaching()
    g, ref = _pair(gpu_lib, monkeypatch, setup)
    orc = Simulation(oracle_lib, setup.case)
    cases.load(orc, setup)
    try:
        touched = []
        for call in range(8):
            for s in (g, ref, orc):
                s.cloud_evolution()
            m = _same(g, ref)
            touched.append(bool(np.abs(m.fcol).any()))
            mo, bo, bg = orc.download_markers(), orc.download_bodies(), g.download_bodies()
            assert rel(m.x, mo.x) < 1e-12
            assert rel(m.v, mo.v) < 1e-10
            for k in ("x", "v"):
                assert rel(getattr(bg, k), getattr(bo, k)) < 1e-10, k
        # periodic: no contact in the first call (the bodies start out of
        # reach), and contact forces once they meet (the oracle: calls 3 and
        # 4, after which they part again and leave culling reach in call 6);
        # walls: wall contact from the start, pair contact once they meet
        assert any(touched), touched
        if not walls:
            assert not touched[0], touched
        for a, b in zip(g.pairs(), ref.pairs()):
            np.testing.assert_array_equal(a, b)
    finally:
        for s in (g, ref, orc):
            s.close()


def test_steps_gated_equal_always_binned(gpu_lib, monkeypatch):
    """Whole coupled steps (captured graphs after the first) with fluid: the
    bodies come within culling reach in the third step and touch in the
    eighth (oracle)."""
    setup = _approaching(distance=9.3, speed=0.12, fluid=True)
    g, ref = _pair(gpu_lib, monkeypatch, setup)
    try:
        for step in range(8):
            ig, ir = g.step(), ref.step()
            assert (ig.iterations, ig.rel_residual) == (ir.iterations, ir.rel_residual)
            _same(g, ref)
            for a, b in zip(g.download_fields(), ref.download_fields()):
                np.testing.assert_array_equal(a, b)
    finally:
        g.close()
        ref.close()


def test_batch_gated_equal_always_binned(gpu_lib, monkeypatch):
    """The batched DEM sub-steps gate each case's rebuild by its own culling
    flag: a batch of the approaching spheres (their gate opens in the third
    step) and a cluster in contact from the start (its gate always open)
    equals each case stepped alone by a library that bins after every
    sub-step."""
    from ibmdem.abi import Batch
    setups = [_approaching(distance=9.3, speed=0.12, fluid=True), cases.small_periodic(gap=0.75)]
    with Batch(gpu_lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
        b.run(8)
        got = [(b[q].download_fields(), b[q].download_markers(), b[q].download_bodies()) for q in range(len(b))]
    for s, (fields, mk, bd) in zip(setups, got):
        monkeypatch.setenv("IBMDEM_NO_BIN_GATE", "1")
        ref = Simulation(gpu_lib, s.case)
        monkeypatch.delenv("IBMDEM_NO_BIN_GATE")
        try:
            cases.load(ref, s)
            ref.run(8)
            for a, r in zip(fields, ref.download_fields()):
                np.testing.assert_array_equal(a, r)
            mr, br = ref.download_markers(), ref.download_bodies()
            for k in ("x", "v", "fcol"):
                np.testing.assert_array_equal(getattr(mk, k), getattr(mr, k), err_msg=k)
            for k in ("x", "v", "omega"):
                np.testing.assert_array_equal(getattr(bd, k), getattr(br, k), err_msg=k)
        finally:
            ref.close()
