"""Diagnostics: force coefficients, probe lines, Tecplot field dumps."""
import numpy as np
import pytest

from ibmdem import abi, cases, diagnostics


def _bodies():
    b = abi.BodyArrays.empty(2)
    b.fdrag[:, 1] = [3.0, -1.5, 0.25]
    b.mass[:] = [2.0, 4.0]
    b.x[:, 1] = [1.0, 2.0, 3.0]
    b.v[:, 1] = [0.1, 0.2, 0.3]
    b.omega[:, 1] = [0.0, 0.0, 0.5]
    b.fmom[:, 1] = [0.0, 1.0, 0.0]
    b.fcol[:, 1] = [4.0, 0.0, 0.0]
    return b


def test_force_coefficients_closed_form():
    b = _bodies()
    cd, cl = diagnostics.force_coefficients(b, 1, rho=1.0, U=2.0, D=32.0, L=64.0)
    q = 0.5 * 1.0 * 4.0 * 32.0 * 64.0
    assert cd == 3.0 / q and cl == -1.5 / q
    cd2, cl2 = diagnostics.force_coefficients(b, 1, rho=1.0, U=2.0, D=32.0, L=64.0, drag_axis=2, lift_axis=0)
    assert (cd2, cl2) == (0.25 / q, 3.0 / q)
    with pytest.raises(ValueError):
        diagnostics.force_coefficients(b, 1, rho=1.0, U=0.0, D=1.0, L=1.0)


def test_probe_line_fields():
    b = _bodies()
    vals = diagnostics.probe_line(7, b, 1, rhof=2.0).split()
    assert vals[0] == "7"
    got = np.array([float(x) for x in vals[1:]])
    want = np.concatenate([b.x[:, 1], b.v[:, 1], b.omega[:, 1], b.fdrag[:, 1] / 4.0 * 2.0,
                           b.fmom[:, 1] * 2.0, b.fcol[:, 1] / 4.0])
    np.testing.assert_array_equal(got, want)


def test_tecplot_round_trip(tmp_path):
    c = cases.build_case(6, 5, 4, [1] * 6, h=0.5)
    rng = np.random.default_rng(3)
    fields = [rng.normal(size=(4, 5, 6)) for _ in range(4)]
    p = tmp_path / "f.dat"
    diagnostics.write_tecplot(p, c, fields)
    nx, ny, nz, data = diagnostics.read_tecplot(p)
    assert (nx, ny, nz) == (6, 5, 4)
    # i fastest: row 1 is cell (2,1,1), row nx is cell (1,2,1)
    np.testing.assert_array_equal(data[1, :3], [1.5 * 0.5, 0.25, 0.25])
    np.testing.assert_array_equal(data[6, :3], [0.25, 0.75, 0.25])
    for q in range(3):
        np.testing.assert_array_equal(data[:, 3 + q], fields[q].reshape(-1))
    np.testing.assert_allclose(data[:, 6], np.sqrt(sum(f.reshape(-1) ** 2 for f in fields[:3])), rtol=1e-15)
    np.testing.assert_array_equal(data[:, 7], fields[3].reshape(-1))
    head = open(p).read().splitlines()[:2]
    assert head[0].strip() == "variables=x,y,z,u,v,w,vel,P"
    assert head[1].split() == ["zone", "i=", "6", "j=", "5", "k=", "4"]


def test_divergence_matches_the_library_rhs():
    """diagnostics.divergence equals the library's own divergence (the
    Poisson right-hand side is -div/dt*rhof with c = 1) on an open case."""
    import oracle
    from ibmdem.abi import Simulation
    st = cases.small_open(axis=0, periodic=(2,))
    with Simulation(oracle.library(), st.case) as s:
        cases.load(s, st)
        s.begin_step()
        s.predictor(1.0, 0.0, 1.0)
        s.divergence(1.0)
        _, rhs = s.download_poisson()
        f = s.download_fields()
        planes = {1: s.download_outflow(1)}
    d = diagnostics.divergence(st.case, f[:3], planes)
    np.testing.assert_allclose(-d / st.case.dt * st.case.rhof, rhs, rtol=1e-12, atol=1e-15)


def _probe_run(lib, steps, cap, graphs=True):
    import os
    from ibmdem.abi import Simulation
    setup = cases.small_walls()
    old = os.environ.get("IBMDEM_NO_GRAPH")
    if not graphs:
        os.environ["IBMDEM_NO_GRAPH"] = "1"
    try:
        s = Simulation(lib, setup.case)
    finally:
        if not graphs:
            if old is None:
                del os.environ["IBMDEM_NO_GRAPH"]
            else:
                os.environ["IBMDEM_NO_GRAPH"] = old
    with s:
        cases.load(s, setup)
        s.probe_bodies([1, 0], cap)
        states = []
        for _ in range(steps):
            s.step()
            b = s.download_bodies()
            states.append(np.stack([np.concatenate([getattr(b, f)[:, q] for f in s.PROBE_FIELDS]) for q in (1, 0)]))
        first, rec = s.probe_read()
    return first, rec, np.array(states)


def _check_probe(lib, graphs=True):
    first, rec, states = _probe_run(lib, 6, 4, graphs)
    assert first == 2 and rec.shape == (4, 2, 18)
    np.testing.assert_array_equal(rec, states[2:])


def test_probe_ring_on_the_oracle(oracle_lib):
    """The probe ring keeps the last `capacity` steps, oldest first, each
    record the listed bodies' state after its step."""
    _check_probe(oracle_lib)


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [True, False])
def test_probe_ring_on_the_gpu(gpu_lib, graphs):
    """The same on the GPU, recorded inside the captured step graphs or by
    eager launches, with no host wait."""
    _check_probe(gpu_lib, graphs)


def test_shedding_statistics_of_a_known_signal():
    """A lift signal A sin(2 pi f t) + c and a drag signal with mean 1.3:
    St = f D / U, rms C_L = A / sqrt(2), mean C_D = 1.3, from the window past
    the first half."""
    D, U, f = 32.0, 1.0, 1.0 / 160.0
    t = np.arange(0, 3000.0, 0.072)
    cl = 0.7 * np.sin(2 * np.pi * f * t + 0.3) + 0.01
    cd = 1.3 + 0.05 * np.sin(4 * np.pi * f * t)
    r = diagnostics.shedding_statistics(t, cd, cl, D, U)
    assert r["strouhal"] == pytest.approx(f * D / U, rel=1e-4)
    assert r["rms_cl"] == pytest.approx(0.7 / np.sqrt(2), rel=2e-3)
    assert r["mean_cd"] == pytest.approx(1.3, abs=2e-3)
    assert r["periods"] >= 8
    assert diagnostics.shedding_statistics(t[:100], cd[:100], cl[:100], D, U)["strouhal"] is None
