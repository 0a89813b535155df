"""Case set-up from the reference's own inputs (no GPU): the marker-template
reader, Rodrigues' formula, and the reference's 86-cylinder case
(InitCylinderParcelsDEM, ZJS_INIT_DEM.F90:350-485) built from the
CYLINDER_LP1125.dat template (committed as tests/golden/cylinder_lp1125.npy)."""
import math
import os

import numpy as np
import pytest

from ibmdem import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def template():
    return np.load(os.path.join(GOLDEN, "cylinder_lp1125.npy"))


def test_template_reader(tmp_path):
    p = tmp_path / "t.dat"
    p.write_text("  -4.7629688e-03  -2.0912142e-01   5.0000000e-01\n 1.0 2.0 3.0\n\n")
    t = cases.read_marker_template(str(p))
    np.testing.assert_array_equal(t, [[-4.7629688e-03, -2.0912142e-01, 0.5], [1.0, 2.0, 3.0]])


def test_template_is_a_unit_cylinder():
    """The template the reference scales: radius 1/4, height 1, centred."""
    t = template()
    assert t.shape == (1125, 3)
    r = np.hypot(t[:, 0], t[:, 1])
    assert r.max() == pytest.approx(0.25, abs=1e-7)
    assert t[:, 2].min() == -0.5 and t[:, 2].max() == 0.5


def test_rotation_matrix_is_the_exponential_map():
    from scipy.linalg import expm
    for th in ((0.0, 0.7, 0.0), (0.3, -1.1, 0.4), (1e-12, 0.0, 0.0)):
        K = np.array([[0, -th[2], th[1]], [th[2], 0, -th[0]], [-th[1], th[0], 0]])
        np.testing.assert_allclose(cases.rotation_matrix(th).reshape(3, 3), expm(K), atol=1e-14)


def test_reference_cylinder_case_geometry():
    st = cases.config_reference_cylinders(template())
    dx, rad, h = st.meta["dx"], st.meta["rad"], st.meta["height"]
    assert rad == float(np.float32(2.08e-4)) and h == 4 * rad
    assert st.case.n_bodies == 86 and st.case.n_markers == 86 * 1125
    assert (st.case.nx, st.case.ny, st.case.nz) == (240, 32, 240)
    L = np.array([9e-3, 1.2e-3, 9e-3])
    # every centre inside the domain, on two y planes, all distinct
    c = st.bodies.x.T
    assert np.all((c > 0) & (c < L)) and len({tuple(q) for q in c}) == 86
    assert sorted(set(np.round(c[:, 1] / dx, 3))) == [8.0, 24.0]   # (+1e-10 m offsets)
    m, b = st.markers, st.bodies
    for j in (0, 17, 85):
        sel = m.body == j
        R = np.asarray(b.rot[j]).reshape(3, 3)
        # a tilt about y only
        assert R[1, 1] == pytest.approx(1.0, abs=1e-15) and abs(R[0, 1]) < 1e-15
        loc = R.T @ (m.x[:, sel] - b.x[:, j:j + 1])
        r = np.hypot(loc[0], loc[1])
        # side markers on the retracted radius, caps at the retracted height
        assert r.max() == pytest.approx(rad - 0.3 * dx, rel=1e-7)
        assert np.abs(loc[2]).max() == pytest.approx(h / 2 - 0.3 * dx, rel=1e-12)
        vol_b = math.pi * rad ** 2 * h
        assert b.mass[j] == pytest.approx(2650.0 * vol_b, rel=1e-15)
        assert b.inertia[j][8] == pytest.approx(0.5 * b.mass[j] * rad ** 2, rel=1e-15)
        assert b.inertia[j][0] == pytest.approx(b.mass[j] * (3 * rad ** 2 + h ** 2) / 12, rel=1e-15)
    vol_m = math.pi * dx / 1125 * (rad ** 2 + 2 * rad * h + 0.25 * dx ** 2)
    assert np.all(m.vol == vol_m)
    # markers inside the walled y span
    assert m.x[1].min() > 0 and m.x[1].max() < L[1]
