"""The CPU oracle pinned by closed forms, invariants and brute force.

Each test targets a mistake a restatement could plausibly make (a dropped
term, a wrong staggering offset, a sign, a transposed index) and checks the
oracle against mathematics that fixes the answer independently of the
oracle's own formulas.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from ibmdem import abi, cases
from ibmdem.abi import BC_PERIODIC, BC_WALL


def lib():
    l = oracle.library()
    l.ibmdem_oracle_phi.restype = C.c_double
    l.ibmdem_oracle_phi.argtypes = [C.c_double]
    l.ibmdem_oracle_eta.restype = C.c_double
    l.ibmdem_oracle_eta.argtypes = [C.c_double, C.c_int]
    l.ibmdem_oracle_rotation.restype = None
    l.ibmdem_oracle_rotation.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return l


# ---------------------------------------------------------------------------
# a14: Roma-Peskin-Berger 3-point kernel (interpolation.F90:2-33)
# ---------------------------------------------------------------------------
def test_delta_known_values():
    phi = lib().ibmdem_oracle_phi
    assert phi(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert phi(0.5) == pytest.approx(0.5, abs=1e-15)
    assert phi(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert phi(1.5) == 0.0
    assert phi(1.5000001) == 0.0 and phi(3.0) == 0.0


@pytest.mark.parametrize("s", np.linspace(0.0, 0.999, 37))
def test_delta_moments(s):
    """sum_j phi(s-j) = 1, sum_j (s-j) phi(s-j) = 0, sum_j phi^2 = 1/2 for
    every shift: the defining moment conditions of the 3-point kernel."""
    phi = lib().ibmdem_oracle_phi
    js = np.arange(-3, 4)
    w = np.array([phi(abs(s - j)) for j in js])
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    assert ((s - js) * w).sum() == pytest.approx(0.0, abs=1e-14)
    assert (w * w).sum() == pytest.approx(0.5, abs=1e-14)


def test_eta_quirk_and_textbook():
    """eta = -2 ln e / sqrt(pi_f32^2 + ln e) (ZJS_COLLISION_DEM.F90:30), not
    the textbook sqrt(pi^2 + ln^2 e)."""
    l = lib()
    e = 0.9
    le = math.log(e)
    pf2 = float(np.float32(np.float32(3.1415926536) * np.float32(3.1415926536)))
    assert l.ibmdem_oracle_eta(e, 0) == pytest.approx(-2 * le / math.sqrt(pf2 + le), rel=1e-15)
    assert l.ibmdem_oracle_eta(e, 0) == pytest.approx(0.06743, abs=5e-5)
    assert l.ibmdem_oracle_eta(e, 1) == pytest.approx(-2 * le / math.sqrt(math.pi ** 2 + le * le), rel=1e-15)
    assert l.ibmdem_oracle_eta(e, 1) == pytest.approx(0.06704, abs=5e-5)


def test_rotation_is_exponential_map():
    """CalculateRotateMatrix is Rodrigues' formula: R = exp([theta]x),
    orthogonal with det 1 (couplesolverCylinder.F90:76-97)."""
    from scipy.linalg import expm
    l = lib()
    rng = np.random.default_rng(3)
    for _ in range(10):
        th = rng.normal(size=3) * rng.uniform(0.001, 2.0)
        out = (C.c_double * 9)()
        l.ibmdem_oracle_rotation((C.c_double * 3)(*th), out)
        R = np.array(out).reshape(3, 3)
        K = np.array([[0, -th[2], th[1]], [th[2], 0, -th[0]], [-th[1], th[0], 0]])
        np.testing.assert_allclose(R, expm(K), atol=1e-14)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)


# ---------------------------------------------------------------------------
# helpers to build small states
# ---------------------------------------------------------------------------
def single_marker_case(n, btype, xs, fdrag=None, vol=0.7, **kw):
    case = cases.build_case(*n, btype, n_markers=len(xs), n_bodies=len(xs), **kw)
    m = len(xs)
    mk = abi.MarkerArrays.empty(m)
    bd = abi.BodyArrays.empty(m)
    for i, x in enumerate(xs):
        mk.x[:, i] = x
        mk.vol[i] = vol
        mk.mass[i] = 1.3
        mk.radeff[i] = 0.75
        mk.body[i] = i
        bd.x[:, i] = x
        bd.mass[i] = 10.0
        bd.rho[i] = 2.0
        bd.inertia[i] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        bd.rot[i] = bd.rot_glo[i] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    if fdrag is not None:
        mk.fdrag[:] = fdrag
    return case, mk, bd


def staggered_coords(n, h=1.0):
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    xc, yc, zc = (x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h
    return {"u": (xc - 0.5 * h, yc, zc), "v": (xc, yc - 0.5 * h, zc), "w": (xc, yc, zc - 0.5 * h)}


# ---------------------------------------------------------------------------
# a15-a17: staggered interpolation and spreading
# ---------------------------------------------------------------------------
def test_interpolation_exact_for_linear_fields():
    """Sum r phi = 0 makes interpolation exact for linear fields; a wrong
    staggering offset (u on faces, ZJS_INTERPOLATE_DEM.F90:12-14) breaks it."""
    n = (16, 16, 16)
    rng = np.random.default_rng(5)
    xs = rng.uniform(5.0, 11.0, size=(12, 3))
    case, mk, bd = single_marker_case(n, [BC_PERIODIC] * 6, xs)
    coef = {"u": (0.1, 0.02, 0.03, -0.01), "v": (-0.2, 0.01, -0.04, 0.02), "w": (0.05, -0.03, 0.01, 0.05)}
    st = staggered_coords(n)
    f = {c: coef[c][0] + coef[c][1] * st[c][0] + coef[c][2] * st[c][1] + coef[c][3] * st[c][2] for c in "uvw"}
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.upload_fields(f["u"], f["v"], f["w"], None)
        s.begin_step()
        s.interpolate()
        out = s.download_markers().fdrag * case.dt   # (U - 0)/dt * dt
    for d, c in enumerate("uvw"):
        exact = coef[c][0] + xs @ np.array(coef[c][1:])
        np.testing.assert_allclose(out[d], exact, rtol=1e-12, atol=1e-13)


def test_spread_conserves_force_and_moment():
    """f = -sum fdrag vol delta: sum f h^3 = -sum fdrag vol (partition of
    unity) and sum f x h^3 = -sum fdrag vol x_m (first moment), per staggered
    component, periodic images included."""
    n = (16, 12, 14)
    rng = np.random.default_rng(6)
    xs = rng.uniform(0.0, 1.0, size=(20, 3)) * np.array(n)
    xs[:4, 0] = [0.2, 15.9, 0.6, 15.3]          # across the periodic x face
    fd = rng.normal(size=(3, 20))
    case, mk, bd = single_marker_case(n, [BC_PERIODIC] * 6, xs, fdrag=fd)
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.spread()
        f = s.download_force()
    tot = -(fd * mk.vol).sum(axis=1)
    for d in range(3):
        assert f[d].sum() == pytest.approx(tot[d], rel=1e-12, abs=1e-12)
    # first moment in y for the u component (y is a centre coordinate of u,
    # no marker crosses the y faces here)
    st = staggered_coords(n)
    keep = (xs[:, 1] > 2) & (xs[:, 1] < n[1] - 2)
    fd2 = fd.copy()
    fd2[:, ~keep] = 0.0
    mk.fdrag[:] = fd2
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.spread()
        f = s.download_force()
    mom = (f[0] * st["u"][1]).sum()
    assert mom == pytest.approx(-(fd2[0] * mk.vol * xs[:, 1]).sum(), rel=1e-12)


def test_interpolation_spreading_adjoint():
    """<I u, F> = <u, S F> h^3: spreading is the adjoint of interpolation
    (same weights, ZJS_VOLUMEFRACTION_DEM.F90:12-22 vs :48-64)."""
    n = (12, 12, 12)
    rng = np.random.default_rng(8)
    xs = rng.uniform(0.0, 12.0, size=(15, 3))
    F = rng.normal(size=(3, 15))
    case, mk, bd = single_marker_case(n, [BC_PERIODIC] * 6, xs, fdrag=F)
    u = [rng.normal(size=(12, 12, 12)) for _ in range(3)]
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.upload_fields(*u, None)
        s.spread()
        f = s.download_force()
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        mk2 = mk.copy()
        mk2.fdrag[:] = 0.0
        s.upload_markers(mk2)
        s.upload_fields(*u, None)
        s.begin_step()
        s.interpolate()
        U = s.download_markers().fdrag * case.dt
    lhs = (U * F * mk.vol).sum()
    rhs = -sum((u[d] * f[d]).sum() for d in range(3))
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_wall_mirror_markers_brute_force():
    """Near a wall the reference spreads mirror copies x' = 2*wall - x with
    fdrag unnegated (ZJS_MPI_DEM.F90:244-258); compare with an explicit
    brute-force spread over the markers and their images."""
    n = (8, 8, 8)
    xs = np.array([[0.3, 3.2, 4.1], [4.4, 7.6, 0.45], [7.7, 0.2, 7.9]])
    fd = np.array([[1.0, -0.5, 0.3], [0.2, 0.8, -1.1], [-0.7, 0.4, 0.9]]).T
    case, mk, bd = single_marker_case(n, [BC_WALL] * 6, xs, fdrag=fd, vol=1.0)
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.spread()
        f = s.download_force()
    phi = np.vectorize(lib().ibmdem_oracle_phi)
    L = np.array(n, dtype=float)
    imgs = []
    for m in range(3):
        x = xs[m]
        for sx in (0, 1, 2):
            for sy in (0, 1, 2):
                for sz in (0, 1, 2):
                    p = x.copy()
                    ok = True
                    for d, sd in enumerate((sx, sy, sz)):
                        if sd == 1:
                            if x[d] >= 2: ok = False
                            p[d] = -x[d]
                        elif sd == 2:
                            if x[d] < L[d] - 2: ok = False
                            p[d] = 2 * L[d] - x[d]
                    if ok:
                        imgs.append((p, fd[:, m]))
    st = staggered_coords(n)
    for d, c in enumerate("uvw"):
        ref = np.zeros(n[::-1])
        for p, fm in imgs:
            w = phi(np.abs(st[c][0] - p[0])) * phi(np.abs(st[c][1] - p[1])) * phi(np.abs(st[c][2] - p[2]))
            ref -= fm[d] * w
        np.testing.assert_allclose(f[d], ref, atol=1e-13)


# ---------------------------------------------------------------------------
# a11-a13: the pressure Poisson operator and solve
# ---------------------------------------------------------------------------
def poisson_case(n, btype):
    return cases.build_case(*n, btype, rtol=1e-13, max_iter=400)


def test_poisson_matches_fft_periodic():
    n = (16, 12, 20)
    rng = np.random.default_rng(11)
    b = rng.normal(size=n[::-1])
    b -= b.mean()
    with oracle.simulation(poisson_case(n, [BC_PERIODIC] * 6)) as s:
        s.upload_poisson_rhs(b)
        info = s.poisson()
        x, _ = s.download_poisson()
    assert info.converged
    kz, ky, kx = [2 * np.pi * np.fft.fftfreq(m) for m in n[::-1]]
    lam = (2 - 2 * np.cos(kz))[:, None, None] + (2 - 2 * np.cos(ky))[None, :, None] + (2 - 2 * np.cos(kx))[None, None, :]
    bh = np.fft.fftn(b)
    lam[0, 0, 0] = 1.0
    xh = bh / lam
    xh[0, 0, 0] = 0.0
    ref = np.real(np.fft.ifftn(xh))
    ref -= ref[0, 0, 0]            # gauge P'(1,1,1) = 0 (poisson.F90:9,97-100)
    assert x[0, 0, 0] == 0.0
    np.testing.assert_allclose(x, ref, atol=1e-10 * np.abs(ref).max())


def test_poisson_matches_dct_walls():
    """Walls omit the neighbour (Neumann, initialization.F90:127-147): the
    cell-centred operator diagonalises in the DCT-II basis."""
    from scipy.fft import dctn, idctn
    n = (12, 10, 14)
    rng = np.random.default_rng(12)
    b = rng.normal(size=n[::-1])
    b -= b.mean()
    with oracle.simulation(poisson_case(n, [BC_WALL] * 6)) as s:
        s.upload_poisson_rhs(b)
        info = s.poisson()
        x, _ = s.download_poisson()
    assert info.converged
    lz, ly, lx = [2 - 2 * np.cos(np.pi * np.arange(m) / m) for m in n[::-1]]
    lam = lz[:, None, None] + ly[None, :, None] + lx[None, None, :]
    bh = dctn(b, type=2, norm="ortho")
    lam[0, 0, 0] = 1.0
    xh = bh / lam
    xh[0, 0, 0] = 0.0
    ref = idctn(xh, type=2, norm="ortho")
    ref -= ref[0, 0, 0]
    np.testing.assert_allclose(x, ref, atol=1e-10 * np.abs(ref).max())


def test_poisson_pinned_row_equivalence():
    """The reference pins row 0 to the identity with b[0] = 0
    (poisson.F90:78,126); with an RHS that does not sum to zero the pinned
    system's solution is the singular solve with b[0] := -sum_{i>0} b_i,
    gauged to P'(1,1,1) = 0. Check against a dense solve of the pinned
    nonsymmetric matrix on a tiny box."""
    n = (4, 3, 5)
    N = int(np.prod(n))
    rng = np.random.default_rng(13)
    b = rng.normal(size=n[::-1])          # deliberately not mean-free
    A = np.zeros((N, N))
    nx, ny, nz = n
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = (k * ny + j) * nx + i
                for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    q = (((k + dk) % nz) * ny + (j + dj) % ny) * nx + (i + di) % nx
                    A[r, q] -= 1.0
                    A[r, r] += 1.0
    A[0, :] = 0.0
    A[0, 0] = 1.0
    bb = b.ravel().copy()
    bb[0] = 0.0
    ref = np.linalg.solve(A, bb).reshape(n[::-1])
    with oracle.simulation(poisson_case(n, [BC_PERIODIC] * 6)) as s:
        s.upload_poisson_rhs(b)
        assert s.poisson().converged
        x, _ = s.download_poisson()
    np.testing.assert_allclose(x, ref, atol=1e-10 * np.abs(ref).max())


# ---------------------------------------------------------------------------
# a2-a9: fluid operators
# ---------------------------------------------------------------------------
def taylor_green_error(n):
    """2-D Taylor-Green vortex (the reference's own benchmark,
    initialization.F90:299-310, fluid.F90:406-457) on [0,2pi]^2 x 2 cells,
    run to t = 0.4 with dt ~ h^2; max error of u against the exact decay."""
    nu = 0.1
    h = 2 * np.pi / n
    dt = 0.05 * h * h / nu
    steps = int(round(0.4 / dt))
    dt = 0.4 / steps
    case = cases.build_case(n, n, 2, [BC_PERIODIC] * 6, h=h, visc=nu, dt=dt, rtol=1e-13, max_iter=300)
    case.n_markers = 0
    case.n_bodies = 0
    st = staggered_coords((n, n, 2), h)
    u = np.cos(st["u"][0]) * np.sin(st["u"][1])
    v = -np.sin(st["v"][0]) * np.cos(st["v"][1])
    xc, yc = st["w"][0], st["w"][1]
    p = -0.25 * (np.cos(2 * xc) + np.cos(2 * yc))
    with oracle.simulation(case) as s:
        s.upload_fields(u, v, np.zeros_like(u), p)
        s.run(steps)
        uo, vo, wo, po = s.download_fields()
    t = steps * dt
    ue = np.cos(st["u"][0]) * np.sin(st["u"][1]) * np.exp(-2 * nu * t)
    return np.abs(uo - ue).max()


def test_taylor_green_second_order():
    e1 = taylor_green_error(12)
    e2 = taylor_green_error(24)
    assert e2 < e1
    order = math.log(e1 / e2, 2)
    assert 1.7 < order < 2.5, (e1, e2, order)


def taylor_green_run(n, dt, T, scheme):
    nu = 0.1
    h = 2 * np.pi / n
    steps = int(round(T / dt))
    case = cases.build_case(n, n, 2, [BC_PERIODIC] * 6, h=h, visc=nu, dt=T / steps, rtol=1e-13,
                            max_iter=300, time_scheme=scheme)
    case.n_markers = 0
    case.n_bodies = 0
    st = staggered_coords((n, n, 2), h)
    u = np.cos(st["u"][0]) * np.sin(st["u"][1])
    v = -np.sin(st["v"][0]) * np.cos(st["v"][1])
    p = -0.25 * (np.cos(2 * st["w"][0]) + np.cos(2 * st["w"][1]))
    with oracle.simulation(case) as s:
        s.upload_fields(u, v, np.zeros_like(u), p)
        s.run(steps)
        return s.download_fields()[0]


def test_time_order_euler_first_rk3_third():
    """Temporal order on a fixed grid against a run with an 8x smaller step:
    forward Euler (FirstOrderForwardEuler) is first order; the low-storage
    RK3 with Wray's coefficients (fluid.F90:333-336) is third order, which a
    wrong coefficient or a stale stage velocity would break."""
    for scheme, lo, hi in ((0, 0.9, 1.4), (1, 2.7, 3.3)):
        ref = taylor_green_run(16, 0.8 / 128, 0.8, scheme)
        e = [np.abs(taylor_green_run(16, dt, 0.8, scheme) - ref).max() for dt in (0.1, 0.05)]
        order = math.log(e[0] / e[1], 2)
        assert lo < order < hi, (scheme, e, order)


def test_projection_makes_velocity_divergence_free():
    setup = cases.small_periodic(n=(20, 16, 18), n_spheres=3, D=6.0, gap=0.7, rtol=1e-12)
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        s.begin_step()
        for _ in range(2):
            s.interpolate()
            s.spread()
        s.predictor()
        s.bulk()
        s.divergence()
        _, rhs0 = s.download_poisson()
        s.poisson()
        s.project()
        s.divergence()
        _, rhs1 = s.download_poisson()
    assert np.abs(rhs1).max() < 1e-10 * np.abs(rhs0).max()


def test_bulk_forcing_sets_mean():
    setup = cases.small_periodic(n=(16, 12, 14), n_spheres=2, D=5.0)
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        s.bulk()
        u, v, w, _ = s.download_fields()
    assert u.mean() == pytest.approx(setup.case.bulk_target[0], abs=1e-15)
    assert w.mean() == pytest.approx(setup.case.bulk_target[2], abs=1e-15)


# ---------------------------------------------------------------------------
# a19-a27: cells, candidate lists and contact
# ---------------------------------------------------------------------------
def test_cells_match_floor():
    setup = cases.small_periodic()
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        cells = s.cells()
        x = s.download_markers().x
    ref = np.floor(x.astype(np.longdouble) / 1.0).astype(np.int64) + 1
    np.testing.assert_array_equal(cells, ref)


def brute_pairs(x, body, L, lim):
    out = set()
    m = x.shape[1]
    for i in range(m):
        d = x[:, i][:, None] - x
        s = -np.round(d / L[:, None]).astype(int)
        d = d + s * L[:, None]
        d2 = (d * d).sum(axis=0)
        for j in np.nonzero((d2 < lim * lim) & (body != body[i]))[0]:
            out.add((i, int(j), tuple(-s[:, j])))
    return out


def test_pair_list_brute_force():
    setup = cases.small_periodic(gap=0.6)
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        cnt, par, sh = s.pairs()
        x = s.download_markers().x
    L = np.array([setup.case.nx, setup.case.ny, setup.case.nz], dtype=float)
    lim = setup.case.collision_dist * (1 + setup.case.collision_buffer)
    got = set()
    for i in range(cnt.size):
        for q in range(cnt[i]):
            code = int(sh[i, q])
            shift = (code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1)
            got.add((i, int(par[i, q]), shift))
    ref = brute_pairs(x, setup.markers.body, L, lim)
    # the oracle stores partner position x_j + shift*L; brute force stores
    # the image shift that brings j next to i
    assert len(ref) > 0
    assert got == ref


def test_two_marker_contact_closed_form():
    """Spring-dashpot law (ZJS_COLLISION_DEM.F90:36-69) for one overlapping
    pair, and f_ji = -f_ij bit for bit (each side computed separately)."""
    n = (16, 16, 16)
    xs = np.array([[8.0, 8.0, 8.0], [9.1, 8.3, 7.9]])
    case, mk, bd = single_marker_case(n, [BC_PERIODIC] * 6, xs)
    mk.v[:, 0] = [0.01, -0.02, 0.0]
    mk.v[:, 1] = [-0.03, 0.0, 0.01]
    mk.mass[:] = [1.3, 0.9]
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.contact()
        f = s.download_markers().fcol
    r = xs[0] - xs[1]
    d = np.linalg.norm(r)
    nn = r / d
    meff = 2 * 1.3 * 0.9 / (1.3 + 0.9)
    kn = case.kn_coeff / (0.1 * 1.0) * meff
    eta = lib().ibmdem_oracle_eta(case.restitution, 0)
    vn = nn @ (mk.v[:, 0] - mk.v[:, 1])
    ref = nn * (kn * (1.5 - d) - eta * math.sqrt(meff * kn) * vn)
    np.testing.assert_allclose(f[:, 0], ref, rtol=1e-13)
    np.testing.assert_array_equal(f[:, 1], -f[:, 0])


def test_wall_contact_closed_form():
    n = (10, 10, 10)
    xs = np.array([[5.0, 5.0, 0.4]])
    case, mk, bd = single_marker_case(n, [BC_WALL] * 6, xs)
    mk.v[:, 0] = [0.0, 0.0, -0.02]
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.contact()
        f = s.download_markers().fcol
    kn = case.kn_coeff / 0.1 * 1.3
    eta = lib().ibmdem_oracle_eta(case.restitution, 0)
    ref = kn * (0.75 - 0.4) - eta * math.sqrt(1.3 * kn) * (-0.02)
    np.testing.assert_allclose(f[:, 0], [0.0, 0.0, ref], rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------
# a28-a33: rigid bodies
# ---------------------------------------------------------------------------
def test_free_fall_explicit_euler():
    """No hydrodynamic or contact force: v after N DEM sub-steps is
    N*ratio*g*dtDEM and x follows the explicit-Euler sum
    (couplesolverCylinder.F90:21,37-39)."""
    n = (8, 8, 64)
    case, mk, bd = single_marker_case(n, [BC_WALL] * 6, np.array([[4.0, 4.0, 40.0]]),
                                      gravity=(0.0, 0.0, -0.01), dt=1.0)
    bd.rho[0] = 2.5
    ratio = (2.5 - 1.0) / 2.5
    dtd = case.dt / case.couple_interval
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        vz, xz = 0.0, 40.0
        for k in range(30):
            s.body_integrate()
            vz = vz + 0.0 * dtd + 0.0 / 10.0 * dtd + ratio * (-0.01) * dtd
            xz = xz + vz * dtd
        b = s.download_bodies()
    assert b.v[2, 0] == vz
    assert b.x[2, 0] == xz
    assert b.v[2, 0] == pytest.approx(30 * ratio * -0.01 * dtd, rel=1e-13)


def test_markers_move_rigidly():
    setup = cases.small_periodic(n=(24, 20, 22), n_spheres=2, D=6.0, gap=3.0)
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        x0 = s.download_markers().x.copy()
        for _ in range(5):
            s.dem_substep()
        x1 = s.download_markers().x
    L = np.array([24.0, 20.0, 22.0])[:, None]
    body = setup.markers.body
    for b in range(2):
        idx = np.nonzero(body == b)[0][:40]
        def dists(x):
            d = x[:, idx][:, :, None] - x[:, idx][:, None, :]
            d -= L[:, :, None] * np.round(d / L[:, :, None])
            return np.sqrt((d * d).sum(axis=0))
        np.testing.assert_allclose(dists(x1), dists(x0), atol=1e-12)


def test_body_force_assembly():
    """AssembleLocalForceMS (ZJS_MS_DEM.F90:494-519): body force = rho_f sum
    fdrag vol, torque = sum r x (fdrag vol) (no rho_f), brute force."""
    setup = cases.small_periodic(n=(24, 20, 22), n_spheres=2, D=6.0, gap=0.6)
    rng = np.random.default_rng(21)
    setup.markers.fdrag[:] = rng.normal(size=setup.markers.fdrag.shape)
    setup.case.rhof = 1.7
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        s.body_assemble()
        b = s.download_bodies()
        mk = s.download_markers()
    L = np.array([24.0, 20.0, 22.0])
    for k in range(2):
        sel = setup.markers.body == k
        fv = mk.fdrag[:, sel] * mk.vol[sel]
        r = mk.x[:, sel] - b.x[:, k][:, None]
        r -= L[:, None] * np.round(r / L[:, None])
        np.testing.assert_allclose(b.fdrag[:, k], 1.7 * fv.sum(axis=1), rtol=1e-12)
        np.testing.assert_allclose(b.fmom[:, k], np.cross(r.T, fv.T).sum(axis=0), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(b.fcol[:, k], mk.fcol[:, sel].sum(axis=1), rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------------------
# a2-a5 in three dimensions: the ABC (Beltrami) flow
# ---------------------------------------------------------------------------
def abc_flow_error(n, A=1.0, B=0.7, Cc=0.5, nu=0.1, T=0.3):
    """Arnold-Beltrami-Childress flow u = (A sin z + C cos y, B sin x + A cos z,
    C sin y + B cos x) on the triply periodic [0, 2pi]^3. Its vorticity equals
    the velocity, so (u.grad)u = grad(|u|^2/2) is absorbed by the pressure and
    the exact solution decays as exp(-nu t): every flux of ConvectionTensor
    (fluid.F90:15-44) including the z-fluxes and the cross terms, and every
    direction of ViscousTensor (fluid.F90:58-66), must be second-order
    accurate for the discrete solution to converge at second order. dt ~ h^2
    keeps the first-order time error at the same order."""
    h = 2 * np.pi / n
    dt = 0.04 * h * h / nu
    steps = int(round(T / dt))
    dt = T / steps
    case = cases.build_case(n, n, n, [BC_PERIODIC] * 6, h=h, visc=nu, dt=dt, rtol=1e-13, max_iter=300)
    case.n_markers = 0
    case.n_bodies = 0
    st = staggered_coords((n, n, n), h)
    def vel(t):
        g = math.exp(-nu * t)
        u = (A * np.sin(st["u"][2]) + Cc * np.cos(st["u"][1])) * g
        v = (B * np.sin(st["v"][0]) + A * np.cos(st["v"][2])) * g
        w = (Cc * np.sin(st["w"][1]) + B * np.cos(st["w"][0])) * g
        return u, v, w
    u0, v0, w0 = vel(0.0)
    with oracle.simulation(case) as s:
        s.upload_fields(u0, v0, w0, np.zeros_like(u0))
        s.run(steps)
        uo, vo, wo, _ = s.download_fields()
    ue, ve, we = vel(steps * dt)
    return max(np.abs(uo - ue).max(), np.abs(vo - ve).max(), np.abs(wo - we).max())


def test_abc_flow_second_order_in_3d():
    e1 = abc_flow_error(12)
    e2 = abc_flow_error(24)
    order = math.log(e1 / e2, 2)
    assert 1.7 < order < 2.5, (e1, e2, order)
    # the error is a discretisation error, far below the decay itself
    assert e2 < 2e-3


# ---------------------------------------------------------------------------
# a29: CloudEvolutionMS, fhydro and fmomHydro (couplesolverCylinder.F90:194-211)
# ---------------------------------------------------------------------------
def test_cloud_evolution_recurrence():
    """With no IBM or contact force, CloudEvolutionMS sets
    fhydro = (v - vold)/dt (1 - ratio) and the ten sub-steps add fhydro dt
    and ratio g dt, so the per-step increment d_n = v_{n+1} - v_n obeys
    d_n = (1 - ratio) d_{n-1} + ratio g dt. From vold = v (d_{-1} = 0):
    d_n = g dt (1 - (1 - ratio)^(n+1)). The angular momentum, with no torque,
    has e_n = L_{n+1} - L_n = (1 - ratio)^(n+1) (L_0 - Lold_0)."""
    n = (16, 16, 64)
    g = (0.0, 0.0, -0.01)
    case, mk, bd = single_marker_case(n, [BC_WALL] * 6, np.array([[8.0, 8.0, 40.0]]), gravity=g, dt=0.5)
    case.rhof = 1.3
    rho = 3.9
    ratio = (rho - 1.3) / rho
    bd.rho[0] = rho
    bd.v[:, 0] = [0.01, -0.02, 0.005]
    bd.vold[:, 0] = bd.v[:, 0]
    bd.angmom[:, 0] = [0.3, -0.1, 0.2]
    bd.angmom_old[:, 0] = [0.1, 0.05, 0.0]
    mk.v[:, 0] = bd.v[:, 0]
    dt = case.dt
    steps = 12
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        vs = [s.download_bodies().v[:, 0].copy()]
        Ls = [s.download_bodies().angmom[:, 0].copy()]
        for _ in range(steps):
            s.cloud_evolution()
            b = s.download_bodies()
            vs.append(b.v[:, 0].copy())
            Ls.append(b.angmom[:, 0].copy())
            assert np.all(b.fcol == 0.0) and np.all(b.fdrag == 0.0)
    vs, Ls = np.array(vs), np.array(Ls)
    q = 1.0 - ratio
    # increments are differences of the state: compare to a few ulps of it
    for k in range(steps):
        d_ref = np.asarray(g) * dt * (1.0 - q ** (k + 1))
        np.testing.assert_allclose(vs[k + 1] - vs[k], d_ref, rtol=1e-12, atol=8e-16 * np.abs(vs).max())
        e_ref = q ** (k + 1) * (bd.angmom[:, 0] - bd.angmom_old[:, 0])
        np.testing.assert_allclose(Ls[k + 1] - Ls[k], e_ref, rtol=1e-12, atol=8e-16 * np.abs(Ls).max())


# ---------------------------------------------------------------------------
# a36: the composed step. Momentum budget in a periodic box.
# ---------------------------------------------------------------------------
def momentum_budget_run(gravity, steps=4):
    """Returns per-step (fluid momentum rho_f sum u dV, body momenta
    sum_b (m_b v_b - rho_f V_b vold_b), gravity impulse) for a periodic box of
    spheres in contact with random fluid and body velocities. Per step the
    spread adds rho_f dt sum f dV = -dt sum_b F_b to the fluid (partition of
    unity), the body update adds dt F_b + (m - rho_f V) g dt + contact
    (pairwise antisymmetric) to m v - rho_f V vold (the fhydro term of
    CloudEvolutionMS is rho_f V (v - vold)/dt since m (1 - ratio) = rho_f V),
    and the convective, viscous and pressure terms sum to zero in a periodic
    box. Bulk forcing is off."""
    setup = cases.small_periodic(n=(24, 20, 22), n_spheres=3, D=6.0, gap=0.6, rtol=1e-13)
    case = setup.case
    for d in range(3):
        case.bulk_switch[d] = 0
        case.gravity[d] = gravity[d]
    case.rhof = 1.3
    setup.bodies.rho[:] *= 1.3
    setup.bodies.mass[:] *= 1.3
    setup.bodies.inertia[:] *= 1.3
    setup.markers.mass[:] *= 1.3
    rf = case.rhof
    dV = case.dx * case.dy * case.dz
    vol_b = setup.bodies.mass / setup.bodies.rho
    out = []
    with oracle.simulation(case) as s:
        cases.load(s, setup)
        for k in range(steps + 1):
            u, v, w, _ = s.download_fields()
            b = s.download_bodies()
            pf = rf * dV * np.array([u.sum(), v.sum(), w.sum()])
            pb = (b.mass[None, :] * b.v - rf * vol_b[None, :] * b.vold).sum(axis=1)
            out.append((pf, pb, b.fcol.sum(axis=1)))
            if k < steps:
                s.step()
    grav = ((setup.bodies.mass - rf * vol_b).sum() * np.asarray(gravity) * case.dt)
    return out, grav


@pytest.mark.parametrize("gravity", [(0.0, 0.0, 0.0), (0.0, 0.0, -2e-3)])
def test_composed_step_momentum_budget(gravity):
    out, grav = momentum_budget_run(gravity)
    scale = max(np.abs(o[0]).max() + np.abs(o[1]).max() for o in out)
    moved = False
    for k in range(len(out) - 1):
        (pf0, pb0, _), (pf1, pb1, _) = out[k], out[k + 1]
        change = (pf1 + pb1) - (pf0 + pb0)
        np.testing.assert_allclose(change, grav, rtol=0, atol=1e-11 * scale)
        moved |= np.abs(pf1 - pf0).max() > 1e-6 * scale
    assert moved, "the fluid and the bodies exchanged no momentum"


def test_cloud_evolution_torque_term():
    """One CloudEvolutionMS call on a lone body whose markers carry a fixed
    fdrag: fmomHydro = (L - Lold)/dt (1 - ratio) + T with T = sum r x
    (fdrag vol) (couplesolverCylinder.F90:198-210, ZJS_MS_DEM.F90:509-510),
    constant over the ten sub-steps (no contact), so
    L_1 - L_0 = (L_0 - Lold_0)(1 - ratio) + T dt; T by brute force from the
    marker geometry."""
    n = (24, 24, 24)
    rng = np.random.default_rng(17)
    centre = np.array([12.0, 11.5, 12.5])
    m = 6
    case = cases.build_case(*n, [BC_WALL] * 6, n_markers=m, n_bodies=1, dt=0.4)
    case.rhof = 1.2
    mk = abi.MarkerArrays.empty(m)
    bd = abi.BodyArrays.empty(1)
    mk.x[:] = centre[:, None] + rng.uniform(-3.0, 3.0, size=(3, m))
    mk.fdrag[:] = rng.normal(size=(3, m))
    mk.vol[:] = rng.uniform(0.5, 1.0, size=m)
    mk.mass[:] = 0.8
    mk.radeff[:] = 0.75
    bd.x[:, 0] = centre
    bd.mass[0] = 50.0
    bd.rho[0] = 3.0
    bd.inertia[0] = [40.0, 0, 0, 0, 45.0, 0, 0, 0, 50.0]
    bd.rot[0] = bd.rot_glo[0] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    bd.angmom[:, 0] = [0.2, -0.1, 0.05]
    bd.angmom_old[:, 0] = [0.1, 0.0, 0.1]
    ratio = (3.0 - 1.2) / 3.0
    r = mk.x - centre[:, None]
    T = np.cross(r.T, (mk.fdrag * mk.vol[None, :]).T).sum(axis=0)
    with oracle.simulation(case) as s:
        s.upload_bodies(bd)
        s.upload_markers(mk)
        s.cloud_evolution()
        b = s.download_bodies()
    want = (bd.angmom[:, 0] - bd.angmom_old[:, 0]) * (1 - ratio) + T * case.dt
    np.testing.assert_allclose(b.angmom[:, 0] - bd.angmom[:, 0], want, rtol=1e-12,
                               atol=1e-14 * np.abs(bd.angmom).max())
    assert np.abs(T).max() > 0.1
