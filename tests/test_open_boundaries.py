"""Inflow and outflow sides on the CPU oracle, pinned by what the
mathematics fixes (the reference's fluid has no open side, so there is no
reference output to compare with):

  * a uniform stream through a transversely periodic channel is an exact
    steady state (every difference vanishes, the Poisson right-hand side is
    zero);
  * a tangential velocity sin(pi x / 2L), zero on the inflow face and with
    zero slope on the outflow face, is an exact eigenvector of the discrete
    diffusion with the open-side ghost rules, so each step multiplies it by
    1 - 4 nu dt sin^2(pi h / 4L) / h^2;
  * after a projection the field is discretely divergence-free in every cell,
    the last layer included (its outer face is the outflow plane), so the
    outflow flux equals the inflow flux;
  * the Poisson operator with P' = 0 on the outflow face equals a dense
    matrix assembled here from that definition (ghost = -P'), solved with
    numpy.
"""
import ctypes as C
import os

import numpy as np
import pytest

import ibmdem
import oracle
from ibmdem import abi, cases
from ibmdem.abi import BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, IbmdemError, Simulation


@pytest.fixture(scope="module")
def lib():
    return oracle.library()


def _div(fields, plane, axis, bt, h=(1.0, 1.0, 1.0)):
    """Cell divergence of a staggered field; the high face of the open axis
    is the outflow plane, high wall faces are 0, periodic faces wrap."""
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
        nxt = np.concatenate([np.take(a, range(1, a.shape[na]), axis=na), hi], axis=na)
        d = d + (nxt - a) / h[c]
    return d


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_stream_is_steady(lib, axis):
    n = (12, 10, 8)
    bt = [BC_PERIODIC] * 6
    bt[2 * axis], bt[2 * axis + 1] = BC_INFLOW, BC_OUTFLOW
    U = 0.07
    inflow = [0.0] * 3
    inflow[axis] = U
    case = cases.build_case(*n, bt, visc=0.05, dt=0.2, inflow=inflow)
    u, v, w, p = cases.zero_fields(case)
    [u, v, w][axis][:] = U
    with Simulation(lib, case) as s:
        s.upload_fields(u, v, w, p)
        for _ in range(4):
            info = s.step()
            assert info.iterations == 0          # zero right-hand side
        got = s.download_fields()
        plane = s.download_outflow(2 * axis + 1)
    for c in range(3):
        np.testing.assert_array_equal(got[c], U if c == axis else 0.0)
    np.testing.assert_array_equal(got[3], 0.0)
    np.testing.assert_array_equal(plane, U)


@pytest.mark.parametrize("comp", [1, 2])
def test_tangential_diffusion_eigenmode(lib, comp):
    """x inflow (U = 0) / outflow, y and z periodic; v (or w) =
    A sin(k x_c), k = pi / (2L), x_c the cell centre: the inflow ghost
    -v(1) and the outflow ghost +v(n) make it an exact eigenvector of the
    discrete Laplacian."""
    nx, ny, nz = 16, 6, 4
    h, nu, dt = 1.0, 0.1, 0.5
    bt = [BC_INFLOW, BC_OUTFLOW] + [BC_PERIODIC] * 4
    case = cases.build_case(nx, ny, nz, bt, visc=nu, dt=dt, h=h)
    u, v, w, p = cases.zero_fields(case)
    L = nx * h
    k = np.pi / (2 * L)
    xc = (np.arange(nx) + 0.5) * h
    prof = 0.3 * np.sin(k * xc)
    [u, v, w][comp][:] = prof[None, None, :]
    nsteps = 5
    with Simulation(lib, case) as s:
        s.upload_fields(u, v, w, p)
        for _ in range(nsteps):
            s.step()
        got = s.download_fields()
    lam = -4.0 / (h * h) * np.sin(k * h / 2) ** 2
    want = prof * (1.0 + dt * nu * lam) ** nsteps
    np.testing.assert_allclose(got[comp][0, 0, :], want, rtol=1e-13, atol=1e-16)
    for c in range(3):
        if c != comp:
            assert np.abs(got[c]).max() == 0.0
    # the mode decayed measurably (another ghost at either end would put an
    # O(nu dt / h^2) error on the first or last cell instead)
    assert np.abs(got[comp][0, 0, :] - prof).max() > 1e-4


@pytest.mark.parametrize("axis,periodic", [(0, (2,)), (2, ()), (1, (0,))])
def test_projection_divergence_free_and_mass_balance(lib, axis, periodic):
    st = cases.small_open(axis=axis, periodic=periodic)
    bt = list(st.case.btype)
    with Simulation(lib, st.case) as s:
        cases.load(s, st)
        for _ in range(2):
            info = s.step()
            assert info.converged
        u, v, w, p = s.download_fields()
        plane = s.download_outflow(2 * axis + 1)
    f = (u, v, w)
    div = _div(f, plane, axis, bt)
    # the solve stops at rtol 1e-12 of the right-hand side's 2-norm over
    # about 1.5e4 cells; a wrong outflow term leaves O(1e-4)
    scale = np.abs(f[axis]).max()
    assert np.abs(div).max() < 1e-10 * scale
    U = st.case.inflow[axis]
    # the inflow face holds U; outflow flux = inflow flux
    np.testing.assert_array_equal(np.take(f[axis], 0, axis=2 - axis), U)
    assert abs(plane.sum() - U * plane.size) < 1e-11 * U * plane.size
    # the plane is not just a copy of the last face: the projection moved it
    assert np.abs(plane - np.take(f[axis], -1, axis=2 - axis)).max() > 1e-8


def _dense_operator(n, bt, h):
    """-lap_h with Neumann walls/inflow, wrapped periodic sides and ghost =
    -x beyond an outflow side, assembled from that definition."""
    nx, ny, nz = n
    N = nx * ny * nz
    A = np.zeros((N, N))
    idx = lambda i, j, k: (k * ny + j) * nx + i
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                pos = [i, j, k]
                for d in range(3):
                    ih2 = 1.0 / h[d] ** 2
                    for step in (-1, 1):
                        q = list(pos)
                        q[d] += step
                        side = 2 * d + (0 if step < 0 else 1)
                        if 0 <= q[d] < n[d]:
                            A[r, r] += ih2
                            A[r, idx(*q)] -= ih2
                        elif bt[side] == BC_PERIODIC:
                            q[d] %= n[d]
                            A[r, r] += ih2
                            A[r, idx(*q)] -= ih2
                        elif bt[side] == BC_OUTFLOW:
                            A[r, r] += 2.0 * ih2      # (x - (-x)) / h^2
    return A


@pytest.mark.parametrize("bt", [
    [BC_INFLOW, BC_OUTFLOW, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC],
    [BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC, BC_INFLOW, BC_OUTFLOW],
    [BC_WALL, BC_OUTFLOW, BC_WALL, BC_OUTFLOW, BC_WALL, BC_WALL],
])
def test_dirichlet_poisson_equals_dense_solve(lib, bt):
    n = (8, 6, 4)
    h = (1.0, 0.5, 2.0)
    case = cases.build_case(*n, bt, rtol=1e-13, max_iter=300)
    case.dx, case.dy, case.dz = h
    rng = np.random.default_rng(11)
    rhs = rng.normal(size=(n[2], n[1], n[0]))
    A = _dense_operator(n, bt, h)
    want = np.linalg.solve(A, rhs.reshape(-1)).reshape(rhs.shape)
    with Simulation(lib, case) as s:
        s.upload_poisson_rhs(rhs)
        info = s.poisson()
        assert info.converged
        got, _ = s.download_poisson()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_outflow_plane_io(lib):
    st = cases.small_open(axis=2)
    with Simulation(lib, st.case) as s:
        cases.load(s, st)
        pl = s.download_outflow(5)
        # upload of w resets the plane to the last interior face
        np.testing.assert_array_equal(pl, st.fields[2][-1])
        new = np.arange(pl.size, dtype=np.float64).reshape(pl.shape)
        s.upload_outflow(5, new)
        np.testing.assert_array_equal(s.download_outflow(5), new)
        # sides 1 and 3 are walls here; 0, 2, 4 are low sides
        for side in (1, 3):
            with pytest.raises(IbmdemError):
                s.download_outflow(side)
        buf = np.zeros(st.case.nx * st.case.ny * st.case.nz)
        for side in (0, 2, 4, 6):
            with pytest.raises(IbmdemError):
                s.call("ibmdem_download_outflow", side, buf.ctypes.data_as(C.POINTER(C.c_double)))


def _libs():
    out = [("oracle", oracle.library())]
    if os.path.exists(ibmdem.GPU_LIB):
        out.append(("gpu", C.CDLL(ibmdem.GPU_LIB)))
    return out


@pytest.mark.parametrize("name,l", _libs())
def test_open_side_validation(name, l):
    """Both libraries reject, before any device work: inflow on a high side,
    outflow on a low side, an open side mixed with periodic, bulk forcing
    with an open side."""
    def rc(bt, **kw):
        case = cases.build_case(8, 8, 8, bt, **kw)
        h = C.c_void_p()
        r = l.ibmdem_create(C.byref(case), C.byref(h))
        if r == 0:
            l.ibmdem_destroy(h)
        return r
    P, W, I, O = BC_PERIODIC, BC_WALL, BC_INFLOW, BC_OUTFLOW
    ERR_ARG, ERR_UNSUPPORTED = 2, 8
    assert rc([O, I, W, W, W, W]) == ERR_ARG
    assert rc([I, I, W, W, W, W]) == ERR_ARG
    assert rc([O, O, W, W, W, W]) == ERR_ARG
    assert rc([I, P, W, W, W, W]) == ERR_ARG
    assert rc([I, O, P, P, P, P], bulk_switch=(1, 0, 0)) == ERR_UNSUPPORTED
    if name == "oracle":
        assert rc([I, O, W, W, P, P]) == abi.IBMDEM_OK
        assert rc([W, O, P, P, I, W]) == abi.IBMDEM_OK
