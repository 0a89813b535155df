"""Failure statuses of the C ABI (no counterpart in the reference, which
checks nothing: ZJS_UPDATE_LIST_DEM.F90:46-51 tests nothing that can fail,
and the bins and pair lists of ZJS_VARS.F90:4-5 have no capacity checks).

Each case is driven through the C ABI on the CPU oracle (CPU tests) and on
the HIP library (GPU tests): a marker through a wall gives ERR_OUTSIDE, a
candidate list longer than maxCollisionNo gives ERR_CAPACITY, a non-finite
velocity or marker position gives ERR_NAN. The kernels clamp every index
they derive from such states, so none of these cases touches memory out of
bounds; the status is read from device flags after the call."""
import numpy as np
import pytest

from ibmdem import abi, cases
from ibmdem.abi import ERR_CAPACITY, ERR_NAN, ERR_NOT_CONVERGED, ERR_OUTSIDE, IbmdemError, Simulation


def _libs(request):
    return request.getfixturevalue("gpu_lib" if request.param == "gpu" else "oracle_lib")


@pytest.fixture(params=["cpu", pytest.param("gpu", marks=pytest.mark.gpu)])
def lib(request):
    return _libs(request)


def test_marker_uploaded_outside_a_wall(lib):
    setup = cases.small_walls()
    setup.markers.x[2, 5] = -0.25   # below the bottom wall
    with Simulation(lib, setup.case) as s:
        s.upload_bodies(setup.bodies)
        with pytest.raises(IbmdemError) as e:
            s.upload_markers(setup.markers)
        assert e.value.code == ERR_OUTSIDE


def test_body_driven_through_a_wall(lib):
    """The lower sphere gets a downward velocity that carries its markers
    through the bottom wall within one DEM sub-step."""
    setup = cases.small_walls()
    b = setup.bodies
    dtd = setup.case.dt / setup.case.couple_interval
    b.v[:, 0] = (0.0, 0.0, -8.0 / dtd)
    b.vold[:, 0] = b.v[:, 0]
    setup.markers.v[:, setup.markers.body == 0] = b.v[:, :1]
    with Simulation(lib, setup.case) as s:
        cases.load(s, setup)
        with pytest.raises(IbmdemError) as e:
            s.dem_substep()
        assert e.value.code == ERR_OUTSIDE


def test_candidate_list_beyond_max_pairs(lib):
    """maxCollisionNo = 1 in a cluster whose marker shells are inside the
    candidate distance: the list of some marker is longer."""
    setup = cases.small_periodic(gap=0.6)
    setup.case.max_pairs = 1
    with Simulation(lib, setup.case) as s:
        s.upload_bodies(setup.bodies)
        with pytest.raises(IbmdemError) as e:
            # the oracle builds its lists at upload, the GPU library scans
            # candidates in the contact kernel
            s.upload_markers(setup.markers)
            s.upload_fields(*setup.fields)
            s.contact()
        assert e.value.code == ERR_CAPACITY


def test_pair_listing_beyond_the_callers_capacity(lib):
    setup = cases.small_periodic(gap=0.6)
    with Simulation(lib, setup.case) as s:
        cases.load(s, setup)
        cnt, _, _ = s.pairs()
        assert cnt.max() > 1
        with pytest.raises(IbmdemError) as e:
            s.pairs(cap=1)
        assert e.value.code == ERR_CAPACITY
        # the context stays usable after the listing error
        s.step()


def test_nan_in_the_velocity(lib):
    setup = cases.small_walls()
    setup.case.poisson_max_iter = 20
    u, v, w, p = (f.copy() for f in setup.fields)
    u[3, 4, 5] = np.nan
    with Simulation(lib, setup.case) as s:
        s.upload_bodies(setup.bodies)
        s.upload_markers(setup.markers)
        s.upload_fields(u, v, w, p)
        with pytest.raises(IbmdemError) as e:
            s.step()
        assert e.value.code == ERR_NAN


def test_nan_in_a_marker_position(lib):
    setup = cases.small_walls()
    setup.markers.x[0, 7] = np.nan
    with Simulation(lib, setup.case) as s:
        s.upload_bodies(setup.bodies)
        with pytest.raises(IbmdemError) as e:
            s.upload_markers(setup.markers)
        assert e.value.code == ERR_NAN


def test_nan_in_a_body_state(lib):
    """A non-finite body velocity makes the Newton-Euler update and the
    markers that follow it non-finite: the sub-step reports ERR_NAN."""
    setup = cases.small_walls()
    setup.bodies.v[0, 1] = np.inf
    with Simulation(lib, setup.case) as s:
        cases.load(s, setup)
        with pytest.raises(IbmdemError) as e:
            s.dem_substep()
        assert e.value.code == ERR_NAN


def _batch_setups():
    return [cases.small_periodic(gap=0.75), cases.small_walls(), cases.small_open(axis=2)]


def test_batch_failure_leaves_the_other_cases_complete(lib):
    """Case 1 cannot converge (one CG iteration allowed): the batch reports
    it, and every other case has completed the step exactly as when it is
    stepped alone."""
    setups = _batch_setups()
    setups[1].case.poisson_max_iter = 1
    with abi.Batch(lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
        with pytest.raises(IbmdemError) as e:
            b.run(2)
        assert e.value.code == ERR_NOT_CONVERGED
        got = {q: (b[q].download_fields(), b[q].download_markers().x) for q in (0, 2)}
    for q in (0, 2):
        with Simulation(lib, setups[q].case) as s:
            cases.load(s, setups[q])
            s.step()
            want = (s.download_fields(), s.download_markers().x)
        for a, c in zip(got[q][0], want[0]):
            np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(got[q][1], want[1])


def test_batch_failure_inside_a_batched_group(lib):
    """A case that fails in a step whose solve runs batched (a NaN uploaded
    into its velocity after two good steps): the batch reports ERR_NAN and
    the other cases of its group complete that step exactly as when they
    are stepped alone."""
    setups = _batch_setups()
    with abi.Batch(lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
        b.run(2)
        u, v, w, p = b[1].download_fields()
        u[3, 4, 5] = np.nan
        b[1].upload_fields(u, v, w, p)
        with pytest.raises(IbmdemError) as e:
            b.run(1)
        assert e.value.code == ERR_NAN
        got = {q: (b[q].download_fields(), b[q].download_markers().x) for q in (0, 2)}
    for q in (0, 2):
        with Simulation(lib, setups[q].case) as s:
            cases.load(s, setups[q])
            s.run(3)
            want = (s.download_fields(), s.download_markers().x)
        for a, c in zip(got[q][0], want[0]):
            np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(got[q][1], want[1])
