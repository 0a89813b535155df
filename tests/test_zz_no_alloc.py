"""ibmdem_download_force_noalloc: the force download without device
allocation (a fixed device scratch array, filled one slab of z planes at a
time) must give exactly what ibmdem_download_force gives, on two calls in a
row, the second on a smaller grid (the scratch keeps the first call's
contents, which must not show through), and must reject a grid whose z plane
does not fit the scratch."""
import numpy as np
import pytest

from ibmdem import cases
from ibmdem.abi import ERR_UNSUPPORTED, IbmdemError, Simulation

pytestmark = pytest.mark.gpu


def _force_both(lib, setup, steps):
    with Simulation(lib, setup.case) as s:
        cases.load(s, setup)
        s.run(steps)
        s.begin_step()
        s.interpolate()
        s.spread()
        return s.download_force(), s.download_force_noalloc()


def test_noalloc_force_download_equals_download_force(gpu_lib):
    # first the 256^3 suspension (26 slabs through the scratch), then a
    # smaller grid in one slab
    for setup, steps in ((cases.config_suspension(), 2), (cases.small_walls(), 1)):
        want, got = _force_both(gpu_lib, setup, steps)
        for a, b in zip(got, want):
            assert np.abs(b).max() > 0
            np.testing.assert_array_equal(a, b)


def test_noalloc_force_download_rejects_a_plane_beyond_the_scratch(gpu_lib):
    case = cases.build_case(1024, 1024, 2, [0] * 6)
    with Simulation(gpu_lib, case) as s:
        with pytest.raises(IbmdemError) as e:
            s.download_force_noalloc()
        assert e.value.code == ERR_UNSUPPORTED
