"""The batched form: independent cases of uneven size advanced together.

A batch step must give each case exactly what stepping it alone gives (the
same kernels in the same order on the case's own stream); the autotuner may
change a case's multigrid cycle, which changes only the solver iterates
(the solution agrees to the solver tolerance), and must leave the state as
it found it."""
import numpy as np
import pytest

import oracle
from ibmdem import abi, cases


def _cases():
    return [cases.small_periodic(gap=0.75),
            cases.small_walls(),
            cases.small_periodic(n=(40, 32, 24), n_spheres=4, D=7.0, seed=cases.SEED_BASE + 21, gap=1.0),
            cases.small_open(axis=2)]


def _planes(s):
    return [s.download_outflow(sd) for sd in (1, 3, 5) if s.case.btype[sd] == abi.BC_OUTFLOW]


def _single(lib, setup, n):
    s = abi.Simulation(lib, setup.case)
    cases.load(s, setup)
    info = s.run(n)
    out = (s.download_fields(), s.download_markers().x, s.download_bodies().x, info.iterations)
    s.close()
    return out


def _batched(lib, setups, n, tune=0, fused=True, mg=None, solves=None, groups=1, tiles=None):
    with abi.Batch(lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
            if mg and q in mg:
                b[q].set_mg_params(*mg[q])
            if tiles and q in tiles:
                b[q].set_stencil_tile(tiles[q])
        b.set_fused(fused)
        b.set_solve_groups(groups)
        assert b.solve_groups() == groups
        if tune:
            before = [b[q].download_fields() + _planes(b[q]) for q in range(len(b))]
            b.autotune(tune)
            for q in range(len(b)):
                for x, y in zip(before[q], b[q].download_fields() + _planes(b[q])):
                    np.testing.assert_array_equal(x, y)
        infos = b.run(n)
        if solves is not None:
            solves.append(b.fused_solves())
        return [(b[q].download_fields(), b[q].download_markers().x, b[q].download_bodies().x,
                 infos[q].iterations, b[q].mg_params(), b[q].stencil_tile()) for q in range(len(b))]


def test_mg_params_round_trip_and_range(oracle_lib):
    setup = cases.small_walls()
    s = abi.Simulation(oracle_lib, setup.case)
    assert s.mg_params() == (2, 2, 24)
    s.set_mg_params(3, 1, 12)
    assert s.mg_params() == (3, 1, 12)
    with pytest.raises(abi.IbmdemError):
        s.set_mg_params(0, 1, 12)
    s.close()


def test_batch_equals_single_cases_on_the_oracle(oracle_lib):
    setups = _cases()
    got = _batched(oracle_lib, setups, 2)
    for s, g in zip(setups, got):
        want = _single(oracle_lib, s, 2)
        for a, b in zip(g[0], want[0]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(g[1], want[1])
        np.testing.assert_array_equal(g[2], want[2])


def _assert_same(got, want):
    for a, b in zip(got[0], want[0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[2], want[2])
    assert got[3] == want[3]


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [1, 3])
def test_batch_equals_single_cases_on_the_gpu(gpu_lib, groups):
    """The solves run batched (one launch per solver operation over all
    cases, or over each of `groups` groups) from the second step on; every
    case still gets exactly what it gets alone."""
    setups = _cases()
    solves = []
    got = _batched(gpu_lib, setups, 4, solves=solves, groups=groups)
    assert solves[0] >= 3 * groups   # steps 2..4 (the first learns the iteration counts)
    for s, g in zip(setups, got):
        _assert_same(g, _single(gpu_lib, s, 4))


@pytest.mark.gpu
def test_batched_solve_equals_per_case_solves(gpu_lib):
    """ibmdem_batch_set_fused(0) runs each case's step on its own stream,
    (2) batches only the solve, (3) the solve and the post-solve part, (4)
    the forcing loop as well, (1) the whole step; all give the same bits."""
    setups = _cases()
    solves = []
    a = _batched(gpu_lib, setups, 3, fused=0, solves=solves)
    assert solves[0] == 0
    for mode in (2, 3, 4, 1):
        for x, y in zip(a, _batched(gpu_lib, setups, 3, fused=mode)):
            _assert_same(x, y)


def _spaced(setup, h):
    """The setup on a grid of spacing h (not a power of two): positions
    scaled with the grid, so every marker keeps its cell."""
    c = setup.case
    c.dx = c.dy = c.dz = float(h)
    setup.markers.x *= h
    setup.bodies.x *= h
    return setup


@pytest.mark.gpu
def test_batched_first_part_variants(gpu_lib):
    """The whole step batched over cases that take different kernel variants
    of its first part: bulk forcing (periodic), a grid spacing that is not a
    power of two (closed box), no particles at all (open channel), and the
    four predictor tile schedules; each case equals stepping it alone with
    the same schedule."""
    setups = [cases.small_periodic(gap=0.75),
              _spaced(cases.small_walls(), 1.25),
              cases.small_open(axis=0, spheres=False),
              cases.small_periodic(n=(40, 32, 24), n_spheres=4, D=7.0, seed=cases.SEED_BASE + 21, gap=1.0)]
    tiles = {0: 1, 1: 2, 2: 3, 3: 0}
    solves = []
    got = _batched(gpu_lib, setups, 4, solves=solves, tiles=tiles)
    assert solves[0] >= 3
    for q, (s, g) in enumerate(zip(setups, got)):
        with abi.Simulation(gpu_lib, s.case) as one:
            cases.load(one, s)
            one.set_stencil_tile(tiles[q])
            info = one.run(4)
            want = (one.download_fields(), one.download_markers().x, one.download_bodies().x, info.iterations)
        _assert_same(g, want)


@pytest.mark.gpu
def test_batched_post_part_with_probes(gpu_lib):
    """The post-solve part batched with probe rings recorded for some cases:
    the rings and the state equal stepping each case alone."""
    setups = _cases()
    probes = {0: [0, 1], 2: [3]}
    steps = 4
    with abi.Batch(gpu_lib, [s.case for s in setups]) as b:
        for q, s in enumerate(setups):
            cases.load(b[q], s)
            if q in probes:
                b[q].probe_bodies(probes[q], 8)
        b.run(steps)
        assert b.fused_solves() >= steps - 1
        got = [(b[q].download_bodies().x, b[q].probe_read() if q in probes else None) for q in range(len(b))]
    for q, s in enumerate(setups):
        with abi.Simulation(gpu_lib, s.case) as one:
            cases.load(one, s)
            if q in probes:
                one.probe_bodies(probes[q], 8)
            one.run(steps)
            np.testing.assert_array_equal(got[q][0], one.download_bodies().x)
            if q in probes:
                ring_b, ring_1 = got[q][1], one.probe_read()
                assert ring_b[0] == ring_1[0]
                np.testing.assert_array_equal(ring_b[1], ring_1[1])


@pytest.mark.gpu
def test_batched_mixed_precision_cases(gpu_lib):
    """Mixed-precision cases (FP32 V-cycle, FP64 CG) batched, with uneven
    grids whose V-cycles turn at different levels."""
    setups = [cases.small_periodic(gap=0.75, precision=abi.MIXED),
              cases.small_walls(precision=abi.MIXED),
              cases.small_periodic(n=(64, 48, 40), n_spheres=5, D=8.0, seed=cases.SEED_BASE + 22, gap=1.0,
                                   precision=abi.MIXED),
              cases.small_open(axis=2, precision=abi.MIXED)]
    solves = []
    got = _batched(gpu_lib, setups, 4, solves=solves)
    assert solves[0] >= 3
    for s, g in zip(setups, got):
        _assert_same(g, _single(gpu_lib, s, 4))


@pytest.mark.gpu
def test_batched_groups_with_different_cycles(gpu_lib):
    """Cases with different smoothing counts solve in separate groups; cases
    of one group may differ in their coarsest-level sweeps."""
    setups = _cases()
    mg = {0: (1, 2, 12), 1: (2, 2, 12), 3: (1, 2, 24)}
    solves = []
    got = _batched(gpu_lib, setups, 3, mg=mg, solves=solves)
    assert solves[0] >= 4   # two groups, steps 2 and 3
    for q, (s, g) in enumerate(zip(setups, got)):
        with abi.Simulation(gpu_lib, s.case) as one:
            cases.load(one, s)
            if q in mg:
                one.set_mg_params(*mg[q])
            info = one.run(3)
            want = (one.download_fields(), one.download_markers().x, one.download_bodies().x, info.iterations)
        _assert_same(g, want)


@pytest.mark.gpu
def test_batch_autotune_keeps_the_solution(gpu_lib, oracle_lib):
    setups = _cases()
    tuned = _batched(gpu_lib, setups, 3, tune=2)
    for s, g in zip(setups, tuned):
        want = _single(oracle_lib, s, 3)
        assert g[4] in [(1, 1, 24), (2, 2, 24), (3, 3, 24), (2, 2, 12), (1, 2, 24)]
        assert g[5] in (0, 1, 2, 3)
        for a, b in zip(g[0], want[0]):
            scale = max(np.abs(b).max(), 1e-300)
            assert np.abs(a - b).max() <= 1e-8 * scale
        assert np.abs(g[2] - want[2]).max() <= 1e-10 * np.abs(want[2]).max()
