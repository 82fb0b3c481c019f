"""Batching form (rpfvm_batch_*) against the single-problem device path.

Each problem of a batch must evolve exactly as an rpfvm_ctx with the same
configuration driven to the same stop time: the batch shares the face and cell
arithmetic, so its interior is compared bit for bit (FP64 and mixed form), over
mixed solvers, uneven sizes, every boundary type and each launch shape the
autotuner may choose.  The FP64 single-problem path is itself pinned to the
reference by the parity suite.
"""
import numpy as np
import pytest

from overheatingproblem_amd import cases
from overheatingproblem_amd.capi import STAGING_DIRECT, STAGING_LDS, STAGING_SWEEPS, Batch, Solver
from overheatingproblem_amd.config import (BC_FIXED_VALUE as BC_FIXED, BC_INFLOW, BC_OUTFLOW,
                                           BC_PERIODICITY as BC_PERIODIC,
                                           BC_REFLECTION as BC_REFLECT, DEBUG_LARGE_ARRAY_FORMS,
                                           PRECISION_MIXED)

pytestmark = pytest.mark.gpu
GC = 4


def single(cfg, U0, t_end):
    with Solver(cfg) as s:
        s.set_state(U0)
        t, n = s.advance_to(t_end, 100000)
        return s.get_state(), n, s.sync()


def interior(U):
    return U[:, GC:-GC, GC:-GC]


def small_batch(precision=0, characteristic=True, seed=21):
    """Six uneven problems with all solvers and boundary types (seeds whose problems the
    reference algorithm keeps finite and positive to t = 0.05; pinned on the CPU by
    test_batch_cases_cpu.py)."""
    rng = np.random.default_rng(seed)
    shapes = [(70, 51), (33, 129), (64, 64), (100, 17), (9, 40), (127, 66)]
    schemes = [2, 1, 3, 10, 2, 3]
    bcs = [(BC_OUTFLOW,) * 4, (BC_REFLECT,) * 4, (BC_PERIODIC,) * 4,
           (BC_INFLOW, BC_OUTFLOW, BC_REFLECT, BC_FIXED), (BC_FIXED, BC_REFLECT, BC_OUTFLOW, BC_PERIODIC),
           (BC_REFLECT, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC)]
    cfgs, states = [], []
    for (nx, ny), sch, bc in zip(shapes, schemes, bcs):
        quad = {k: (float(rng.uniform(0.3, 2.0)), float(rng.uniform(-0.8, 0.8)),
                    float(rng.uniform(-0.8, 0.8)), float(rng.uniform(0.2, 2.0)))
                for k in ("NE", "NW", "SW", "SE")}
        c, U = cases.four_quadrant(nx, ny, quad, x0=0.43, y0=0.57, perturb=1e-3, seed=int(nx),
                                   scheme=sch, characteristic=characteristic, bc=bc, cfl=0.45)
        c.precision = precision
        c.inflow_state[:] = [1.0, 0.5, -0.2, 3.0]
        for s in range(4):
            c.fixed_state[s][:] = [0.9 + 0.1 * s, 0.1, 0.2, 2.5]
        cfgs.append(c)
        states.append(U)
    return cfgs, states


# seeds whose problems the reference algorithm keeps finite and positive to ORDER_T_END at
# each reconstruction order (pinned on the CPU by test_batch_cases_cpu.py; seed 20 at
# WENO3 is not: its reflective HLL box blows up in the reference arithmetic too)
ORDER_SEEDS = {1: 18, 2: 19, 3: 21}
ORDER_T_END = 0.03


def order_batch(precision, order):
    """small_batch at a lower reconstruction order (TVD component-wise, the others
    characteristic)."""
    cfgs, states = small_batch(precision, order != 2, seed=ORDER_SEEDS[order])
    for c in cfgs:
        c.recon_order = order
    return cfgs, states


HC_SCHEMES = (11, 22, 33, 22, 33, 11)

# states of the seeded generator away from near vacuum (rho >= 0.3, p >= 0.2, |u|, |v| <= 1)
NARROW_RANGES = {"rho": (0.3, 2.0), "p": (0.2, 2.0), "uv": (-1.0, 1.0)}


def hc_batch(characteristic, precision=0):
    """small_batch with the heat-capacity solvers HLLHC/HLLCHC/RoeHC."""
    cfgs, states = small_batch(precision, characteristic)
    for c, sch in zip(cfgs, HC_SCHEMES):
        c.scheme = sch
    return cfgs, states


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_SWEEPS])
@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
@pytest.mark.parametrize("characteristic", [True, False])
def test_batch_matches_single_problems(precision, characteristic, staging):
    cfgs, states = small_batch(precision, characteristic)
    t_end = 0.05
    with Batch(cfgs) as b:
        b.set_tuning(staging=staging)
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.advance_to(t_end, 100000)
        t, steps, faults = b.status()
        got = [b.get_state(k) for k in range(len(cfgs))]
    report = []
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        ref, n, f = single(c, U0, t_end)
        diff = float(np.abs(interior(got[k]) - interior(ref)).max())
        report.append((k, c.nx, c.ny, c.scheme, tuple(c.bc), int(faults[k]) & 3, f & 3,
                       int(steps[k]), n, float(t[k]), diff))
    bad = [r for r in report if r[5] != r[6] or r[7] != r[8] or r[9] != t_end or r[10] != 0.0]
    assert not bad, "\n".join(map(str, bad))

@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_LDS, STAGING_SWEEPS])
@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_batch_reconstruction_orders_match_single_problems(order, precision, staging):
    """Godunov, TVD and WENO3 batches: each problem bit-identical to running it alone."""
    cfgs, states = order_batch(precision, order)
    t_end = ORDER_T_END
    with Batch(cfgs) as b:
        b.set_tuning(staging=staging)
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.advance_to(t_end, 100000)
        t, steps, faults = b.status()
        got = [b.get_state(k) for k in range(len(cfgs))]
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        ref, n, f = single(c, U0, t_end)
        assert (int(faults[k]) & 3, int(steps[k]), float(t[k])) == (f & 3, n, t_end), k
        assert np.isfinite(interior(ref)).all(), k
        assert np.array_equal(interior(got[k]), interior(ref)), k


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_LDS, STAGING_SWEEPS])
@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
@pytest.mark.parametrize("characteristic", [True, False])
def test_batch_heat_capacity_solvers_match_single_problems(characteristic, precision, staging):
    """The heat-capacity solvers in FP64 and mixed batches, each problem bit-identical to
    an rpfvm_ctx of the same form (the mixed batch's FP64 fix-ups included: fault bit 4)."""
    cfgs, states = hc_batch(characteristic, precision)
    t_end = ORDER_T_END
    with Batch(cfgs) as b:
        b.set_tuning(staging=staging)
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.advance_to(t_end, 100000)
        t, steps, faults = b.status()
        got = [b.get_state(k) for k in range(len(cfgs))]
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        ref, n, f = single(c, U0, t_end)
        assert (int(faults[k]) & 3, int(steps[k]), float(t[k])) == (f & 3, n, t_end), k
        assert np.isfinite(interior(ref)).all(), k
        assert np.array_equal(interior(got[k]), interior(ref)), k


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_LDS, STAGING_SWEEPS])
@pytest.mark.parametrize("rows,waves", [(4, 1), (16, 8), (64, 3)])
def test_batch_launch_shapes_agree(staging, rows, waves):
    cfgs, states = small_batch(0, True, seed=11)
    t_end = 0.02
    outs = []
    for shape in ((STAGING_DIRECT, 32, 8), (staging, rows, waves)):
        with Batch(cfgs) as b:
            b.set_tuning(staging=shape[0], rows=shape[1], waves_per_cu=shape[2])
            for k, U in enumerate(states):
                b.set_state(k, U)
            b.advance_to(t_end, 100000)
            outs.append([interior(b.get_state(k)) for k in range(len(cfgs))])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
def test_batch_sweeps_large_array_forms_agree(precision):
    """The two-sweep schedule's y sweep staged through LDS with buffer addressing (used
    when every problem's arrays are below 4 GB) and its global-address form (chosen when
    one problem carries RPFVM_DEBUG_LARGE_ARRAY_FORMS) give the same states."""
    cfgs, states = small_batch(precision, True, seed=11)
    t_end = 0.02
    outs = []
    for flag in (0, DEBUG_LARGE_ARRAY_FORMS):
        cs = [c.copy() for c in cfgs]
        cs[3].debug_flags = flag
        with Batch(cs) as b:
            b.set_tuning(staging=STAGING_SWEEPS, rows=16, waves_per_cu=16)
            for k, U in enumerate(states):
                b.set_state(k, U)
            b.advance_to(t_end, 100000)
            outs.append([interior(b.get_state(k)) for k in range(len(cs))])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_SWEEPS])
def test_batch_graph_replay_matches_eager(staging):
    """Graph replays with host work and synchronisation in between give the eager
    launches' results at every step."""
    cfgs, states = small_batch(0, True, seed=3)
    res = {}
    for use_graph in (0, 1):
        with Batch(cfgs) as b:
            b.set_tuning(use_graph=use_graph, staging=staging)
            for k, U in enumerate(states):
                b.set_state(k, U)
            snaps = []
            for _ in range(5):
                b.step(1.0)
                snaps.append([interior(b.get_state(k)).copy() for k in range(len(cfgs))])
                b.status()
            res[use_graph] = snaps
    for s0, s1 in zip(res[0], res[1]):
        for a, c in zip(s0, s1):
            assert np.array_equal(a, c)


def test_batch_tuner_keeps_state_and_picks_a_valid_shape():
    cfgs, states = small_batch(0, True, seed=8)
    t_end = 0.03
    with Batch(cfgs) as b:
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.step(t_end)
        tuned = b.tune()
        assert tuned["ms_per_stage"] > 0 and tuned["staging"] in (0, 1, 2)
        b.advance_to(t_end, 100000)
        got = [interior(b.get_state(k)) for k in range(len(cfgs))]
    with Batch(cfgs) as b:
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.advance_to(t_end, 100000)
        ref = [interior(b.get_state(k)) for k in range(len(cfgs))]
    for a, c in zip(got, ref):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_SWEEPS])
def test_batch_rk_orders(staging):
    for order in (1, 2):
        cfgs, states = small_batch(0, True, seed=order)
        for c in cfgs:
            c.time_order = order
        with Batch(cfgs) as b:
            b.set_tuning(staging=staging)
            for k, U in enumerate(states):
                b.set_state(k, U)
            b.advance_to(0.01, 100000)
            got = [b.get_state(k) for k in range(len(cfgs))]
        for k, (c, U0) in enumerate(zip(cfgs, states)):
            ref, _, _ = single(c, U0, 0.01)
            assert np.array_equal(interior(got[k]), interior(ref)), (order, k)


@pytest.mark.parametrize("staging", [STAGING_DIRECT, STAGING_SWEEPS])
@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
def test_batch_seeded_workload_prefix(precision, staging):
    """The first problems of the seeded 64-problem generator (sizes 256..1536) with
    states kept clear of near vacuum and no safeguard (so that every schedule runs them),
    a few steps: identical to the single-problem path."""
    cfgs, states = cases.batch_case(count=64, positivity=False, ranges=NARROW_RANGES)
    pick = [0, 1, 2, 3]
    cfgs = [cfgs[k] for k in pick]
    states = [states[k] for k in pick]
    for c in cfgs:
        c.precision = precision
    with Batch(cfgs) as b:
        b.set_tuning(staging=staging)
        for k, U in enumerate(states):
            b.set_state(k, U)
        for _ in range(4):
            b.step(0.25)
        t, steps, _ = b.status()
        got = [interior(b.get_state(k)) for k in range(len(cfgs))]
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        with Solver(c) as s:
            s.set_state(U0)
            for _ in range(4):
                s.time_advance(s.get_delta_t())
            assert np.array_equal(got[k], interior(s.get_state()))
        assert steps[k] == 4


# ----------------------------------------------------------- positivity safeguard
def wide_batch(precision, seed=32, time_order=3):
    """small_batch with the positivity safeguard and the seeded workload's state ranges
    (rho ~ U[0.1,2], p ~ U[0.02,2], u, v ~ U[-1.5,1.5]), which reach near vacuum: the
    safeguard repairs cells in several of these problems (pinned on the CPU by
    test_batch_cases_cpu.py)."""
    rng = np.random.default_rng(seed)
    shapes = [(70, 51), (33, 129), (64, 64), (100, 17), (9, 40), (127, 66)]
    schemes = [1, 2, 3, 1, 10, 1]
    bcs = [(BC_OUTFLOW,) * 4, (BC_REFLECT,) * 4, (BC_PERIODIC,) * 4,
           (BC_INFLOW, BC_OUTFLOW, BC_REFLECT, BC_FIXED), (BC_FIXED, BC_REFLECT, BC_OUTFLOW, BC_PERIODIC),
           (BC_REFLECT, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC)]
    cfgs, states = [], []
    for (nx, ny), sch, bc in zip(shapes, schemes, bcs):
        quad = {k: (float(rng.uniform(0.1, 2.0)), float(rng.uniform(-1.5, 1.5)),
                    float(rng.uniform(-1.5, 1.5)), float(rng.uniform(0.02, 2.0)))
                for k in ("NE", "NW", "SW", "SE")}
        c, U = cases.four_quadrant(nx, ny, quad, x0=0.43, y0=0.57, perturb=1e-3, seed=int(nx),
                                   scheme=sch, characteristic=True, bc=bc, cfl=0.45,
                                   positivity=True, time_order=time_order)
        c.precision = precision
        c.inflow_state[:] = [1.0, 0.5, -0.2, 3.0]
        for s in range(4):
            c.fixed_state[s][:] = [0.9 + 0.1 * s, 0.1, 0.2, 2.5]
        cfgs.append(c)
        states.append(U)
    return cfgs, states


WIDE_T_END = 0.04


@pytest.mark.parametrize("time_order", [3, 2, 1])
@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
def test_batch_safeguard_matches_single_problems(precision, time_order):
    """With the positivity safeguard each problem is bit-identical to the safeguarded
    single-problem path (same dts, step counts and fault bits), RK1 (copy back), RK2 and
    RK3 (last stage into the fourth pool, swapped in)."""
    cfgs, states = wide_batch(precision, time_order=time_order)
    with Batch(cfgs) as b:
        assert b.get_tuning()["staging"] == STAGING_SWEEPS
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.advance_to(WIDE_T_END, 100000)
        t, steps, faults = b.status()
        got = [b.get_state(k) for k in range(len(cfgs))]
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        print(f"problem {k}: {c.nx}x{c.ny} scheme {c.scheme} order {time_order}", flush=True)
        ref, n, f = single(c, U0, WIDE_T_END)
        assert np.isfinite(interior(ref)).all(), k
        assert (int(faults[k]) & 9, int(steps[k]), float(t[k])) == (f & 9, n, WIDE_T_END), k
        assert np.array_equal(interior(got[k]), interior(ref)), k


@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
def test_batch_safeguard_graph_replay_matches_eager(precision):
    """The two step graphs (one per pool assignment) replayed with host reads in between
    give the eager launches' states at every step."""
    cfgs, states = wide_batch(precision)
    res = {}
    for use_graph in (0, 1):
        with Batch(cfgs) as b:
            b.set_tuning(use_graph=use_graph)
            for k, U in enumerate(states):
                b.set_state(k, U)
            snaps = []
            for _ in range(5):
                b.step(1.0)
                snaps.append([interior(b.get_state(k)).copy() for k in range(len(cfgs))])
                b.status()
            res[use_graph] = snaps
    for s0, s1 in zip(res[0], res[1]):
        for a, c in zip(s0, s1):
            assert np.array_equal(a, c)


def test_batch_safeguard_schedule_and_tuning():
    """The safeguard runs in the two-sweep schedule: the tuner keeps to it and a one-pass
    schedule is refused; tuning leaves the state alone."""
    from overheatingproblem_amd.capi import RpfvmError
    cfgs, states = wide_batch(0)
    with Batch(cfgs) as b:
        with pytest.raises(RpfvmError, match="two-sweep"):
            b.set_tuning(staging=STAGING_DIRECT)
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.step(WIDE_T_END)
        tuned = b.tune()
        assert tuned["staging"] == STAGING_SWEEPS
        b.advance_to(WIDE_T_END, 100000)
        got = [interior(b.get_state(k)) for k in range(len(cfgs))]
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        ref, _, _ = single(c, U0, WIDE_T_END)
        assert np.array_equal(got[k], interior(ref)), k


@pytest.mark.parametrize("precision", [PRECISION_MIXED, 0])
def test_batch_seeded_workload_all_problems(precision):
    """Every problem of the seeded 64-problem workload (rho ~ U[0.1,2], p ~ U[0.02,2],
    u, v ~ U[-1.5,1.5], sizes 256..1536, safeguard on) for 6 steps, bit-identical to the
    safeguarded single-problem path; two of its problems need cell repairs from the first
    steps on."""
    cfgs, states = cases.batch_case(count=64)
    for c in cfgs:
        c.precision = precision
    nsteps = 6
    with Batch(cfgs) as b:
        for k, U in enumerate(states):
            b.set_state(k, U)
        b.tune()
        for _ in range(nsteps):
            b.step(0.25)
        t, steps, faults = b.status()
        got = [interior(b.get_state(k)).copy() for k in range(len(cfgs))]
    print(f"batch done: steps {set(steps.tolist())}, fault bits {set(faults.tolist())}", flush=True)
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        with Solver(c) as s:
            s.set_state(U0)
            for _ in range(nsteps):
                s.time_advance(s.get_delta_t())
            ref = interior(s.get_state())
            f = s.sync()
        assert steps[k] == nsteps and (int(faults[k]) & 9) == (f & 9), k
        assert np.isfinite(ref).all(), k
        assert np.array_equal(got[k], ref), k
