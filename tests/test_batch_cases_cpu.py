"""The problem sets of the batch GPU tests are stable under the reference algorithm:
every problem, run by the oracle (FP64, the reference's arithmetic), stays finite with
positive density and pressure to t = 0.05.  (Seed 5 of the same generator is not: its
reflective HLL box blows up in the reference arithmetic too.)"""
import numpy as np
import pytest

from overheatingproblem_amd.cases import cons_to_prim
from test_gpu_batch import (ORDER_SEEDS, ORDER_T_END, WIDE_T_END, hc_batch, order_batch, small_batch,
                            wide_batch)

SEEDS_USED = (21, 11, 3, 8, 1, 2)


def runs_clean(orc, c, U, t_end):
    U = U.copy()
    t = 0.0
    for _ in range(5000):
        if not t < t_end:
            break
        dt = orc.delta_t(c, U)
        if t + dt > t_end:
            dt = t_end - t
        if t + dt >= t_end:
            dt = t_end - t
        if not (np.isfinite(dt) and dt > 0 and t + dt > t):
            return False
        orc.time_advance(c, U, dt)
        t += dt
    rho, _, _, p = cons_to_prim(U[:, 4:-4, 4:-4], c.gamma)
    return bool(t == t_end and np.isfinite(U[:, 4:-4, 4:-4]).all() and (rho > 0).all() and (p > 0).all())


@pytest.mark.parametrize("seed", [21])
@pytest.mark.parametrize("characteristic", [True, False])
def test_default_batch_set_is_stable(oracle, seed, characteristic):
    cfgs, states = small_batch(0, characteristic, seed=seed)
    assert all(runs_clean(oracle, c, U, 0.05) for c, U in zip(cfgs, states))


def test_unstable_seed_is_detected(oracle):
    cfgs, states = small_batch(0, True, seed=5)
    assert not runs_clean(oracle, cfgs[1], states[1], 0.05)


@pytest.mark.parametrize("order", sorted(ORDER_SEEDS))
def test_order_batch_sets_are_stable(oracle, order):
    cfgs, states = order_batch(0, order)
    assert all(runs_clean(oracle, c, U, ORDER_T_END) for c, U in zip(cfgs, states))


@pytest.mark.parametrize("characteristic", [True, False])
def test_heat_capacity_batch_sets_are_stable(oracle, characteristic):
    cfgs, states = hc_batch(characteristic)
    assert all(runs_clean(oracle, c, U, ORDER_T_END) for c, U in zip(cfgs, states))


def test_batch_rejects_what_it_does_not_run():
    """Configuration checks happen before any device work."""
    from overheatingproblem_amd.capi import Batch, RpfvmError
    cfgs, _ = small_batch(0, True)
    cfgs[2].positivity = 1
    with pytest.raises(RpfvmError, match="must match problem 0"):
        Batch(cfgs)
    cfgs, _ = small_batch(0, True)
    cfgs[2].recon_order = 3
    with pytest.raises(RpfvmError, match="must match problem 0"):
        Batch(cfgs)
    cfgs, _ = small_batch(0, True)
    cfgs[1].time_order = 2
    with pytest.raises(RpfvmError, match="must match problem 0"):
        Batch(cfgs)


# Problems of the seeded 64-problem workload (rho ~ U[0.1,2], p ~ U[0.02,2],
# u, v ~ U[-1.5,1.5]) that leave the admissible set under the reference's arithmetic:
# index -> (n, scheme, step at which the unprotected run fails)
C4_FRAGILE = {57: (256, 3, 2), 62: (320, 1, 0)}


def admissible_state(c, U):
    rho, _, _, p = cons_to_prim(U[:, 4:-4, 4:-4], c.gamma)
    return bool(np.isfinite(U).all() and (rho > 0).all() and (p > 0).all())


@pytest.mark.parametrize("k", sorted(C4_FRAGILE))
def test_seeded_workload_needs_and_keeps_the_safeguard(oracle, k):
    """Without the safeguard the reference arithmetic reaches rho <= 0 or p <= 0 on
    these problems within a few steps; with it (first-order face fallback + HLLE cell
    repair) they stay admissible.  (Its conservation is pinned on a periodic box by
    test_oracle_positivity.py; here the quadrant states flow through the outflow walls.)"""
    from overheatingproblem_amd import cases
    cfgs, states = cases.batch_case(64)
    c, U0 = cfgs[k], states[k]
    n, scheme, fail_step = C4_FRAGILE[k]
    assert (c.nx, c.scheme, c.positivity) == (n, scheme, 1)
    bare = c.copy()
    bare.positivity = 0
    U = U0.copy()
    failed = None
    for s in range(fail_step + 1):
        oracle.time_advance(bare, U, oracle.delta_t(bare, U))
        if not admissible_state(bare, U):
            failed = s
            break
    assert failed == fail_step
    U = U0.copy()
    repairs = 0
    for _ in range(12):
        oracle.time_advance(c, U, oracle.delta_t(c, U))
        repairs += oracle.last_repairs()
        assert admissible_state(c, U)
    assert repairs > 0


def test_wide_batch_exercises_the_safeguard(oracle):
    """The safeguarded small batch of the GPU tests: with the safeguard every problem stays
    admissible to its end time, cells are repaired in at least three of them, and without
    it the reference arithmetic leaves the admissible set in those problems."""
    cfgs, states = wide_batch(0)
    repaired = []
    for k, (c, U0) in enumerate(zip(cfgs, states)):
        U = U0.copy()
        t, reps = 0.0, 0
        while t < WIDE_T_END:
            dt = oracle.delta_t(c, U)
            if t + dt >= WIDE_T_END:
                dt = WIDE_T_END - t
            oracle.time_advance(c, U, dt)
            reps += oracle.last_repairs()
            t += dt
        assert admissible_state(c, U), k
        if reps:
            repaired.append(k)
            bare = c.copy()
            bare.positivity = 0
            assert not runs_clean(oracle, bare, U0, WIDE_T_END), k
    assert len(repaired) >= 3, repaired
