"""GPU: ensembles of independent instances (gsk_ensemble_*) against the members stepped alone.

Bar: bit-exact. An ensemble launches the same kernels with the same arguments as
gsk_step on each member, whatever the schedule (sequential, concurrent streams,
replayed hipGraph), and the noise counter lives on the device, so every member's
fields must equal those of an independently stepped handle bit for bit. One member
is also checked against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

import gsk
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


C = gsk.cubic_modulus(0.7, 0.3, 0.2)
V = 3


def member_configs():
    """Mixed sizes and parameters, noise on (exercises the device step counter)."""
    cfgs = []
    spec = [(32, 0.15, 0.6, 0.005), (64, 0.2, 1.0, 0.01), (96, 0.25, 1.4, 0.008), (32, 0.18, 0.9, 0.007),
            (64, 0.22, 1.2, 0.006)]
    for i, (n, A, kap, dt) in enumerate(spec):
        cfgs.append(gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, A=A, kappa=kap, dt=dt, noise=0.003, seed=100 + i,
                                       w_el=5.0))
    return cfgs


def init_members(sims):
    for i, s in enumerate(sims):
        s.init_uniform("eta", -0.05, 0.05, 1000 + i)


def alone(cfgs, steps):
    out = []
    for i, c in enumerate(cfgs):
        with gsk.Sim(c) as s:
            s.build_btensor(C, gsk.mart_eigenstrains(V))
            s.init_uniform("eta", -0.05, 0.05, 1000 + i)  # as init_members does for member i
            s.step(steps)
            out.append(s.get_field("eta"))
    return out


@pytest.mark.parametrize("mode,branches,spg", [(gsk.SCHED_SEQUENTIAL, 1, 1), (gsk.SCHED_STREAMS, 2, 1),
                                               (gsk.SCHED_GRAPH, 2, 3), (gsk.SCHED_GRAPH, 5, 1),
                                               (gsk.SCHED_GRAPH, 1, 4), (gsk.SCHED_BATCHED, 2, 1),
                                               (gsk.SCHED_BATCHED_GRAPH, 2, 3), (gsk.SCHED_BATCHED_GRAPH, 1, 4)])
def test_ensemble_bit_exact_vs_alone(gpu, mode, branches, spg):
    say(f"ensemble mode={mode} branches={branches} steps/graph={spg}")
    cfgs = member_configs()
    steps = 7
    want = alone(cfgs, steps)
    with gsk.Ensemble(cfgs) as e:
        e.build_btensor(C, gsk.mart_eigenstrains(V))
        init_members(e.members)
        e.set_schedule(mode, branches, spg)
        e.step(4)
        e.step(3)  # a second call reuses the graph and runs the remainder eagerly
        for i, s in enumerate(e.members):
            assert s.step_index == steps
            assert np.array_equal(s.get_field("eta").view(np.uint32), want[i].view(np.uint32)), f"member {i}"


@pytest.mark.parametrize("mode", [gsk.SCHED_BATCHED, gsk.SCHED_BATCHED_GRAPH])
def test_ensemble_batched_with_large_members(gpu, mode):
    """Batched schedules share launches only among small grids; members above 96^3 run on the
    branch streams beside the batched classes. Both kinds must match members stepped alone."""
    say(f"ensemble batched mode={mode} with 128^3 members")
    spec = [(64, 0.2, 1.0, 0.01), (128, 0.18, 0.8, 0.006), (64, 0.22, 1.2, 0.008), (128, 0.25, 1.3, 0.007)]
    cfgs = [gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, A=A, kappa=k, dt=dt, noise=0.003, seed=200 + i, w_el=5.0)
            for i, (n, A, k, dt) in enumerate(spec)]
    steps = 5
    want = alone(cfgs, steps)
    with gsk.Ensemble(cfgs) as e:
        e.build_btensor(C, gsk.mart_eigenstrains(V))
        init_members(e.members)
        e.set_schedule(mode, 2, 2)
        e.step(steps)
        for i, s in enumerate(e.members):
            assert s.step_index == steps
            assert np.array_equal(s.get_field("eta").view(np.uint32), want[i].view(np.uint32)), f"member {i}"


def test_ensemble_autotune_restores_state(gpu):
    say("ensemble autotune")
    cfgs = member_configs()
    with gsk.Ensemble(cfgs) as e:
        e.build_btensor(C, gsk.mart_eigenstrains(V))
        init_members(e.members)
        e.step(2)
        before = [s.get_field("eta") for s in e.members]
        table = e.autotune(trial_steps=2)
        modes = {int(r[0]) for r in table}
        assert modes == {gsk.SCHED_SEQUENTIAL, gsk.SCHED_STREAMS, gsk.SCHED_GRAPH, gsk.SCHED_BATCHED,
                         gsk.SCHED_BATCHED_GRAPH}
        assert all(r[3] > 0 for r in table)
        best = min(table, key=lambda r: r[3])
        m, b, k = e.schedule()
        assert (m, k) == (int(best[0]), int(best[2]))
        for i, s in enumerate(e.members):
            assert s.step_index == 2
            assert np.array_equal(s.get_field("eta"), before[i])
        e.step(5)
        want = alone(cfgs, 7)
        for i, s in enumerate(e.members):
            assert np.array_equal(s.get_field("eta").view(np.uint32), want[i].view(np.uint32)), f"member {i}"


def test_ensemble_member_matches_oracle(gpu):
    say("ensemble member 0 vs FP64 oracle")
    cfgs = member_configs()
    steps = 20
    with gsk.Ensemble(cfgs) as e:
        e.build_btensor(C, gsk.mart_eigenstrains(V))
        init_members(e.members)
        eta0 = e.members[0].get_field("eta").astype(np.float64)
        e.set_schedule(gsk.SCHED_GRAPH, 3, 5)
        e.step(steps)
        got = e.members[0].get_field("eta", np.float64)
    c = cfgs[0]
    p = O.mart_params(A=c.A, B=c.B, C=c.C, kappa=c.kappa, L=c.L, dt=c.dt, w_el=c.w_el, noise=c.noise, dx=c.dx,
                      kconv=c.kconv, seed=c.seed)
    want = O.mart_run(eta0, O.btensor(c.n, C, gsk.mart_eigenstrains(V)), p, steps)
    err = float(np.linalg.norm((got - want).ravel()) / np.linalg.norm(want.ravel()))
    assert err < 2e-5, err


def test_ensemble_rejects_plain_path(gpu):
    cfg = gsk.default_config(gsk.PHYS_MART, n=16, nvar=2, path=gsk.PATH_PLAIN64)
    with pytest.raises(gsk.GskError):
        gsk.Ensemble([cfg])


@pytest.mark.parametrize("mode", ["batched_graph", "autotuned"])
def test_ensemble_config3_sizes_bit_exact(gpu, mode):
    """The 96-member workload's three sizes in one ensemble: 2 x 64^3, 2 x 128^3 and 2 x 192^3
    (192 = 3 * 64 runs the radix-12 plans), with per-member A, kappa, dt and seeds as the
    workload draws them. Under the batched hipGraph schedule (64^3 members share launches, the
    others run on the branch streams) and under the schedule the autotuner picks, every member
    must equal itself stepped alone bit for bit; the first 64^3 member is also checked against
    the FP64 oracle."""
    say(f"ensemble 2x64^3 + 2x128^3 + 2x192^3, {mode}")
    spec = [(64, 0.15, 0.6, 0.005), (128, 0.2, 1.0, 0.01), (192, 0.25, 1.4, 0.008), (64, 0.18, 0.9, 0.007),
            (128, 0.22, 1.2, 0.006), (192, 0.16, 0.7, 0.009)]
    cfgs = [gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, A=A, kappa=k, dt=dt, noise=0.002, seed=300 + i, w_el=1.0)
            for i, (n, A, k, dt) in enumerate(spec)]
    steps = 7
    want = alone(cfgs, steps)
    with gsk.Ensemble(cfgs) as e:
        e.build_btensor(C, gsk.mart_eigenstrains(V))
        init_members(e.members)
        eta0 = e.members[0].get_field("eta").astype(np.float64)
        if mode == "autotuned":
            table = e.autotune(trial_steps=2)
            say(f"  autotuned schedule {e.schedule()} of {len(table)} candidates")
        else:
            e.set_schedule(gsk.SCHED_BATCHED_GRAPH, 4, 3)
        e.step(steps)
        for i, s in enumerate(e.members):
            assert s.step_index == steps
            assert np.array_equal(s.get_field("eta").view(np.uint32), want[i].view(np.uint32)), f"member {i}"
        got0 = e.members[0].get_field("eta", np.float64)
    c = cfgs[0]
    p = O.mart_params(A=c.A, B=c.B, C=c.C, kappa=c.kappa, L=c.L, dt=c.dt, w_el=c.w_el, noise=c.noise, dx=c.dx,
                      kconv=c.kconv, seed=c.seed)
    ref = O.mart_run(eta0, O.btensor(c.n, C, gsk.mart_eigenstrains(V)), p, steps)
    err = float(np.linalg.norm((got0 - ref).ravel()) / np.linalg.norm(ref.ravel()))
    assert err < 2e-6, err
