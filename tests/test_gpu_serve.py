"""GPU tests of the continuous-batching server.

Bars: every request's final state is bit-identical whether it runs alone or
in a mix (direct requests share one batched launch, Barnes-Hut requests share
lanes); a Barnes-Hut request equals the library's own BH + leapfrog path bit
for bit; direct requests agree with the FP64 oracle to FP32 tolerance;
admission respects the HBM budget."""
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIELDS = ("x", "y", "z", "vx", "vy", "vz")

REQS = [  # (n, steps, seed)
    (1000, 5, 1), (3000, 8, 2), (20000, 3, 3), (700, 6, 4), (40000, 4, 5), (50000, 6, 6),
]


def systems(nb):
    return [nb.ic_plummer(n, seed) for n, _, seed in REQS]


def run_server(nb, items, stagger=0, budget=200e9, lanes=2, kdir=None, kbh=None, tick=None,
               order=None):
    srv = nb.Server(hbm_budget_bytes=budget, direct_threshold=32768, bh_lanes=lanes,
                    direct_steps_per_tick=kdir, bh_steps_per_tick=kbh, tick_budget=tick)
    ids = []
    pending = list(items)
    results = {}
    ticks = 0
    while pending or len(results) < len(items):
        if pending and (stagger == 0 or ticks % stagger == 0):
            sysd, steps = pending.pop(0)
            ids.append(srv.submit(sysd, steps))
            if stagger == 0:
                continue
        srv.tick()
        ticks += 1
        for rid in srv.poll():
            results[rid] = srv.result(rid)
            if order is not None:
                order.append(ids.index(rid))
        assert ticks < 10000
    return [results[i] for i in ids], srv


def test_same_alone_and_in_the_mix(nb):
    sy = systems(nb)
    items = [(s, steps) for s, (_, steps, _) in zip(sy, REQS)]
    mix, _ = run_server(nb, items)
    staggered, _ = run_server(nb, items, stagger=2)
    for k, it in enumerate(items):
        alone, _ = run_server(nb, [it])
        for f in FIELDS:
            assert np.array_equal(alone[0][f], mix[k][f]), (k, f)
            assert np.array_equal(alone[0][f], staggered[k][f]), (k, f)


def test_bh_request_equals_library_path(nb):
    n, steps, seed = REQS[4]
    s = nb.ic_plummer(n, seed)
    res, _ = run_server(nb, [(s, steps)])
    f = nb.BarnesHutForceCalculator(0.005, 0.5, 8, "quadrupole", G=1.0)
    p = nb.Particles.from_arrays(s["m"], s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
    it = nb.LeapfrogIntegrator()
    for _ in range(steps):
        it.step(f, p, 1e-4)
    g = p.download()
    for fl in FIELDS:
        assert np.array_equal(g[fl], res[0][fl]), fl


def test_direct_request_vs_oracle(nb, O):
    n, steps, seed = REQS[1]
    s = nb.ic_plummer(n, seed)
    res, _ = run_server(nb, [(s, steps)])
    ref = O.leapfrog_direct(s["m"], s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], 1e-4, steps,
                            1.0, 0.005 ** 2, False)
    # FP32 forces (relative error ~1e-6 of |a| <~ 100 here) integrated over
    # steps*dt = 8e-4 bound the velocity error by ~1e-7 and positions by ~1e-10.
    for fl, r in zip(FIELDS, ref[:6]):
        tol = 1e-10 if fl in ("x", "y", "z") else 1e-7
        assert np.abs(res[0][fl] - r).max() < tol, fl


def test_budget_admission(nb):
    sy = systems(nb)
    items = [(sy[0], 3), (sy[1], 3), (sy[3], 3)]
    # Room for the first two direct requests (96 kB + 288 kB of 432 kB): the third waits, then runs.
    budget = 3000 * 96 * 1.5
    res, srv = run_server(nb, items, budget=budget)
    alone = [run_server(nb, [it])[0][0] for it in items]
    for a, b in zip(alone, res):
        assert all(np.array_equal(a[f], b[f]) for f in FIELDS)
    st = srv.status()
    assert st["active"] == 0 and st["pending"] == 0 and st["bytes_in_use"] == 0
    too_big = nb.Server(hbm_budget_bytes=1e5)
    with pytest.raises(nb.NBError) as e:
        too_big.submit(sy[2], 1)
    assert e.value.code == 5  # NB_ERR_BUDGET


def test_admission_order_and_queueing(nb):
    """Every request is reported admitted exactly once, in submission order
    (admission is first come, first served), and before it completes.  The
    budget holds one of these direct requests at a time (96 B per particle:
    96 kB, 288 kB and 67 kB against 300 kB), so each waits for the one before
    it to finish."""
    sy = systems(nb)
    srv = nb.Server(hbm_budget_bytes=300e3, direct_threshold=32768, bh_lanes=1)
    ids = [srv.submit(sy[0], 3), srv.submit(sy[1], 3), srv.submit(sy[3], 2)]
    admitted, completed, tick_admitted, tick_done = [], [], {}, {}
    for tick in range(1, 100):
        srv.tick()
        for rid in srv.poll_admitted():
            admitted.append(rid)
            tick_admitted[rid] = tick
        for rid in srv.poll():
            assert rid in tick_admitted
            completed.append(rid)
            tick_done[rid] = tick
            srv.result(rid)
        if len(completed) == len(ids):
            break
    assert admitted == ids and sorted(completed) == sorted(ids)
    assert tick_admitted[ids[1]] >= tick_done[ids[0]] and tick_admitted[ids[2]] >= tick_done[ids[1]]
    assert srv.poll_admitted() == []


def test_direct_steps_per_tick_does_not_change_results(nb):
    """Direct requests take K steps per tick (bounded by the fewest steps any
    of them has left); every K gives the same final states, bit for bit, and
    K outside [1, 1024] is rejected."""
    sy = systems(nb)
    items = [(sy[0], 5), (sy[1], 8), (sy[3], 6), (sy[4], 4)]
    base, _ = run_server(nb, items, stagger=2, kdir=1)
    for k in (3, 8):
        res, srv = run_server(nb, items, stagger=2, kdir=k)
        for a, b in zip(base, res):
            assert all(np.array_equal(a[f], b[f]) for f in FIELDS)
        assert srv.status()["particle_steps"] == sum(len(sd["m"]) * st for sd, st in items)
    for bad in (0, 1025):
        with pytest.raises(nb.NBError):
            nb.Server(direct_steps_per_tick=bad)


def test_zero_step_request_and_result_collection(nb):
    """A request of zero steps completes on the first tick with its initial
    state.  Asking for a result too early is an error that keeps the request
    collectable; a result is collected once; an unknown id is an error."""
    sy = systems(nb)
    srv = nb.Server(direct_threshold=32768, bh_lanes=1, direct_steps_per_tick=4)
    early = srv.submit(sy[1], 20)  # more steps than one tick takes
    with pytest.raises(nb.NBError) as e:
        srv.result(early)  # still pending
    assert e.value.code == 6  # NB_ERR_STATE
    rid = srv.submit(sy[0], 0)
    srv.tick()
    assert srv.poll() == [rid]
    res = srv.result(rid)
    for f in FIELDS:
        assert np.array_equal(res[f], sy[0][f]), f
    for again in (rid, 12345):
        with pytest.raises(nb.NBError) as e:
            srv.result(again)
        assert e.value.code == 6
    while srv.status()["finished_uncollected"] == 0:
        srv.tick()
    assert srv.poll() == [early]
    assert srv.result(early)["x"].shape == (len(sy[1]["m"]),)


def test_bh_lane_workspace_accounting(nb):
    """Barnes-Hut lanes keep a workspace sized for the largest request they
    have run: admission charges every lane's growth up front and never
    releases it, while a request's own state is released when it finishes."""
    n, steps = 3000, 3
    sysd = nb.ic_plummer(n, 21)
    state = n * 80
    # one step per tick (no tick budget), so the request is still active
    # after the first tick
    probe = nb.Server(hbm_budget_bytes=200e9, direct_threshold=1000, bh_lanes=2, tick_budget=0)
    probe.submit(sysd, steps)
    probe.tick()
    lanes_plus_one = probe.status()["bytes_in_use"]
    assert lanes_plus_one > state
    while probe.status()["active"]:
        probe.tick()
    assert probe.status()["bytes_in_use"] == lanes_plus_one - state  # lanes keep theirs
    alone = None
    for rid in probe.poll():
        alone = probe.result(rid)

    # Room for the lanes and one request's state: the second request waits.
    srv = nb.Server(hbm_budget_bytes=lanes_plus_one + state // 2, direct_threshold=1000,
                    bh_lanes=2, tick_budget=0)
    ids = [srv.submit(sysd, steps), srv.submit(sysd, steps)]
    results, ticks = {}, 0
    while len(results) < 2:
        srv.tick()
        st = srv.status()
        assert st["active"] <= 1 and st["bytes_in_use"] <= lanes_plus_one
        for rid in srv.poll():
            results[rid] = srv.result(rid)
        ticks += 1
        assert ticks < 100
    for rid in ids:
        assert all(np.array_equal(results[rid][f], alone[f]) for f in FIELDS)

    with pytest.raises(nb.NBError) as e:  # lanes alone exceed this budget
        nb.Server(hbm_budget_bytes=state * 2, direct_threshold=1000, bh_lanes=2).submit(sysd, 1)
    assert e.value.code == 5



def test_bh_steps_per_tick_does_not_change_results(nb):
    """Without a tick budget, Barnes-Hut requests take up to B steps per tick
    (fused leapfrog runs); every B gives the same final states, bit for bit,
    as the default budgeted schedule, and the particle-step count is exact."""
    sy = systems(nb)
    items = [(sy[4], 7), (sy[0], 5), (sy[5], 4), (sy[1], 8)]
    base, _ = run_server(nb, items, stagger=2)
    for k in (1, 3, 8):
        res, srv = run_server(nb, items, stagger=2, kbh=k, tick=0)
        for a, b in zip(base, res):
            assert all(np.array_equal(a[f], b[f]) for f in FIELDS)
        assert srv.status()["particle_steps"] == sum(len(sd["m"]) * st for sd, st in items)
    for bad in (0, 1025):
        with pytest.raises(nb.NBError):
            nb.Server(bh_steps_per_tick=bad)


def test_failed_bh_request_fails_alone(nb, monkeypatch):
    """A Barnes-Hut request whose walk frontier overflows (NB_WALK_FCAP, a
    testing override that keeps every kernel in bounds) finishes at once with
    NB_ERR_CAPACITY, its state attached to the error; the direct requests of
    the same ticks finish normally, bit-identical to running alone, and its
    memory is released."""
    sy = systems(nb)
    direct_items = [(sy[0], 5), (sy[1], 8)]
    alone = [run_server(nb, [it])[0][0] for it in direct_items]
    monkeypatch.setenv("NB_WALK_FCAP", "4")
    srv = nb.Server(hbm_budget_bytes=200e9, direct_threshold=32768, bh_lanes=2)
    bad = srv.submit(sy[5], 50)
    ids = [srv.submit(s, k) for s, k in direct_items]
    done, ticks = {}, 0
    while len(done) < 3:
        srv.tick()
        ticks += 1
        for rid in srv.poll():
            try:
                done[rid] = srv.result(rid)
            except nb.NBError as e:
                assert rid == bad and e.code == 4  # NB_ERR_CAPACITY
                # flag 1 (frontier), and 16 when the NaNs reached the step's
                # second evaluation as non-finite positions
                assert re.search(r"flags (1|17):", str(e)), str(e)
                assert e.state["x"].shape == (len(sy[5]["m"]),)
                done[rid] = None
                assert ticks == 1  # failed in the tick of its first step
        assert ticks < 100
    for rid, a in zip(ids, alone):
        assert all(np.array_equal(done[rid][f], a[f]) for f in FIELDS)
    st = srv.status()
    assert st["active"] == 0 and st["pending"] == 0


def test_million_particle_bh_request_equals_library_path(nb):
    """A Barnes-Hut request at the top of the serving range (N = 10^6, the
    size that sets p99), 3 steps through the server while a direct request
    shares its ticks, is bit-identical to the library's own Barnes-Hut +
    leapfrog path."""
    n, steps = 1_000_000, 3
    s = nb.ic_plummer(n, 77)
    small = nb.ic_plummer(2000, 78)
    res, srv = run_server(nb, [(s, steps), (small, 5)], lanes=2, kbh=2)
    f = nb.BarnesHutForceCalculator(0.005, 0.5, 8, "quadrupole", G=1.0)
    p = nb.Particles.from_arrays(s["m"], s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"])
    nb.LeapfrogIntegrator().run(f, p, 1e-4, steps)
    g = p.download()
    for fl in FIELDS:
        assert np.array_equal(g[fl], res[0][fl]), fl
    assert srv.status()["particle_steps"] == n * steps + 2000 * 5


def test_shortest_remaining_work_first(nb):
    """With a per-tick budget the Barnes-Hut requests run shortest remaining
    work first: submitted together, they finish in order of n x steps, and
    every final state is bit-identical to the every-request-every-tick
    schedule; a negative budget is rejected."""
    sy = systems(nb)
    items = [(sy[5], 6), (nb.ic_plummer(35000, 9), 2), (sy[4], 4)]  # 300k, 70k, 160k particle-steps
    base, _ = run_server(nb, items)
    order = []
    res, srv = run_server(nb, items, lanes=1, tick=120000, order=order)
    assert order == [1, 2, 0]
    for a, b in zip(base, res):
        assert all(np.array_equal(a[f], b[f]) for f in FIELDS)
    assert srv.status()["particle_steps"] == sum(len(sd["m"]) * st for sd, st in items)
    with pytest.raises(nb.NBError):
        nb.Server(tick_budget=-1.0)
