"""The server's admission planner (plan_admission, run by md_server_* at every iteration
boundary), through its host diagnostic md_host_admission_plan: no device needed.

* Regression: a mesh chosen as an eviction victim in one pass must not be admitted as resident
  by a later request of the same pass (its arrays are freed before that request's targets are
  uploaded). Such a request waits and is restored at a later boundary.
* Invariants over seeded random request streams across many boundaries: the resident bytes
  never exceed the budget and equal the sum over resident meshes; every admitted request's mesh
  is resident and had no request in flight; no mesh is both evicted and restored in one pass;
  victims are idle and least recently used first; max_active holds.
"""
import numpy as np

import mdarap


def test_victim_is_not_admitted_in_the_same_pass():
    # three meshes of 100 bytes, budget 200: meshes 0 and 1 resident and idle, 2 evicted;
    # queue: a request on 2 (needs room: mesh 0, least recently used, is the victim), then one on 0
    take, ev, rs, rb = mdarap.host_admission_plan([100, 100, 100], [1, 2, 0], [1, 1, 0], [0, 0, 0], [2, 0],
                                                  budget=200)
    assert take.tolist() == [True, False]
    assert ev == [0] and rs == [2] and rb == 200
    # next boundary (0 now evicted, 2 busy): the request on 0 evicts the idle mesh 1
    take, ev, rs, rb = mdarap.host_admission_plan([100, 100, 100], [1, 2, 3], [0, 1, 1], [0, 0, 1], [0],
                                                  active=1, budget=200)
    assert take.tolist() == [True] and ev == [1] and rs == [0] and rb == 200


def test_blocked_request_keeps_order_and_resident_requests_pass():
    # budget 150: mesh 2 (100) cannot fit while busy mesh 0 (100) holds memory; a later request
    # on evicted mesh 3 (40, would fit) waits behind it; a request on resident idle mesh 1 passes
    take, ev, rs, rb = mdarap.host_admission_plan([100, 10, 100, 40], [5, 1, 0, 0], [1, 1, 0, 0], [1, 0, 0, 0],
                                                  [2, 3, 1], active=1, budget=150)
    assert take.tolist() == [False, False, True]
    assert ev == [] and rs == [] and rb == 110


def test_unlimited_budget_restores_without_evicting():
    take, ev, rs, rb = mdarap.host_admission_plan([100, 100, 100], [0, 0, 0], [0, 1, 0], [0, 0, 0], [0, 2, 0, 1])
    assert take.tolist() == [True, True, False, True]   # the second request on 0 waits (0 is busy)
    assert ev == [] and rs == [0, 2] and rb == 300


def test_max_active():
    take, _, _, _ = mdarap.host_admission_plan([1] * 4, [0] * 4, [1] * 4, [0, 1, 0, 0], [0, 2, 3], active=1,
                                               max_active=2)
    assert take.tolist() == [True, False, False]


def test_random_streams_keep_the_invariants():
    rng = np.random.default_rng(2024)
    for trial in range(40):
        nm = int(rng.integers(2, 12))
        sizes = rng.integers(1, 100, nm).astype(np.int64)
        budget = int(rng.integers(int(sizes.max()), int(sizes.sum()) + 1))
        max_active = int(rng.integers(0, 4))
        resident = np.zeros(nm, bool)
        used = 0
        for k in rng.permutation(nm):          # registration: resident while it fits
            if used + sizes[k] <= budget:
                resident[k] = True
                used += sizes[k]
        last_use = np.zeros(nm, np.uint64)
        busy = np.zeros(nm, bool)
        queue = []
        for _ in range(60):
            queue += rng.integers(0, nm, int(rng.integers(0, 5))).tolist()
            active = int(busy.sum())
            take, ev, rs, rb = mdarap.host_admission_plan(sizes, last_use, resident, busy, queue, active=active,
                                                          max_active=max_active, budget=budget)
            assert not set(ev) & set(rs)
            admitted = [m for m, t in zip(queue, take) if t]
            # least recently used first: a mesh that stays resident and idle through the pass
            # was a candidate at every choice, so no victim was used more recently than it
            kept_idle = [k for k in range(nm) if resident[k] and not busy[k] and k not in ev and k not in admitted]
            for v in ev:
                assert resident[v] and not busy[v]
                assert all(last_use[v] <= last_use[k] for k in kept_idle)
                resident[v] = False
            for k in rs:
                assert not resident[k]
                resident[k] = True
            assert rb == int(sizes[resident].sum()) <= budget
            assert len(set(admitted)) == len(admitted)
            if max_active > 0:
                assert active + len(admitted) <= max_active
            for m in admitted:
                assert resident[m] and not busy[m]
                busy[m] = True
            clock = int(last_use.max())
            for m in admitted:
                clock += 1
                last_use[m] = clock
            queue = [m for m, t in zip(queue, take) if not t]
            # some in-flight requests complete
            for m in np.flatnonzero(busy):
                if rng.random() < 0.5:
                    busy[m] = False
        # drain: with nothing in flight every request is eventually admitted
        for _ in range(len(queue) + nm + 2):
            if not queue:
                break
            take, ev, rs, rb = mdarap.host_admission_plan(sizes, last_use, resident, busy * 0, queue, budget=budget)
            for v in ev:
                resident[v] = False
            for k in rs:
                resident[k] = True
            clock = int(last_use.max())
            for m, t in zip(queue, take):
                if t:
                    clock += 1
                    last_use[m] = clock
            queue = [m for m, t in zip(queue, take) if not t]
        assert not queue, trial
