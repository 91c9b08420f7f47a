"""The full-size workloads of the large configurations, on the GPU through the C ABI.

(The 3.2k UV sphere over 10 iterations and the 67k bar over 50 iterations are in
test_gpu_parity.py.)

* 164k noisy icosphere (7 subdivisions, sigma 0.01): 100 iterations, FP64 at 1e-9 and FP32 +
  FP64 refinement at 1e-5 against the FP64 oracle; the Jacobi-PCG global step (FP32, 1000 CG
  iterations per solve, FP64 residual) at 1e-5.
* 256 uneven meshes (icospheres and bars, 1k-200k vertices, seed 4242) batched over 20
  iterations with the autotuned sweep layout and graph replay: every mesh's result equals
  the same mesh iterated alone (bitwise), its ARAP energy does not increase from iteration to
  iteration (within 1e-6 of the first energy: FP32 solves), and 8 meshes across the size
  range, the largest included, match the oracle at 1e-5.
* 64 resident meshes (5k-164k vertices, seed 777) served under Poisson drag requests at
  2000/s whose handle targets follow seeded 4-keyframe SE3 trajectories (5 iterations per
  request): each result is bitwise the one the request gives alone, and the requests on 4
  meshes replayed on the oracle in fixed-rest mode match at 1e-5.
"""
import math
import sys
import time

import numpy as np
import pytest

import meshgen
import mdarap
import oracle
import trajectory as tj

pytestmark = pytest.mark.gpu

POS_TOL_FP64 = 1e-9
POS_TOL_FP32 = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    if mdarap.device_count() == 0:
        pytest.fail("no GPU visible: the HIP path cannot run")
    c = mdarap.Context(0)
    yield c
    c.close()


# ---- 164k noisy icosphere, 100 iterations --------------------------------------------------

@pytest.fixture(scope="module")
def ico164k():
    pr = meshgen.noisy_icosphere_problem(7, 0.01, 1234)
    idx, tgt = pr.constraints()
    t0 = time.time()
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    assert a.deform(100)
    say("oracle: 164k icosphere, 100 iterations in %.1f s" % (time.time() - t0))
    return pr, a.mesh.copy(), a.energy()


@pytest.mark.timeout(300)
def test_icosphere164k_100_iterations(ctx, ico164k):
    pr, ref, e_ref = ico164k
    idx, tgt = pr.constraints()
    for prec, tol in ((mdarap.PREC_FP64, POS_TOL_FP64), (mdarap.PREC_FP32_REFINE, POS_TOL_FP32)):
        say("case: icosphere164k 100 iterations, precision", prec)
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=prec))
        m.set_constraints(idx, tgt)
        assert m.deform(100)
        err = meshgen.relative_error(m.mesh, ref)
        say("  rel err %.3e, energy %.10e (oracle %.10e)" % (err, m.energy(), e_ref))
        assert err < tol
        assert m.energy() == pytest.approx(e_ref, rel=1e-9 if prec == mdarap.PREC_FP64 else 1e-6)
        m.close()


@pytest.mark.timeout(300)
def test_icosphere164k_pcg_100_iterations(ctx, ico164k):
    """The Jacobi-PCG alternative at the bench's setting (1000 CG iterations per solve)."""
    pr, ref, _ = ico164k
    idx, tgt = pr.constraints()
    say("case: icosphere164k PCG, 100 iterations")
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, solver=mdarap.SOLVER_PCG,
                                                       pcg_iters=1000, pcg_tol=1e-7))
    m.set_constraints(idx, tgt)
    assert m.deform(100)
    err = meshgen.relative_error(m.mesh, ref)
    say("  rel err %.3e" % err)
    assert err < POS_TOL_FP32
    m.close()


# ---- 256 uneven meshes batched --------------------------------------------------------------

def _prepared(ctx, probs, **kw):
    out = []
    for pr in probs:
        idx, tgt = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, **kw))
        m.set_constraints(idx, tgt)
        out.append(m)
    mdarap.prepare_many(ctx, out)
    return out


@pytest.mark.timeout(600)
def test_batch256_uneven_full_size(ctx):
    iters = 20
    probs = meshgen.mixed_batch(256, 1000, 200000, seed=4242)
    say("case: batch of 256 uneven meshes, %d vertices" % sum(p.n for p in probs))
    meshes = _prepared(ctx, probs, level_pairs=False)
    b = mdarap.Batch(ctx, meshes)
    budget, ms = b.autotune()
    say("  batch budget", budget, "ms per batched iteration %.3f" % ms)
    energies = np.zeros((iters + 1, len(meshes)))
    energies[0] = [m.energy() for m in meshes]
    for i in range(iters):
        b.iterate(1)
        energies[i + 1] = [m.energy() for m in meshes]
    got = [m.download() for m in meshes]
    b.close()
    # energy of every mesh: non-increasing over the iterations (FP32 solves: within 1e-6 of
    # the first iteration's energy)
    e1 = energies[1]
    worst = np.max((energies[2:] - energies[1:-1]) / e1)
    say("  largest relative energy increase %.3e" % worst)
    assert worst <= 1e-6
    for m in meshes:
        m.close()
    # the same meshes, same factor and layout, each iterated alone
    alone = _prepared(ctx, probs, level_pairs=False, subtree_budget=budget)
    for k, (m, g) in enumerate(zip(alone, got)):
        m.iterate(iters)
        assert np.array_equal(m.download(), g), (k, probs[k].name)
        m.close()
    # the oracle on 8 meshes across the size range, the largest included
    order = np.argsort([p.n for p in probs], kind="stable")
    picks = sorted(set(int(order[int(round(q * (len(probs) - 1)))]) for q in np.linspace(0.0, 1.0, 8)))
    for k in picks:
        pr = probs[k]
        idx, tgt = pr.constraints()
        a = oracle.Arap(pr.xyz, pr.tri)
        a.set_constraints(idx, tgt)
        assert a.deform(iters)
        err = meshgen.relative_error(got[k], a.mesh)
        say("  oracle check %s: rel err %.3e" % (pr.name, err))
        assert err < POS_TOL_FP32


# ---- 64 resident meshes served ---------------------------------------------------------------

def _served_setup(count, seed):
    """The served workload of bench.py: meshes from one generator seeded with `seed`, handle
    trajectories and the request stream from a second one with the same seed."""
    g = np.random.default_rng(seed)
    sizes = np.exp(g.uniform(math.log(5000), math.log(164000), size=count))
    probs = [meshgen.mixed_problem(k, int(s), g) for k, s in enumerate(sizes)]
    rng = np.random.default_rng(seed)
    utils, trajs = [], []
    for pr in probs:
        origin = np.eye(4)
        origin[:3, 3] = pr.xyz[pr.handles].mean(axis=0)  # handles pivot about their centroid
        utils.append(tj.DeformationUtil(pr.xyz, pr.handles, origin=origin))
        trajs.append(tj.random_trajectory(rng))
    return probs, utils, trajs, rng


def _served_meshes(ctx, probs):
    out = []
    for pr in probs:
        idx, _ = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri,
                        mdarap.config(precision=mdarap.PREC_FP32_REFINE, rest_mode=mdarap.REST_FIXED))
        m.set_constraints(idx, pr.xyz[idx])  # anchors and handles at rest first
        out.append(m)
    mdarap.prepare_many(ctx, out)
    return out


@pytest.mark.timeout(600)
def test_served_64_meshes_poisson_stream(ctx):
    iters, rate, nreq = 5, 2000.0, 200
    probs, utils, trajs, rng = _served_setup(64, 777)
    say("case: served stream, 64 meshes, %d vertices, %d requests" % (sum(p.n for p in probs), nreq))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    picks = rng.integers(len(probs), size=nreq)
    times = rng.uniform(0.0, 1.0, size=nreq)
    reqs = [(int(k), utils[k].handles, utils[k].targets(trajs[k](float(t)))) for k, t in zip(picks, times)]
    # in the mix: Poisson arrivals, requests on different meshes overlap on the device
    meshes = _served_meshes(ctx, probs)
    srv = mdarap.Server(ctx)
    ids = [srv.add_mesh(m) for m in meshes]
    tickets = []
    start = time.perf_counter()
    for i, (k, h, t) in enumerate(reqs):
        now = time.perf_counter() - start
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        tickets.append(srv.submit(ids[k], h, t, iters))
    mix = [srv.wait(tk, probs[k].n) for tk, (k, _, _) in zip(tickets, reqs)]
    srv.close()
    lat = np.array([l for _, l in mix])
    say("  in the mix: p50 %.2f ms, p99 %.2f ms" % (np.percentile(lat, 50), np.percentile(lat, 99)))
    for m in meshes:
        m.close()
    # alone: one request at a time on fresh meshes
    meshes = _served_meshes(ctx, probs)
    srv = mdarap.Server(ctx)
    ids = [srv.add_mesh(m) for m in meshes]
    for i, (k, h, t) in enumerate(reqs):
        pos, _ = srv.wait(srv.submit(ids[k], h, t, iters), probs[k].n)
        assert np.array_equal(pos, mix[i][0]), i
    srv.close()
    for m in meshes:
        m.close()
    # the oracle in fixed-rest mode replays every request of 4 meshes (the smallest three with
    # requests, and the largest with requests) in submission order
    used = sorted(set(k for k, _, _ in reqs), key=lambda k: probs[k].n)
    for k in used[:3] + used[-1:]:
        pr = probs[k]
        idx, _ = pr.constraints()
        a = oracle.Arap(pr.xyz, pr.tri)
        a.set_constraints(idx, pr.xyz[idx])
        assert a.deform(0, rest_mode=1)
        count = 0
        for (kk, h, t), (pos, _) in zip(reqs, mix):
            if kk != k:
                continue
            a.set_constraints(h, t)
            assert a.deform(iters, rest_mode=1)
            err = meshgen.relative_error(pos, a.mesh)
            assert err < POS_TOL_FP32, (pr.name, count, err)
            count += 1
        say("  oracle replay %s: %d requests" % (pr.name, count))
