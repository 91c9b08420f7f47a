#!/usr/bin/env python3
"""ARAP deformation benchmark on one MI355X (gfx950).

A step is one ARAP iteration (local rotation step + rotated-edge residual + block LDL^T
forward/backward sweeps) over the whole mesh, with every input resident in HBM.

Legs:
  plain   (the value): the 163,842-vertex noisy icosphere (7 subdivisions, sigma 0.01,
          seed 1234; 16,384 anchors at the bottom, 8,192 handles lifted 0.4), FP32 local step
          and FP32 block LDL^T with one FP64 residual refinement per iteration; W untimed and
          K timed iterations replayed from hipGraphs; iterations/s.
  check   position error of that path (and of the FP64 path) against the FP64 CPU oracle
          after the same number of iterations from the same start (100, as many as fit the
          CPU time cap), the ARAP energy and iterations/s of each; the same for the Jacobi-PCG alternative
          of the global step (FP32, 1000 CG iterations per solve), with its own time per
          iteration.
  cpu     the oracle (single-threaded C restatement of inc/deform/arap.h) on the same mesh:
          setup plus a bounded number of iterations, timed on this host.
  served  64 resident meshes (log-uniform 5k-164k vertices), Poisson drag requests at
          2000/s; each request moves its handles to a point of a seeded 4-keyframe SE3
          trajectory and runs 5 warm-started iterations; p50/p99 latency (arrival to result
          on the host) and aggregate iterations/s; then a saturation sweep (4k-32k requests/s
          until p99 passes 16 ms, then three bisection steps: the capacity is the highest rate
          measured within 16 ms), compared with the same meshes iterated one by one.
  frames  the reference's interactive loop, setConstraint + deform(5) per frame in the
          reference's rest mode (each frame re-runs the cotan weights and the refactorization),
          on the 67k bar (FP64) and the 164k icosphere (FP64, FP32 + refinement): frames/s
          and the library's milliseconds per frame by phase.
  batched 256 meshes of 1k-200k vertices iterated together (20 iterations) against the
          same meshes iterated one by one; each side with its own factor shape (the batch
          without level pairs, the one-by-one meshes with them); --no-batched skips it.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import platform
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("mesh-deform_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

import meshgen  # noqa: E402
import mdarap  # noqa: E402
import trajectory as tj  # noqa: E402

METRIC = "ARAP iters/s @164k verts + p99 drag latency, served vs plain, rel err≤1e-5"
PEAK_HBM_GBS = 8000.0
PCG_ITERS = 1000  # CG iterations per solve of the Jacobi-PCG alternative (1e-5 on the 164k icosphere)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def plain_leg(ctx, pr, steps, warmup, tune=True, budget=None, fixed_ratio=None):
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, subtree_budget=budget,
                                                        amalg_ratio=fixed_ratio))
    m.set_constraints(idx, tgt)
    t0 = time.perf_counter()
    m.prepare(pr.xyz)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    # factor and sweep layout: the fastest (amalgamation ratio, subtree budget), measured on
    # this mesh; the mesh is left freshly prepared with them
    ratio = fixed_ratio
    if tune and budget is None and fixed_ratio is None:
        ratio, budget, _ = m.autotune_factor(iters=30)
    else:
        budget = m.stats()["subtree_budget"]
    tune_s = time.perf_counter() - t0
    m.iterate(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    m.iterate(steps)  # returns after the stream is synchronised
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # HIP events on the library's stream: graph replays, then one pass with events between
    # the phases (local, residual, forward sweeps, backward sweeps + update)
    ev_total_ms, phase_ms = m.time_iterations(0, steps)
    st = m.stats()
    st["subtree_budget"] = budget
    st["amalg_ratio"] = ratio if ratio is not None else mdarap.config().amalg_ratio
    st["level_pairs"] = mdarap.config().level_pairs
    st["autotune_s"] = tune_s
    return m, st, setup_s, wall, ev_total_ms, phase_ms


def check_leg(ctx, pr, iters, cpu_seconds_cap, budget, ratio):
    import oracle

    idx, tgt = pr.constraints()
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    t0 = time.perf_counter()
    assert a.deform(0)
    setup_s = time.perf_counter() - t0
    # bounded sample: as many iterations as fit the cap, at least 3
    t0 = time.perf_counter()
    assert a.deform(1)
    one = time.perf_counter() - t0
    n_it = int(max(3, min(iters, cpu_seconds_cap / max(one, 1e-6))))
    t0 = time.perf_counter()
    assert a.deform(n_it - 1)
    rest_s = time.perf_counter() - t0
    cpu_iter_s = (one + rest_s) / n_it
    errs, energies, rates = {}, {"oracle": a.energy()}, {}
    for name, prec in (("fp32_refine", mdarap.PREC_FP32_REFINE), ("fp64", mdarap.PREC_FP64)):
        # the plain leg's factor and sweep layout
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=prec, subtree_budget=budget, amalg_ratio=ratio))
        m.set_constraints(idx, tgt)
        assert m.deform(n_it)
        errs[name] = meshgen.relative_error(m.mesh, a.mesh)
        energies[name] = m.energy()
        ev_ms, _ = m.time_iterations(1, 20)  # HIP events over 20 further iterations
        rates[name] = round(2e4 / ev_ms, 2)
        m.close()
    # the Jacobi-PCG alternative of the global step (FP32, warm-started from the previous
    # iterate, FP64 residual): CG iterations per solve for the 1e-5 bound on this mesh
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, solver=mdarap.SOLVER_PCG,
                                                       pcg_iters=PCG_ITERS, pcg_tol=1e-7))
    m.set_constraints(idx, tgt)
    assert m.deform(n_it)
    errs["fp32_refine_pcg"] = meshgen.relative_error(m.mesh, a.mesh)
    energies["fp32_refine_pcg"] = m.energy()
    pcg_ms, _ = m.time_iterations(1, 10)  # HIP events over 10 further iterations
    m.close()
    pcg = {"cg_iterations_per_solve": PCG_ITERS, "ms_per_iteration": round(pcg_ms / 10, 3),
           "iters_per_s": round(1e4 / pcg_ms, 2)}
    return dict(setup_s=setup_s, iters=n_it, iter_s=cpu_iter_s, errors=errs, energy=energies, pcg=pcg, rates=rates)


def frames_leg(ctx, pr, frames, iters, precision):
    """The reference's interactive loop (examples/deform_sphere.cpp:71-91,
    examples/deform_bar.cpp:84-94): every frame moves the handles with setConstraint and calls
    deform(iters). In the reference's rest mode every such call is dirty: it takes the current
    mesh as the rest pose, recomputes the cotan weights and refactors (arap.h:84,102-120); the
    constrained set does not change, so the host analysis of the first call is kept. Frames/s
    of the whole loop (host copies of the Python binding included) and the library's own
    milliseconds per frame by phase (medians)."""
    idx, _ = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=precision))
    m.set_constraints(idx, pr.xyz[idx])
    assert m.deform(0)  # the first call analyses the constrained set
    first = m.stats()["deform_ms"]
    h = pr.handles
    start = pr.xyz[h].copy()
    ph = []
    t0 = time.perf_counter()
    for f in range(frames):
        s = (f + 1.0) / frames
        m.set_constraints(h, start + s * (pr.handle_targets - start))
        assert m.deform(iters)
        ph.append(m.stats()["deform_ms"])
    wall = time.perf_counter() - t0
    m.close()
    ph = np.median(np.array(ph), axis=0)
    return dict(workload=pr.name, n=pr.n, precision="fp64" if precision == mdarap.PREC_FP64 else "fp32_refine",
                frames=frames, iterations_per_frame=iters, frames_per_s=round(frames / wall, 2),
                library_ms_per_frame=round(float(ph[5]), 3),
                phases_ms={"upload_and_cotan": round(float(ph[1]), 3), "factor": round(float(ph[2]), 3),
                           "iterations": round(float(ph[3]), 3), "download": round(float(ph[4]), 3)},
                first_call_analysis_ms=round(float(first[0]), 1))


def served_problems(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))
    return [meshgen.mixed_problem(k, int(s), rng) for k, s in enumerate(sizes)]


def served_stream(ctx, probs, meshes, utils, trajs, rng, rate, seconds, iters, budget=0):
    """One Poisson stream of drag requests at `rate` per second for `seconds` against a fresh
    server holding the meshes (their state carries over from earlier streams). Latency is
    arrival to result on the host; the aggregate is ARAP iterations per second of wall time
    from the first arrival to the last result."""
    srv = mdarap.Server(ctx, hbm_budget_bytes=int(budget))
    ids = [srv.add_mesh(m) for m in meshes]
    # Poisson arrivals; targets precomputed so the submit loop only submits
    nreq = max(1, int(rate * seconds))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    picks = rng.integers(len(probs), size=nreq)
    times = rng.uniform(0.0, 1.0, size=nreq)
    reqs = [(int(k), utils[k].handles, utils[k].targets(trajs[k](float(t)))) for k, t in zip(picks, times)]
    lat = np.zeros(nreq)
    lag = np.zeros(nreq)
    tickets = [None] * nreq
    done = threading.Event()

    def waiter():
        for i in range(nreq):
            while tickets[i] is None:
                time.sleep(0.0002)
            _, l = srv.wait(tickets[i], None, timeout_ms=120000.0)
            lat[i] = l
        done.set()

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    start = time.perf_counter()
    for i, (k, h, t) in enumerate(reqs):
        now = time.perf_counter() - start
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        lag[i] = max(0.0, time.perf_counter() - start - arrivals[i])
        tickets[i] = srv.submit(ids[k], h, t, iters)
    done.wait(timeout=600)
    elapsed = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    total = lat + lag * 1e3
    return dict(rate_per_s=rate, requests=nreq, p50_ms=round(float(np.percentile(total, 50)), 4),
                p99_ms=round(float(np.percentile(total, 99)), 4), max_ms=round(float(total.max()), 4),
                iters_per_s=round(float(nreq * iters / elapsed), 1),
                submit_lag_p99_ms=round(float(np.percentile(lag, 99) * 1e3), 4),
                launches=st["iterations"], mean_requests_per_launch=round(st["mesh_iterations"] / max(1, st["iterations"]), 2),
                evictions=st["evictions"], restores=st["restores"], peak_resident_bytes=st["peak_resident_bytes"])


def served_leg(ctx, count, seconds, rate, iters, seed, sweep=(4000, 8000, 16000, 32000), p99_bound_ms=16.0):
    """The served workload at `rate` (the headline), then a saturation sweep over higher rates
    until p99 exceeds the bound, then three bisection steps between the last rate that met the
    bound and the first that did not (the capacity: the highest rate measured within it)."""
    rng = np.random.default_rng(seed)
    probs = served_problems(count, 5000, 164000, seed)
    meshes, utils, trajs = [], [], []
    t0 = time.perf_counter()
    for pr in probs:
        idx, _ = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri,
                        mdarap.config(precision=mdarap.PREC_FP32_REFINE, rest_mode=mdarap.REST_FIXED))
        m.set_constraints(idx, pr.xyz[idx])
        meshes.append(m)
        origin = np.eye(4)
        origin[:3, 3] = pr.xyz[pr.handles].mean(axis=0)  # handles pivot about their centroid
        utils.append(tj.DeformationUtil(pr.xyz, pr.handles, origin=origin))
        trajs.append(tj.random_trajectory(rng))
    mdarap.prepare_many(ctx, meshes)
    for m in meshes:
        m.autotune(iters=3)
    setup_s = time.perf_counter() - t0
    footprint = sum(m.stats()["device_bytes"] for m in meshes)
    # the same meshes iterated one by one (each alone, its own graphs), `iters` iterations each
    for m in meshes:
        m.iterate(1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for m in meshes:
        m.iterate(iters)
    plain_same = count * iters / (time.perf_counter() - t0)
    head = served_stream(ctx, probs, meshes, utils, trajs, rng, rate, seconds, iters)
    sweep_out, capacity = [], head if head["p99_ms"] < p99_bound_ms else None
    lo, hi = (rate if capacity is not None else None), None
    for r in sweep:
        log("served sweep: %.0f requests/s" % r)
        x = served_stream(ctx, probs, meshes, utils, trajs, rng, r, min(seconds, 1.0), iters)
        sweep_out.append(x)
        if x["p99_ms"] >= p99_bound_ms:
            hi = r
            break
        capacity, lo = x, r
    if lo is not None and hi is not None:
        for _ in range(3):
            r = round((lo + hi) / 2 / 250) * 250
            if r <= lo or r >= hi:
                break
            log("served sweep (bisection): %.0f requests/s" % r)
            x = served_stream(ctx, probs, meshes, utils, trajs, rng, r, min(seconds, 1.0), iters)
            sweep_out.append(x)
            if x["p99_ms"] >= p99_bound_ms:
                hi = r
            else:
                capacity, lo = x, r
        sweep_out.sort(key=lambda x: x["rate_per_s"])
    sizes = [p.n for p in probs]
    for m in meshes:
        m.close()
    out = dict(meshes=count, total_vertices=int(sum(sizes)), iters_per_request=iters, setup_s=round(setup_s, 3),
               device_bytes=int(footprint), plain_one_by_one_iters_per_s=round(plain_same, 1))
    out.update(head)
    out["sweep"] = sweep_out
    out["capacity"] = None if capacity is None else dict(rate_per_s=capacity["rate_per_s"],
                                                          p99_ms=capacity["p99_ms"],
                                                          iters_per_s=capacity["iters_per_s"],
                                                          vs_plain_one_by_one=round(capacity["iters_per_s"] / plain_same, 3))
    return out


def batched_leg(ctx, count, iters, seed, tune=True, baseline=True):
    probs = meshgen.mixed_batch(count, 1000, 200000, seed=seed)

    def prepared(level_pairs):
        out = []
        for pr in probs:
            idx, tgt = pr.constraints()
            m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE,
                                                               level_pairs=level_pairs))
            m.set_constraints(idx, tgt)
            out.append(m)
        mdarap.prepare_many(ctx, out)
        return out

    # plain: each mesh with its own tuned sweep layout and the library's default factor shape
    # (level pairs: fewer launches, the better choice for a mesh iterated alone), iterated alone
    # (its own graphs), back to back
    tune_plain_s, plain_s = 0.0, None
    if baseline:
        meshes = prepared(None)
        t0 = time.perf_counter()
        if tune:
            for m in meshes:
                m.autotune(iters=3)
        tune_plain_s = time.perf_counter() - t0
        for m in meshes:
            m.iterate(1)  # graph capture
        ctx.synchronize()
        t0 = time.perf_counter()
        for m in meshes:
            m.iterate(iters)
        plain_s = time.perf_counter() - t0
        for m in meshes:
            m.close()
    # batched: the same meshes prepared for batch use (no level pairs: a batch is bound by
    # bytes rather than launches, and pairs add panel bytes), one launch per kernel for all
    t0 = time.perf_counter()
    meshes = prepared(False)
    setup_s = time.perf_counter() - t0
    b = mdarap.Batch(ctx, meshes)
    t0 = time.perf_counter()
    budget = b.autotune()[0] if tune else meshes[0].stats()["subtree_budget"]
    tune_batch_s = time.perf_counter() - t0
    b.iterate(2)  # warm-up (graph capture)
    ms = b.time_iterations(0, iters)
    b.close()
    total_iters = count * iters
    # canonical bytes of one batched iteration (every mesh once) over its event time
    batch_bytes = sum(canonical_bytes(pr, m.stats())["B_iter"] for pr, m in zip(probs, meshes))
    achieved = batch_bytes / (ms / iters * 1e-3) / 1e9
    return dict(meshes=count, total_vertices=int(sum(p.n for p in probs)), iterations=iters,
                batched_iters_per_s=total_iters / (ms / 1e3),
                batched_ms_per_iteration=round(ms / iters, 4), canonical_bytes_per_batched_iteration=int(batch_bytes),
                achieved_GBs=round(achieved, 1), hbm_frac=round(achieved / PEAK_HBM_GBS, 4),
                plain_iters_per_s=None if plain_s is None else total_iters / plain_s,
                speedup=None if plain_s is None else (total_iters / (ms / 1e3)) / (total_iters / plain_s),
                batch_subtree_budget=budget, batch_level_pairs=0, plain_level_pairs=mdarap.config().level_pairs,
                setup_s=setup_s, autotune_plain_s=tune_plain_s, autotune_batch_s=tune_batch_s)


def canonical_bytes(pr, st):
    """Algorithmic bytes of one ARAP iteration by the canonical per-pass counts (FP32 arrays,
    int32 indices; each array once per pass): local step + rotated-edge pass
    (120 + 16 d) n + 36 nf with d = 2E/n, the FP64 residual of the refinement step
    12 nnz(A) + 124 nf, and one direct sweep pair with the FP32 factor 16 nL + 68 nf
    (nL = strict nnz of the exact factor: the explicit zeros that relaxed amalgamation
    stores are not counted). An iteration here runs one residual and one sweep
    pair (the residual form folds the right-hand side into the refinement), so
    B_iter = B_local+rhs + B_ref with B_ref = 12 nnz(A) + 124 nf + B_solve."""
    n, nf = pr.n, st["nf"]
    e = np.sort(np.concatenate([pr.tri[:, [0, 1]], pr.tri[:, [1, 2]], pr.tri[:, [2, 0]]]), axis=1)
    e = np.unique(e, axis=0)
    con = np.zeros(n, bool)
    con[pr.constraints()[0]] = True
    e_ff = int(np.count_nonzero(~con[e[:, 0]] & ~con[e[:, 1]]))
    nnz_a = nf + 2 * e_ff
    n_l = st["factor_nnz_exact"]
    d = 2.0 * e.shape[0] / n
    b_local = (120 + 16 * d) * n + 36 * nf
    b_solve = 16 * n_l + 68 * nf
    b_ref = 12 * nnz_a + 124 * nf + b_solve
    return dict(B_iter=int(b_local + b_ref), B_local_rhs=int(b_local), B_solve=int(b_solve), B_ref=int(b_ref),
                n=n, nf=nf, E=int(e.shape[0]), nnzA_full=int(nnz_a), nL=int(n_l), mean_valence=round(d, 4))


def traffic_probe(ctx, iters, budget=None, ratio=None):
    """Workload for rocprofv3 --pmc passes: the library's plain copy kernel over 256 MiB
    (calibration: each dispatch reads and writes exactly 256 MiB), then `iters` ARAP
    iterations of the plain leg's mesh and configuration."""
    nbytes = 256 << 20
    gbps = mdarap.copy_probe(ctx, nbytes, 3)
    pr = meshgen.noisy_icosphere_problem(7, 0.01, 1234)
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, subtree_budget=budget,
                                                        amalg_ratio=ratio))
    m.set_constraints(idx, tgt)
    m.prepare(pr.xyz)
    if budget is None and ratio is None:
        m.autotune_factor()
    m.iterate(2)
    ctx.synchronize()
    log("traffic probe: calibration copies done, running %d iterations" % iters)
    m.iterate(iters)
    ctx.synchronize()
    used = m.stats()["subtree_budget"]
    m.close()
    print(json.dumps({"probe": "traffic", "copy_bytes": nbytes, "copies": 4, "copy_GBs": round(gbps, 1),
                      "iterations": iters, "warmup_iterations": 2, "subtree_budget": used,
                      "amalg_ratio": ratio, "level_pairs": mdarap.config().level_pairs}), flush=True)


def traffic_summary(fetch_csv, write_csv, iters):
    """HBM bytes per ARAP iteration from two rocprofv3 --pmc passes of --traffic-probe:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), summed over the library's kernels of the
    last `iters` iterations, each scaled by bytes / counter of the 256 MiB calibration copy."""
    import csv

    def load(path, counter):
        rows = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                d = int(r["Dispatch_Id"])
                rows[d] = (r["Kernel_Name"], rows.get(d, ("", 0.0))[1] + float(r["Counter_Value"]))
        return [rows[d] for d in sorted(rows)]

    out = {}
    for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        disp = load(path, name)
        calib = [v for k, v in disp if "k_copy_probe" in k]
        ours = [(k, v) for k, v in disp if "md::" in k and "k_copy_probe" not in k]
        copy_kib = float(np.median(calib))  # each copy dispatch moves 256 MiB each way
        scale = (256 << 20) / (copy_kib * 1024.0)
        per_iter = None
        # the library's dispatches of the timed iterations: the last `iters` local steps start them
        starts = [i for i, (k, _) in enumerate(ours) if "k_local" in k]
        first = starts[-iters]
        per_iter = sum(v for _, v in ours[first:]) * 1024.0 * scale / iters
        out[name] = dict(copy_counter_kib=copy_kib, scale=scale, bytes_per_iter=per_iter,
                         dispatches_per_iter=(len(ours) - first) / iters)
    out["traffic_bytes_per_iter"] = out["FETCH_SIZE"]["bytes_per_iter"] + out["WRITE_SIZE"]["bytes_per_iter"]
    return out


def read_traffic(path, budget, ratio, pairs):
    """traffic_bytes_per_iter from a committed traffic summary, when it was measured on the
    factor and sweep layout the plain leg runs. The summary holds one or more pairs of lines
    'traffic_bytes_per_iter: N' then 'traffic_config: subtree_budget=B amalg_ratio=R level_pairs=P'."""
    found, val = {}, None
    try:
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith("traffic_bytes_per_iter:"):
                    val = float(line.split(":", 1)[1])
                elif line.startswith("traffic_config:") and val is not None:
                    kv = dict(t.split("=") for t in line.split(":", 1)[1].split())
                    cfg = (int(kv["subtree_budget"]), round(float(kv["amalg_ratio"]), 6), int(kv.get("level_pairs", 0)))
                    found[cfg] = val
                    val = None
    except (OSError, ValueError, KeyError):
        return None
    return found.get((int(budget), round(float(ratio), 6), int(pairs)))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--check-iters", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--served-meshes", type=int, default=64)
    ap.add_argument("--served-seconds", type=float, default=2.0)
    ap.add_argument("--served-rate", type=float, default=2000.0)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-frames", action="store_true", help="skip the interactive-frames leg")
    ap.add_argument("--frames", type=int, default=20)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--batched", action="store_true", help="(default) kept for older command lines")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--batched-meshes", type=int, default=256)
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg (for profiling)")
    ap.add_argument("--no-plain-baseline", action="store_true",
                    help="batched leg without the one-by-one baseline (profiling runs: one graph per mesh)")
    ap.add_argument("--subtree-budget", type=int, default=None,
                    help="fixed sweep layout of the plain leg and the traffic probe (no autotune)")
    ap.add_argument("--amalg-ratio", type=float, default=None,
                    help="fixed factor amalgamation of the plain leg and the traffic probe (no autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default sweep layout (profiling runs: the tuner re-captures graphs)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the workload of the rocprofv3 --pmc traffic passes")
    ap.add_argument("--probe-iters", type=int, default=5)
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two --pmc counter_collection.csv files of --traffic-probe (no GPU)")
    ap.add_argument("--traffic-summary", default=os.path.join(ROOT, "profiles", "traffic_icosphere164k_fp32_refine.md"))
    args = ap.parse_args()
    if args.traffic_from:
        print(json.dumps(traffic_summary(args.traffic_from[0], args.traffic_from[1], args.probe_iters), indent=1))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (meshes would be "
              "sharded across GPUs by device footprint). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if mdarap.device_count() == 0:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    ctx = mdarap.Context(0)
    if args.traffic_probe:
        traffic_probe(ctx, args.probe_iters, args.subtree_budget, args.amalg_ratio)
        ctx.close()
        return
    if args.batched_only:
        # calibration for counter passes: the plain copy kernel moving 256 MiB each way, 4 times
        mdarap.copy_probe(ctx, 256 << 20, 3)
        log("batched leg: %d meshes" % args.batched_meshes)
        print(json.dumps({"batched": batched_leg(ctx, args.batched_meshes, 20, 4242, not args.no_autotune,
                                                 not args.no_plain_baseline)}), flush=True)
        ctx.close()
        return
    pr = meshgen.noisy_icosphere_problem(7, 0.01, 1234)
    log("plain leg: %s n=%d F=%d" % (pr.name, pr.n, pr.F))
    m, st, setup_s, wall, ev_ms, phase = plain_leg(ctx, pr, args.steps, args.warmup, not args.no_autotune,
                                                   args.subtree_budget, args.amalg_ratio)
    value = args.steps / wall
    ms_step = wall * 1e3 / args.steps
    ev_step_ms = ev_ms / args.steps
    ph_step = phase / args.steps
    solve_ms = ph_step[2] + ph_step[3]
    solve_bytes = st["bytes_forward"] + st["bytes_backward"]
    iter_bytes = st["bytes_local"] + st["bytes_residual"] + solve_bytes
    cb = canonical_bytes(pr, st)
    achieved = cb["B_iter"] / (ev_step_ms * 1e-3) / 1e9
    traffic = read_traffic(args.traffic_summary, st["subtree_budget"], st["amalg_ratio"], st["level_pairs"])
    log("plain: %.1f iters/s (%.4f ms/iter wall, %.4f ms event); phases ms %s" %
        (value, ms_step, ev_step_ms, np.round(ph_step, 4).tolist()))
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "noisy icosphere, 7 subdivisions (n=163842, F=327680), radial noise "
                               "sigma=0.01 seed=1234, 16384 anchors (lowest z), 8192 handles lifted 0.4; "
                               "FP32 local step + FP32 block LDL^T, FP64 factor and one FP64 residual "
                               "refinement per iteration",
                   "n": pr.n, "faces": pr.F, "free": st["nf"], "supernodes": st["supernodes"],
                   "levels": st["levels"], "panel_entries": st["panel_entries"],
                   "factor_nnz": st["factor_nnz"], "setup_s": round(setup_s, 3),
                   "subtree_budget": st["subtree_budget"], "amalg_ratio": st["amalg_ratio"],
                   "level_pairs": st["level_pairs"], "autotune_s": round(st["autotune_s"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None if traffic is None else int(traffic),
                     "kernel": "one ARAP iteration (k_local, k_sub_fwd with the residual, k_fwd per level, k_bwd "
                               "per level, k_sub_bwd with the p' update), replayed from hipGraphs of 8 iterations; "
                               "HIP events on the library stream, time per iteration",
                     "algorithmic_bytes": cb["B_iter"], "ms_per_iteration": round(ev_step_ms, 5),
                     "canonical": cb,
                     "counted_bytes": {"iteration": int(iter_bytes), "sweeps": int(solve_bytes),
                                       "sweeps_GBs": round(solve_bytes / (solve_ms * 1e-3) / 1e9, 1)}},
        "phases_ms_per_iter": {"local": round(ph_step[0], 5), "residual": round(ph_step[1], 5),
                               "forward": round(ph_step[2], 5), "backward": round(ph_step[3], 5)},
    }
    m.close()
    if not args.no_check:
        log("check + cpu leg (oracle, bounded to ~%.0f s)" % args.cpu_seconds)
        ck = check_leg(ctx, pr, args.check_iters, args.cpu_seconds, st["subtree_budget"], st["amalg_ratio"])
        out["rel_err_vs_fp64_oracle"] = {k: float("%.3e" % v) for k, v in ck["errors"].items()}
        out["rel_err_iters"] = ck["iters"]
        out["arap_energy"] = {k: float("%.10e" % v) for k, v in ck["energy"].items()}
        out["pcg_alternative"] = ck["pcg"]
        out["iters_per_s_by_precision"] = ck["rates"]  # the check meshes, same factor and layout
        out["cpu_baseline"] = {"value": round(1.0 / ck["iter_s"], 3), "unit": "iters/s", "cores": 1,
                               "kind": "port", "host_cpu": cpu_model(), "host_nproc": os.cpu_count(),
                               "sample": "oracle (FP64 C restatement of arap.h, up-looking sparse LDL^T, "
                                         "Jacobi SVD) on the same 163842-vertex mesh: %d iterations after a "
                                         "%.2f s setup, host %s" % (ck["iters"], ck["setup_s"], cpu_model()),
                               "setup_s": round(ck["setup_s"], 3)}
    if not args.no_frames:
        log("frames leg: setConstraint + deform(5) per frame, reference rest mode")
        out["frames"] = [frames_leg(ctx, meshgen.bar_problem(), args.frames, 5, mdarap.PREC_FP64),
                         frames_leg(ctx, pr, args.frames, 5, mdarap.PREC_FP64),
                         frames_leg(ctx, pr, args.frames, 5, mdarap.PREC_FP32_REFINE)]
    if not args.no_served:
        log("served leg: %d meshes, %.1f s of arrivals at %.0f/s" %
            (args.served_meshes, args.served_seconds, args.served_rate))
        out["served"] = served_leg(ctx, args.served_meshes, args.served_seconds, args.served_rate, 5, 777)
        out["served"]["plain_iters_per_s_at_164k"] = round(value, 2)
    if not args.no_batched:
        log("batched leg: %d meshes" % args.batched_meshes)
        out["batched"] = batched_leg(ctx, args.batched_meshes, 20, 4242, not args.no_autotune)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
