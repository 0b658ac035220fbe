"""Benchmark: CASPT2 energies/s (+ p50/p99 latency), split-precision vs plain FP64, on one GPU.

One step = one complete CASPT2 energy (CASCI: determinant strings, CI matrix, roots;
CASPT2: Fock build and diagonalisation, CI rotation, densities, integral transforms,
subspaces A-H) with the integrals already resident in HBM.  The default workload is the
324-spinor one (30 inactive / 14 active with 6 electrons (3003 determinants) / 280
secondary, complex128, seeded synthetic Kramers-paired integrals).

    python bench.py [--steps K] [--warmup W] [--workload split|fp64|small|batch|serve|eig|mdcint]
                    [--served N] [--slots S] [--split SLICES] [--secondary fp64|''] [--stream 0|1]
                    [--requests N] [--rate R] [--no-autotune] [--no-cpu-baseline]

Prints ONE JSON line.  value = energies/s of the K timed steps with split-precision (INT8
Ozaki) integral transforms; beside it: the same K steps in plain FP64 ("fp64"), |dE2| per
subspace of both against the oracle's stored energies ("vs_oracle"), the served run (N scan
points through the request server with cross-request eigensolve batching, split precision)
and its rate against plain FP64, the roofline of the dominant kernel, the same figures for
the 192-spinor workload ("secondary"), a Poisson request stream over resident molecules
("serve_stream") and the CPU baseline (the numpy oracle, one energy).
"""
from __future__ import annotations

import argparse
import json
from math import comb
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# hardware queues per process for the concurrent slots of the served run (HIP reads this
# when the runtime starts, i.e. at the first library call), where the environment does not
# set it already (4, HIP's own default, where it does: 4 to 32 gave the same served rate)
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

PEAK_FP64_TFLOPS = 78.6      # MI355X dense FP64 (matrix and vector)
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E


def workload_spec(name):
    from dirac_caspt2_amd import synth
    return {"small": synth.CONFIG_SMALL, "fp64": synth.CONFIG_FP64, "split": synth.CONFIG_SPLIT}[name]


def cpu_baseline(sp, steps=1):
    """The numpy oracle on the same workload, timed on this host (bounded sample: one energy)."""
    from oracle import caspt2_ref as R
    try:
        from threadpoolctl import threadpool_info
        cores = max((i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"), default=1)
    except Exception:
        cores = os.cpu_count() or 1
    s = sp.spec
    p = R.Problem(s.ninact, s.nact, s.nsec, s.nelec, sp.h, R.FactorIntegrals(sp.B), sp.ecore, s.selectroot,
                  eshift=s.eshift)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = R.run_caspt2(p)
    dt = (time.perf_counter() - t0) / steps
    return res, {"value": 1.0 / dt, "unit": "energies/s", "cores": int(cores), "kind": "port",
                 "sample": f"{steps} full CASPT2 energy of the same workload with the numpy oracle "
                           f"({dt:.2f} s per energy, BLAS threads = {cores})"}


def served_run(spec, sp, nreq, slots, lanes, plain_points, eig_batch=0, eig_window=0.002, split=0, native_min=0):
    """Scan points of the same molecule through the request server, all submitted at once:
    energies/s over the whole batch, per-request latency, and the largest |dE2| against the
    plain FP64 path on the points in plain_points ({index: Caspt2Energies})."""
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Server
    direction = synth.scan_displacements(spec, spec.seed + 1)
    hs = [synth.scan_point(sp, lam, direction) for lam in np.linspace(0.0, 1.0, nreq)]
    with Server(0, nslots=slots, lanes_per_slot=lanes, eig_batch=eig_batch, eig_window=eig_window) as srv:
        if native_min:
            srv.set_option("native_min", native_min)
        mid = srv.add_molecule((spec.ninact, spec.nact, spec.nsec, spec.nelec), sp.B)
        for k in range(slots):                                   # warm every slot
            srv.submit(mid, hs[k % nreq], sp.ecore, spec.selectroot, spec.eshift, split_slices=split)
        for k in range(slots):
            r = srv.wait(120.0)
            assert r is not None and r.status == 0, r
        t0 = time.perf_counter()
        ids = {srv.submit(mid, hs[i], sp.ecore, spec.selectroot, spec.eshift, split_slices=split): i
               for i in range(nreq)}
        res = {}
        while len(res) < nreq:
            r = srv.wait(120.0)
            if r is None:
                raise RuntimeError("served run timed out")
            if r.status:
                raise RuntimeError(f"served request failed: {r.error}")
            res[ids[r.id]] = r
        total = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([res[i].latency for i in range(nreq)]) * 1e3
    svc = np.array([res[i].t_end - res[i].t_start for i in range(nreq)]) * 1e3
    d = max(abs(res[i].e2[X] - e.e2[X]) for i, e in plain_points.items() for X in "ABCDEFGH")
    dt = max(abs(res[i].total - e.total) for i, e in plain_points.items())
    return {"requests": nreq, "slots": slots, "lanes_per_slot": lanes, "energies_per_s": nreq / total,
            "batch_seconds": total, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "service_p50_ms": float(np.percentile(svc, 50)), "peak_footprint_bytes": st["peak"],
            "max_abs_dE2_vs_plain_fp64": d, "max_abs_dtotal_vs_plain_fp64": dt,
            "points_compared": len(plain_points),
            "precision": f"split-precision transforms, {split} INT8 slices" if split else "fp64",
            "eig_batch": eig_batch,
            "eig_window_s": eig_window, "eig_batches": st["eig_batches"], "eig_jobs": st["eig_jobs"]}


def split_steps(ctx, spec, args, fp64_out):
    """The same energies with split-precision (Ozaki INT8) integral transforms."""
    ctx.set_option("split_slices", args.split)
    for _ in range(max(1, args.warmup)):
        ctx.casci(spec.selectroot)
        out = ctx.caspt2(spec.selectroot, spec.eshift)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        ctx.casci(spec.selectroot)
        out = ctx.caspt2(spec.selectroot, spec.eshift)
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    fs, fd = ctx.transform_flops()
    ctx.set_option("split_slices", 0)
    return {"slices": args.split, "energies_per_s": args.steps / total, "ms_per_step": 1e3 * total / args.steps,
            "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
            "transform_seconds_last_step": out.timings[3], "fp64_transform_seconds_last_step": fp64_out.timings[3],
            "split_flop_share": fs / max(fs + fd, 1.0),
            "max_abs_dE2_vs_fp64_gpu": max(abs(out.e2[X] - fp64_out.e2[X]) for X in "ABCDEFGH"),
            "abs_dtotal_vs_fp64_gpu": abs(out.total - fp64_out.total),
            "stage_seconds_last_step": list(out.timings), "_e2": dict(out.e2), "_total": out.total}


def plain_points(ctx, spec, sp, nreq, idx):
    """Plain FP64 energies of the scan points idx (untimed), for the served-run comparison."""
    from dirac_caspt2_amd import synth
    direction = synth.scan_displacements(spec, spec.seed + 1)
    lams = np.linspace(0.0, 1.0, nreq)
    out = {}
    for i in idx:
        ctx.set_one_electron(synth.scan_point(sp, lams[i], direction), sp.ecore)
        ctx.casci(spec.selectroot)
        out[i] = ctx.caspt2(spec.selectroot, spec.eshift)
    return out


def copy_calibration(ctx, mib):
    from dirac_caspt2_amd.api import DeviceArray
    n = mib << 20
    src, dst = DeviceArray(ctx, nbytes=n), DeviceArray(ctx, nbytes=n)
    for _ in range(3):
        ctx.lib.dc2_memcpy(ctx.h, dst.ptr, src.ptr, n, 3)
    ctx.synchronize()
    src.free()
    dst.free()


def measured_traffic(workload, kernel, key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known size),
    or None when no summary for this kernel exists.  kernel: "k_zgemm" (key: the op pair) or
    "k_hetrd" (key: the matrix size)."""
    path = os.path.join(HERE, "profiles", f"{kernel}_{workload}_hbm_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        w = line.split()
        if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == str(key):
            return float(w[2])
    return None


def stage_hbm(workload):
    """Achieved HBM bandwidth of the bandwidth-bound stage kernels from the committed rocprofv3
    summary of this workload (profiles/stage_hbm_<workload>.txt: counter bytes per launch and
    the traced mean launch duration), or None."""
    path = os.path.join(HERE, "profiles", f"stage_hbm_{workload}.txt")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        w = line.split()
        if len(w) == 4 and w[0] == "stage_hbm":
            b, t = float(w[2]), float(w[3])
            out[w[1]] = {"bytes_per_launch": b, "mean_launch_us": 1e6 * t, "achieved_GBs": b / t / 1e9,
                         "frac_of_hbm_peak": b / t / 1e9 / PEAK_HBM_GBS}
    return out or None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def generate_all(specs, threads=8):
    from concurrent.futures import ThreadPoolExecutor

    from dirac_caspt2_amd import synth
    with ThreadPoolExecutor(threads) as ex:
        out = []
        for k, sp in enumerate(ex.map(synth.generate, specs)):
            out.append(sp)
            if k % 8 == 7:
                log(f"generated {k + 1}/{len(specs)}")
    return out


def batch_workload(args):
    """64 independent seeded instances (nsec ~ U[40,300], nact in {8,10,12} with nact/2
    electrons, ninact ~ U[4,24]) solved (a) one after another by the plain FP64 path and
    (b) all at once through the request server with cross-request eigensolve batching."""
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, Server
    from dirac_caspt2_amd.problem import load_factors
    specs = synth.batch_instances(args.instances)
    sps = generate_all(specs)
    plain = []
    t_plain = 0.0
    for k, (s, sp) in enumerate(zip(specs, sps)):
        ctx = Caspt2Context(0)
        load_factors(ctx, (s.ninact, s.nact, s.nsec, s.nelec), sp.h, sp.B, sp.ecore, on_device=True)
        if k == 0:                                  # warm the kernels once
            ctx.casci(1)
            ctx.caspt2(1, 0.0)
        t = time.perf_counter()
        ctx.casci(s.selectroot)
        plain.append(ctx.caspt2(s.selectroot, s.eshift))
        t_plain += time.perf_counter() - t
        ctx.close()
        if k % 16 == 15:
            log(f"plain {k + 1}/{len(specs)}")
    with Server(0, nslots=args.slots, lanes_per_slot=args.slot_lanes, eig_batch=args.eig_batch,
                eig_window=args.eig_window) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for s, sp in zip(specs, sps)]
        log("molecules resident")
        for k in range(min(args.slots, len(specs))):      # warm the slots
            srv.submit(ids[k], sps[k].h, sps[k].ecore, 1, 0.0, split_slices=args.split)
        for k in range(min(args.slots, len(specs))):
            srv.wait(300.0)
        t0 = time.perf_counter()
        sub = {srv.submit(ids[k], sp.h, sp.ecore, s.selectroot, s.eshift, split_slices=args.split): k
               for k, (s, sp) in enumerate(zip(specs, sps))}
        res = {}
        while len(res) < len(specs):
            r = srv.wait(300.0)
            if r is None or r.status:
                raise RuntimeError(f"batched request failed: {None if r is None else r.error}")
            res[sub[r.id]] = r
        t_batch = time.perf_counter() - t0
        st = srv.stats()
    even = [k for k, s in enumerate(specs) if s.nelec % 2 == 0]
    d_even = max(abs(res[k].e2[X] - plain[k].e2[X]) for k in even for X in "ABCDEFGH")
    d_odd = max((abs(res[k].e2[X] - plain[k].e2[X]) for k in range(len(specs)) if k not in even
                 for X in "ABCDEFGH"), default=0.0)
    lat = np.array([r.latency for r in res.values()]) * 1e3
    return {"metric": "CASPT2 energies/s, 64 independent instances batched vs sequential plain solves",
            "value": len(specs) / t_batch, "unit": "energies/s", "n_gpus": 1, "higher_is_better": True,
            "sequential_plain_energies_per_s": len(specs) / t_plain, "speedup": t_plain / t_batch,
            "batch_seconds": t_batch, "sequential_seconds": t_plain,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "slots": args.slots, "eig_batch": args.eig_batch, "eig_batches": st["eig_batches"],
            "eig_jobs": st["eig_jobs"], "split_slices": args.split,
            "max_abs_dE2_even_electron_instances": d_even,
            "max_abs_dE2_odd_electron_instances": d_odd,
            "note": "odd-electron instances (nact 10, 5 electrons) have Kramers-degenerate CASCI roots; "
                    "their energies depend on the eigenvector chosen inside the pair, as in the reference",
            "dtype": "f64 (complex128)", "data": "synthetic (seeded)",
            "config": {"workload": f"batch: {len(specs)} instances, seed 20243"}}


def serve_workload(args):
    """Poisson stream of CASPT2 requests over resident molecules (scan points), continuous
    batching with cross-request eigensolve batching and split-precision transforms."""
    import threading

    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, Server
    mols = synth.serve_molecules(args.molecules)
    sps = generate_all([s for s, _ in mols])
    dirs = [synth.scan_displacements(s, s.seed + 1) for s, _ in mols]
    rng = np.random.default_rng(20245)
    n = args.requests
    ras_mols = [k for k, (_, r) in enumerate(mols) if r is not None]
    cas_mols = [k for k, (_, r) in enumerate(mols) if r is None]
    share = getattr(args, "ras_share", 0.2)
    which = np.where(rng.uniform(size=n) < share, rng.choice(ras_mols, n), rng.choice(cas_mols, n))
    lams = rng.uniform(0.0, 1.0, n)
    gaps = rng.exponential(1.0 / args.rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    with Server(0, nslots=args.slots, lanes_per_slot=args.slot_lanes, eig_batch=args.eig_batch,
                eig_window=args.eig_window) as srv:
        ids = [srv.add_molecule((s.ninact, s.nact, s.nsec, s.nelec), sp.B) for (s, _), sp in zip(mols, sps)]
        for k, ((s, ras), sp) in enumerate(zip(mols, sps)):          # warm every molecule once
            srv.submit(ids[k], sp.h, sp.ecore, 1, 0.0, ras, split_slices=args.split)
        for _ in mols:
            r = srv.wait(300.0)
            if r is None or r.status:
                raise RuntimeError(f"warm-up request failed: {None if r is None else r.error}")
        log("molecules resident and warm")
        sub = {}
        lock = threading.Lock()

        def feeder():
            t0 = time.perf_counter()
            for i in range(n):
                dt = arrivals[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                m = int(which[i])
                s, ras = mols[m]
                h = synth.scan_point(sps[m], float(lams[i]), dirs[m])
                rid = srv.submit(ids[m], h, sps[m].ecore, 1, 0.0, ras, split_slices=args.split)
                with lock:
                    sub[rid] = i
        th = threading.Thread(target=feeder)
        t_start = time.perf_counter()
        th.start()
        res = {}
        last = time.perf_counter()
        while len(res) < n:
            r = srv.wait(300.0)
            if r is None or r.status:
                raise RuntimeError(f"served request failed: {None if r is None else r.error}")
            with lock:
                res[sub[r.id]] = r
            if time.perf_counter() - last > 30:
                log(f"served {len(res)}/{n}")
                last = time.perf_counter()
        th.join()
        t_total = time.perf_counter() - t_start
        st = srv.stats()
    lat = np.array([res[i].latency for i in range(n)]) * 1e3
    # spot check: a few requests recomputed alone by the plain FP64 path
    from dirac_caspt2_amd.problem import load_factors
    dmax = 0.0
    for i in range(0, n, max(1, n // 6))[:6]:
        m = int(which[i])
        s, ras = mols[m]
        ctx = Caspt2Context(0)
        load_factors(ctx, (s.ninact, s.nact, s.nsec, s.nelec), synth.scan_point(sps[m], float(lams[i]), dirs[m]),
                     sps[m].B, sps[m].ecore, on_device=True)
        if ras:
            ctx.set_ras(*ras)
        ctx.casci(1)
        e = ctx.caspt2(1, 0.0)
        ctx.close()
        dmax = max(dmax, max(abs(e.e2[X] - res[i].e2[X]) for X in "ABCDEFGH"))
    return {"metric": "served CASPT2 energies/s + p99 latency (Poisson arrivals)", "value": n / t_total,
            "unit": "energies/s", "n_gpus": 1, "higher_is_better": True,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "offered_rate_per_s": args.rate, "requests": n, "molecules": len(mols),
            "ras_share": float(np.mean([mols[int(w)][1] is not None for w in which])),
            "slots": args.slots, "eig_batch": args.eig_batch, "eig_batches": st["eig_batches"],
            "eig_jobs": st["eig_jobs"], "split_slices": args.split, "peak_footprint_bytes": st["peak"],
            "spot_check_max_abs_dE2_vs_plain_fp64": dmax, "dtype": "f64 (complex128)",
            "data": "synthetic (seeded scan points)",
            "config": {"workload": f"serve: {n} requests at {args.rate}/s over {len(mols)} molecules (nsec ~ "
                                   f"U[40,280], nact in {{8,12,14}} with nact/2 electrons, ninact ~ U[4,24]); a "
                                   f"{share:.0%} share of RASPT2 requests with RAS1/RAS2/RAS3 = 4/4/4 spinors, at "
                                   f"most 2 holes/particles (6/8/6 needs densities beyond the full-string-space "
                                   f"engine); {AMPLITUDE_NOTE}"}}


def timed_steps(ctx, spec, steps, warmup, split):
    """warmup untimed energies, then exactly `steps` timed ones (each synchronous: it returns
    after the energies are on the host); split: INT8 slices of the transforms (0: FP64)."""
    ctx.set_option("split_slices", split)
    for _ in range(warmup):
        ctx.casci(spec.selectroot)
        out = ctx.caspt2(spec.selectroot, spec.eshift)
    ctx.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        ctx.casci(spec.selectroot)
        out = ctx.caspt2(spec.selectroot, spec.eshift)
        lat.append(time.perf_counter() - t)
    ctx.synchronize()
    total = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    st = out.timings
    return out, {"energies_per_s": steps / total, "ms_per_step": 1e3 * total / steps, "seconds": total,
                 "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
                 "stage_seconds_last_step": {"casci": st[0], "fock_fockdiag_tracic": st[1], "densities": st[2],
                                             "transforms": st[3],
                                             **{f"subspace_{X}": st[4 + k] for k, X in enumerate("ABCDEFGH")}}}


def golden_diffs(workload, out):
    """|dE2| per subspace and |dtotal| against the oracle's stored energies of this workload."""
    gold = os.path.join(HERE, "tests", "golden", f"synthetic_{workload}.json")
    if not os.path.exists(gold):
        return None
    g = json.load(open(gold))
    return {"abs_dE2": {X: abs(out.e2[X] - g["e2"][X]) for X in "ABCDEFGH"},
            "max_abs_dE2": max(abs(out.e2[X] - g["e2"][X]) for X in "ABCDEFGH"),
            "abs_dtotal": abs(out.total - g["total"])}


def hetrd_roofline(ctx, ht, steps, workload):
    """Roofline of this library's tridiagonalisation over the timed steps: the metric size with
    the most device time; 16/3 n^3 flop per launch (LAPACK's zhetrd count)."""
    from dirac_caspt2_amd.api import hetrd_stats
    best = None
    for n in ht:
        calls, sec = hetrd_stats(n)
        if calls and (best is None or sec > best[2]):
            best = (n, calls, sec)
    if best is None:
        return None
    n, calls, sec = best
    fl = 16.0 / 3.0 * n ** 3
    ach = fl * calls / sec / 1e12
    rows = next(r for lim, r in ((128, 2), (256, 4), (512, 8), (768, 12), (832, 13), (1024, 16), (1280, 20),
                                 (1536, 24), (1792, 28), (2048, 32)) if n <= lim)
    # two columns per wave up to n = 1024, and up to 1280 on the subspace lanes of an energy
    return {"kernel": f"k_hetrd<{rows},{2 if n <= 1280 else 1}> (Householder tridiagonalisation, n = {n}, one grid "
                      f"barrier per column)",
            "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP64_TFLOPS, "traffic": measured_traffic(workload, "k_hetrd", n),
            "limited_by": "latency: one grid barrier and one round of device-coherent exchanges per column, "
                          "not FP64 arithmetic (the roofline's bound) or HBM",
            "launches_per_step": calls / steps,
            "avg_launch_us": 1e6 * sec / calls, "algorithmic_flops_per_launch": fl, "seconds_per_step": sec / steps,
            "n": n}


def gemm_accounting(ctx, step, workload, nsteps):
    """Per-launch GEMM accounting (two HIP events per launch; Davidson graphs off) over extra
    untimed steps: the dominant k_zgemm2 instantiation and the all-GEMM rate."""
    ctx.set_profiling(True)
    for _ in range(nsteps):
        step()
    ctx.synchronize()
    g_all = ctx.gemm_stats("all")
    g_tr = ctx.gemm_stats("transform")
    ops = {f"{oa}{ob}": ctx.gemm_op_stats(oa, ob) for oa in "NTC" for ob in "NRTC"}
    ctx.set_profiling(False)
    dom_op, dom = max(ops.items(), key=lambda kv: kv[1]["seconds"])
    ach = dom["flops"] / dom["seconds"] / 1e12 if dom["seconds"] > 0 else 0.0
    return {"kernel": f"k_zgemm2<'{dom_op[0]}','{dom_op[1]}',16,1> (MFMA FP64 complex GEMM)",
            "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP64_TFLOPS, "traffic": measured_traffic(workload, "k_zgemm", dom_op),
            "launches_per_step": dom["launches"] / nsteps,
            "avg_launch_us": 1e6 * dom["seconds"] / max(dom["launches"], 1),
            "algorithmic_flops_per_launch": dom["flops"] / max(dom["launches"], 1),
            "seconds_per_step": dom["seconds"] / nsteps,
            "all_gemm_launches_per_step": g_all["launches"] / nsteps,
            "all_gemm_achieved_tflops": g_all["flops"] / g_all["seconds"] / 1e12 if g_all["seconds"] > 0 else 0.0,
            "transform_launches_achieved_tflops": g_tr["flops"] / g_tr["seconds"] / 1e12 if g_tr["seconds"] > 0
            else 0.0}


def eig_workload(args):
    """The library's Hermitian eigensolver (dc2_cdiag: tridiagonalisation, tridiagonal divide
    and conquer, back-transformation) on seeded rank-deficient Gram matrices G = X X^H / r, the
    shape of the A/C metrics (diag.F90:66-128 is the routine it replaces).  Wall time per solve
    includes the host-to-device and device-to-host copies of the matrix."""
    from dirac_caspt2_amd.api import Caspt2Context, cdiag, hetrd_stats
    rng = np.random.default_rng(5)
    sizes = ((792, 500), (1274, 900), (2002, 1274))
    out = {}
    with Caspt2Context(0) as ctx:
        ctx.set_option("eigensolver", "native")
        for n, rank in sizes:
            X = rng.normal(size=(n, rank)) + 1j * rng.normal(size=(n, rank))
            A = X @ X.conj().T / rank
            cdiag(ctx, A)
            hetrd_stats(0)
            t = time.perf_counter()
            for _ in range(args.steps):
                w, V, _ = cdiag(ctx, A)
            dt = (time.perf_counter() - t) / args.steps
            calls, sec = hetrd_stats(n)
            top = V[:, -8:]
            out[str(n)] = {"rank": rank, "ms_per_solve": 1e3 * dt, "hetrd_ms": 1e3 * sec / max(calls, 1),
                           "max_residual_top8": float(np.abs(A @ top - top * w[-8:]).max())}
            log(f"n {n}: {1e3 * dt:.2f} ms per solve")
    big = out[str(sizes[-1][0])]
    return {"metric": "Hermitian eigensolves/s of a rank-deficient Gram matrix (n 2002, rank 1274)",
            "value": 1e3 / big["ms_per_solve"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "higher_is_better": True, "dtype": "f64 (complex128)", "data": "synthetic (seeded X, G = X X^H / rank)",
            "config": {"workload": "eig: n/rank 792/500, 1274/900, 2002/1274 through dc2_cdiag, host copies "
                                   "included"},
            "sizes": out}


def mdcint_workload(args):
    """The MDCINT unique set -> MDCINTNEW records conversion on the device (dc2_mdcint_records,
    create_newmdcint.F90:86-177) on a seeded random set at the 324-spinor size: 4000 records of
    2500 entries (10,000,000), Kramers labels in 1..162 with random signs, a random Kramers
    table and CAS order; SCHEME 4 and 6, three calls each, records counted, not copied back.
    Wall time per call includes the host-to-device upload of the entries."""
    from dirac_caspt2_amd.api import Caspt2Context
    rng = np.random.default_rng(11)
    nkr, nrec, nzr = 162, 4000, 2500
    nmo = 2 * nkr
    lab = lambda n: rng.integers(1, nkr + 1, n) * rng.choice([-1, 1], n)
    ikr, jkr = lab(nrec), lab(nrec)
    nz = np.full(nrec, nzr, np.int64)
    ne = int(nz.sum())
    indk, indl = lab(ne), lab(ne)
    rr = rng.standard_normal(ne)
    perm = rng.permutation(nmo) + 1
    kr = np.zeros(2 * nkr + 1, np.int64)
    for m in range(1, nkr + 1):
        kr[nkr + m] = perm[2 * (m - 1)]
        kr[nkr - m] = perm[2 * (m - 1) + 1]
    d2c = np.zeros(nmo + 1, np.int64)
    d2c[1:] = rng.permutation(nmo) + 1
    calls = {}
    with Caspt2Context(0) as ctx:
        for scheme in (4, 6):
            ts = []
            for _ in range(3):
                t = time.perf_counter()
                nout = ctx.mdcint_records(ikr, jkr, nz, indk, indl, rr, None, nkr, kr, d2c, scheme,
                                          want_records=False)
                ts.append(time.perf_counter() - t)
            calls[f"scheme_{scheme}"] = {"records": int(nout), "ms_per_call": 1e3 * min(ts)}
            log(f"scheme {scheme}: {ne} entries -> {nout} records, {1e3 * min(ts):.1f} ms per call")
    best = min(v["ms_per_call"] for v in calls.values())
    return {"metric": "MDCINT entries converted to MDCINTNEW records per second", "value": ne / (best * 1e-3),
            "unit": "entries/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64 / int32",
            "data": "synthetic (seeded random MDCINT set, 10,000,000 entries)",
            "config": {"workload": "mdcint: 4000 records x 2500 entries, nkr 162, SCHEME 4 and 6, host upload "
                                   "included"},
            "calls": calls}


AMPLITUDE_NOTE = ("two-electron factors at amplitude 0.25 x CN(0,1) (amplitude 1 puts a secondary Fock eigenvalue "
                  "below the inactive ones: an intruder state)")


def workload_name(name, spec):
    return (f"{name}: ninact {spec.ninact}, nact {spec.nact}, nsec {spec.nsec}, nelec {spec.nelec}, "
            f"seed {spec.seed}, selectroot {spec.selectroot}; {AMPLITUDE_NOTE}")


def run_workload(name, args, steps, warmup, served, slots, autotune_on, accounting, lanes=1, eig_batch=64):
    """FP64 and split-precision steps of one workload, its served run and (optionally) the
    roofline of its dominant kernels.  Returns (result dict, generated problem, spec)."""
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, hetrd_stats
    from dirac_caspt2_amd.problem import load_factors
    spec = workload_spec(name)
    sp = synth.generate(spec)
    ctx = Caspt2Context(0)
    load_factors(ctx, (spec.ninact, spec.nact, spec.nsec, spec.nelec), sp.h, sp.B, sp.ecore, on_device=True)

    def step():
        ctx.casci(spec.selectroot)
        return ctx.caspt2(spec.selectroot, spec.eshift)

    res = {"workload": workload_name(name, spec)}
    if autotune_on:   # untimed: the GEMM shapes of one step, each tuned on its own
        from dirac_caspt2_amd import autotune
        at = autotune.tune_workload(ctx, step)
        res["gemm_autotune"] = {k: at[k] for k in ("shapes", "changed", "builtin_us", "tuned_us")}
        log(f"{name} autotune: {at['shapes']} GEMM shapes, {at['changed']} with a tuned choice")
    fp64_out, fp64 = timed_steps(ctx, spec, steps, warmup, 0)
    log(f"{name} fp64: {fp64['energies_per_s']:.2f} energies/s")
    hetrd_stats(0)
    split_out, split = timed_steps(ctx, spec, steps, warmup, args.split)
    log(f"{name} split: {split['energies_per_s']:.2f} energies/s")
    fs, fd = ctx.transform_flops()
    split.update({"slices": args.split, "split_flop_share": fs / max(fs + fd, 1.0),
                  "max_abs_dE2_vs_fp64_gpu": max(abs(split_out.e2[X] - fp64_out.e2[X]) for X in "ABCDEFGH"),
                  "abs_dtotal_vs_fp64_gpu": abs(split_out.total - fp64_out.total)})
    res.update({"fp64": fp64, "split": split, "casci_ndet_roots_davidson_iterations": ctx.casci_info(),
                "eig_fallbacks": ctx.eig_fallbacks(), "energies_last_step": {"total": split_out.total,
                                                                              "e2": split_out.e2}})
    go = golden_diffs(name, fp64_out), golden_diffs(name, split_out)
    if go[0] is not None:
        res["vs_oracle"] = {"fp64_gpu": go[0], "split": go[1]}
    if accounting:
        # every eigenproblem size of the path: S (dimn) and B~ (dimm) of each subspace
        # and the determinant-space and tuple bases of the A/C metrics
        ht = sorted({int(v) for X in "ABCDEFGH" for v in split_out.info.get(X, ())[:2] if int(v) >= 3} |
                    {comb(spec.nact, spec.nelec + 1), comb(spec.nact, spec.nelec - 1),
                     spec.nact * comb(spec.nact, 2)})
        hr = hetrd_roofline(ctx, ht, steps, name)
        gr = gemm_accounting(ctx, step, name, max(1, min(steps, 5)))
        if hr is not None and hr["seconds_per_step"] > gr["seconds_per_step"]:
            hr["dominant_gemm"] = gr
            res["roofline"] = hr
        else:
            res["roofline"] = gr
            if hr is not None:
                res["roofline"]["tridiagonalisation"] = hr
        res["roofline"]["stage_hbm"] = stage_hbm(name)
    if served:
        idx = sorted({0, served // 3, (2 * served) // 3, served - 1})
        ctx.set_option("split_slices", 0)
        pts = plain_points(ctx, spec, sp, served, idx)
        ctx.close()
        srv = served_run(spec, sp, served, slots, lanes, pts, eig_batch, args.eig_window, args.split,
                         args.native_min)
        srv["speedup_vs_plain_fp64"] = srv["energies_per_s"] / fp64["energies_per_s"]
        res["served"] = srv
        log(f"{name} served: {srv['energies_per_s']:.2f} energies/s")
    else:
        ctx.close()
    return res, sp, spec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="split", choices=["small", "fp64", "split", "batch", "serve", "eig", "mdcint"],
                    help="split: 324 spinors (30/14/280, 6 electrons, the default line); fp64: 192 spinors "
                         "(16/12/164); small: 52 spinors; batch: 64 independent instances; serve: Poisson stream; "
                         "eig: the eigensolver alone on Gram matrices of the A/C metric shape; mdcint: the "
                         "MDCINT record conversion")
    ap.add_argument("--secondary", default="fp64", help="workload of the secondary object of the line ('' : none)")
    ap.add_argument("--secondary-steps", type=int, default=10)
    ap.add_argument("--instances", type=int, default=64, help="batch workload: independent instances")
    ap.add_argument("--requests", type=int, default=2000, help="serve stream: requests")
    ap.add_argument("--rate", type=float, default=25.0, help="serve stream: Poisson arrival rate (1/s)")
    ap.add_argument("--molecules", type=int, default=12, help="serve stream: resident molecules")
    ap.add_argument("--ras-share", type=float, default=0.2, help="serve stream: share of RASPT2 requests")
    ap.add_argument("--stream", type=int, default=1, help="default line: include the Poisson stream object (0: no)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="skip the GEMM autotuning of the warm-up")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run 3 device-to-device copies of MIB MiB (a plain copy kernel "
                         "of known size, the reference for the HBM counters)")
    ap.add_argument("--served", type=int, default=None, metavar="N",
                    help="N scan points of the same molecule through the request server, all submitted at once "
                         "(default 64 at 324 spinors, 128 at 192; 0: skip)")
    ap.add_argument("--slots", type=int, default=None, help="concurrent slots of the served runs (default: 8 at "
                                                            "324 spinors, 32 for the Poisson stream, 64 otherwise)")
    ap.add_argument("--slot-lanes", type=int, default=None,
                    help="subspace lanes per slot in the served runs (default: 2 at 324 spinors, 1 otherwise)")
    ap.add_argument("--split", type=int, default=6,
                    help="INT8 slices of the split-precision transforms; 6: |dE2| <= 3e-11 Eh against the "
                         "oracle at 324 spinors (5: 3e-9, too close to the 1e-8 bound for the totals; 7: 3e-13)")
    ap.add_argument("--eig-batch", type=int, default=None,
                    help="eigensolves batched across served requests (<=1: off; default: off at 324 spinors and "
                         "for the Poisson stream, where each slot's own eigensolver is faster, 64 otherwise)")
    ap.add_argument("--native-min", type=int, default=0,
                    help="served runs: eigenproblems of at least this size use this library's eigensolver instead "
                         "of the cross-request batched rocSOLVER one (0: never)")
    ap.add_argument("--eig-window", type=float, default=None,
                    help="batching window of the served eigensolves (s); groups are mostly flushed earlier, when "
                         "every running request waits for one.  Default 0.1 for the runs that submit all requests "
                         "at once, 0.003 for the Poisson stream")
    args = ap.parse_args()
    big_lanes, big_eig_batch = args.slot_lanes or 2, 0 if args.eig_batch is None else args.eig_batch
    # the Poisson stream: 32 slots with their own eigensolvers (p50 48 ms, p99 263 ms at 25/s
    # against 164 / 607 ms with 64 slots and cross-request batched rocSOLVER eigensolves)
    stream_eig_batch = 0 if args.eig_batch is None else args.eig_batch
    args.slot_lanes = args.slot_lanes or 1
    args.eig_batch = 64 if args.eig_batch is None else args.eig_batch
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is documented in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    window = args.eig_window
    if args.workload == "eig":
        print(json.dumps(eig_workload(args)))
        return
    if args.workload == "mdcint":
        print(json.dumps(mdcint_workload(args)))
        return
    if args.workload == "batch":
        args.eig_window = 0.1 if window is None else window
        args.slots = args.slots or 64
        print(json.dumps(batch_workload(args)))
        return
    if args.workload == "serve":
        args.eig_window = 0.003 if window is None else window
        args.slots = args.slots or 32
        args.eig_batch = stream_eig_batch
        print(json.dumps(serve_workload(args)))
        return
    args.eig_window = 0.1 if window is None else window
    big = args.workload == "split"
    served = args.served if args.served is not None else (64 if big else 128)
    # 324 spinors: 8 slots of 2 lanes with their own eigensolvers (19.1 energies/s served
    # against 15.8 with 32 slots of 1 lane and cross-request batched rocSOLVER eigensolves)
    slots = args.slots or (8 if big else 64)
    main_res, sp, spec = run_workload(args.workload, args, args.steps, args.warmup, served, slots,
                                      not args.no_autotune, True, big_lanes if big else args.slot_lanes,
                                      big_eig_batch if big else args.eig_batch)
    split, fp64 = main_res["split"], main_res["fp64"]
    result = {
        "metric": "CASPT2 energies/s + p99 latency, split-prec vs FP64 GPU at |\u0394E2|<1e-8 Eh",
        "value": split["energies_per_s"],
        "unit": "energies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": split["ms_per_step"],
        "p50_ms": split["p50_ms"],
        "p99_ms": split["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 (complex128); integral rotation in split precision ({args.split} INT8 slices, INT32 "
                 f"accumulation), eigensolves and E2 in f64",
        "data": "synthetic (seeded Kramers-paired integrals, in HBM before timing)",
        "config": {"workload": main_res["workload"], "global_batch": 1, "parallelism": "single GPU"},
        "step": "one CASPT2 energy (CASCI + rotation + subspaces A-H) with split-precision transforms; the "
                "fp64 object holds the same steps with FP64 transforms",
        "fp64": fp64,
        "split_vs_fp64_rate": split["energies_per_s"] / fp64["energies_per_s"],
    }
    for k in ("split", "vs_oracle", "roofline", "served", "gemm_autotune", "casci_ndet_roots_davidson_iterations",
              "eig_fallbacks", "energies_last_step"):
        if k in main_res:
            result[k] = main_res[k]
    if "served" in result:
        result["served_split_vs_plain_fp64"] = result["served"]["speedup_vs_plain_fp64"]
    if args.copy_calibration:
        from dirac_caspt2_amd.api import Caspt2Context
        with Caspt2Context(0) as cc:
            copy_calibration(cc, args.copy_calibration)
    if args.secondary and args.secondary != args.workload:
        sec, _, _ = run_workload(args.secondary, args, args.secondary_steps, max(1, min(args.warmup, 2)),
                                 128, 64, False, False, args.slot_lanes, args.eig_batch)
        result["secondary"] = sec
    if args.stream and big:
        args.eig_window = 0.003 if window is None else window
        args.slots = 32
        args.eig_batch = stream_eig_batch
        result["serve_stream"] = serve_workload(args)
    if not args.no_cpu_baseline:
        ref, cb = cpu_baseline(sp)
        result["cpu_baseline"] = cb
        out_e2 = main_res["energies_last_step"]["e2"]
        result["parity_vs_oracle_max_abs_e2"] = max(abs(out_e2[X] - ref.e2[X]) for X in "ABCDEFGH")
        result["parity_vs_oracle_total"] = abs(main_res["energies_last_step"]["total"] - ref.total)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
