"""Benchmark: CASPT2 energies/s (+ p50/p99 latency) of the plain-FP64 GPU path.

One step = one complete CASPT2 energy (CASCI: determinant strings, CI matrix,
eigensolve; CASPT2: Fock build and diagonalisation, CI rotation, densities,
integral transforms, subspaces A-H) for the 192-spinor workload
(16 inactive / 12 active with 6 electrons (924 determinants) / 164 secondary,
complex128, seeded synthetic Kramers-paired integrals), with the integrals
already resident in HBM.  Only one GPU is used.

    python bench.py [--steps K] [--warmup W] [--workload fp64|small|split|batch|serve]
                    [--served N] [--slots S] [--split SLICES] [--no-autotune]
                    [--no-cpu-baseline] [--copy-calibration MIB]

Prints ONE JSON line.  With the default workload the line also carries the split-precision
rate of the same steps, the served run (N scan points through the request server), the
roofline of the dominant kernel and the CPU baseline.
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
# when the runtime starts, i.e. at the first library call)
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


def served_run(spec, sp, nreq, slots, lanes, plain_points, eig_batch=0, eig_window=0.002, split=0):
    """Scan points of the same molecule through the request server, all submitted at once:
    energies/s over the whole batch, per-request latency, and the largest |dE2| against the
    plain FP64 path on the points in plain_points ({index: Caspt2Energies})."""
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Server
    direction = synth.scan_displacements(spec, spec.seed + 1)
    hs = [synth.scan_point(sp, lam, direction) for lam in np.linspace(0.0, 1.0, nreq)]
    with Server(0, nslots=slots, lanes_per_slot=lanes, eig_batch=eig_batch, eig_window=eig_window) as srv:
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
    which = rng.integers(0, len(mols), n)
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
            "config": {"workload": f"serve: {n} requests at {args.rate}/s over {len(mols)} molecules"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="fp64", choices=["small", "fp64", "split", "batch", "serve"])
    ap.add_argument("--instances", type=int, default=64, help="batch workload: independent instances")
    ap.add_argument("--requests", type=int, default=2000, help="serve workload: requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve workload: Poisson arrival rate (1/s)")
    ap.add_argument("--molecules", type=int, default=12, help="serve workload: resident molecules")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="skip the GEMM autotuning of the warm-up")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run 3 device-to-device copies of MIB MiB (a plain copy kernel "
                         "of known size, the reference for the HBM counters)")
    ap.add_argument("--served", type=int, default=128, metavar="N",
                    help="after the plain steps, N scan points of the same molecule through the request server "
                         "(0: skip)")
    ap.add_argument("--slots", type=int, default=64, help="concurrent slots of the served run")
    ap.add_argument("--slot-lanes", type=int, default=1, help="subspace lanes per slot in the served run")
    ap.add_argument("--split", type=int, default=6,
                    help="INT8 slices of the split-precision transforms (0: off); 6: |dE2| <= 3e-11 Eh against the "
                         "oracle at 324 spinors (5: 3e-9, too close to the 1e-8 bound for the totals; 7: 3e-13)")
    ap.add_argument("--eig-batch", type=int, default=64, help="eigensolves batched across served requests (<=1: off)")
    ap.add_argument("--eig-window", type=float, default=None,
                    help="batching window of the served eigensolves (s); groups are mostly flushed earlier, when "
                         "every running request waits for one.  Default 0.1 for the runs that submit all requests "
                         "at once (0.003: 72, 0.03: 79, 0.1-0.25: 83-85 energies/s at 64 slots), 0.003 for the "
                         "Poisson stream, whose p50/p99 latency at 15 requests/s it keeps at 161/609 ms (0.1: 361/929)")
    args = ap.parse_args()
    if args.eig_window is None:
        args.eig_window = 0.003 if args.workload == "serve" else 0.1
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is documented in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        print(json.dumps(batch_workload(args)))
        return
    if args.workload == "serve":
        print(json.dumps(serve_workload(args)))
        return
    from dirac_caspt2_amd import synth
    from dirac_caspt2_amd.api import Caspt2Context, hetrd_stats
    from dirac_caspt2_amd.problem import load_factors

    spec = workload_spec(args.workload)
    sp = synth.generate(spec)
    ctx = Caspt2Context(0)
    load_factors(ctx, (spec.ninact, spec.nact, spec.nsec, spec.nelec), sp.h, sp.B, sp.ecore, on_device=True)

    def step():
        ctx.casci(spec.selectroot)
        return ctx.caspt2(spec.selectroot, spec.eshift)

    tuned = None
    if not args.no_autotune:   # untimed: the GEMM shapes of one step, each tuned on its own
        from dirac_caspt2_amd import autotune
        at = autotune.tune_workload(ctx, step)
        tuned = {k: at[k] for k in ("shapes", "changed", "builtin_us", "tuned_us")}
        log(f"autotune: {at['shapes']} GEMM shapes, {at['changed']} with a tuned choice, "
            f"{at['builtin_us']:.0f} -> {at['tuned_us']:.0f} us per set of launches")
    for _ in range(args.warmup):
        out = step()
    ctx.synchronize()
    hetrd_stats(0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        out = step()            # synchronous: returns after the energies are on the host
        lat.append(time.perf_counter() - t)
    ctx.synchronize()
    total = time.perf_counter() - t0
    # the tridiagonalisation is timed with HIP events around every launch, timed steps included
    n_g = max(comb(spec.nact, spec.nelec + 1), comb(spec.nact, spec.nelec - 1))
    ht_calls, ht_sec = hetrd_stats(n_g)
    # per-launch GEMM accounting (two HIP events per launch, Davidson graphs off) over extra
    # untimed steps, so that it does not slow the timed ones
    prof_steps = max(1, min(args.steps, 5))
    ctx.set_profiling(True)
    for _ in range(prof_steps):
        step()
    ctx.synchronize()
    g_all = ctx.gemm_stats("all")
    g_tr = ctx.gemm_stats("transform")
    ops = {f"{oa}{ob}": ctx.gemm_op_stats(oa, ob) for oa in "NTC" for ob in "NRTC"}
    ctx.set_profiling(False)
    dom_op, dom = max(ops.items(), key=lambda kv: kv[1]["seconds"])
    stage = out.timings
    lat_ms = np.array(lat) * 1e3

    # Roofline of the dominant hand-written kernel: the one with the largest HIP-event time per
    # step among the k_zgemm<opA,opB> instantiations (over the accounting steps) and the
    # tridiagonalisation of the largest metric (over the timed steps) (the A/C determinant-space G, n = C(nact,
    # nelec+-1)).  Algorithmic flops per launch: 8*M*N*K per complex GEMM, 16/3 n^3 for the
    # Householder tridiagonalisation of an n x n Hermitian matrix (LAPACK's zhetrd count).
    achieved_g = dom["flops"] / dom["seconds"] / 1e12 if dom["seconds"] > 0 else 0.0
    ach_all = g_all["flops"] / g_all["seconds"] / 1e12 if g_all["seconds"] > 0 else 0.0
    ach_tr = g_tr["flops"] / g_tr["seconds"] / 1e12 if g_tr["seconds"] > 0 else 0.0
    gemm_roof = {"kernel": f"k_zgemm2<'{dom_op[0]}','{dom_op[1]}',16,1> (MFMA FP64 complex GEMM)",
                 "bound": "mfma", "achieved": achieved_g, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": achieved_g / PEAK_FP64_TFLOPS, "traffic": measured_traffic(args.workload, "k_zgemm", dom_op),
                 "launches_per_step": dom["launches"] / prof_steps,
                 "avg_launch_us": 1e6 * dom["seconds"] / max(dom["launches"], 1),
                 "algorithmic_flops_per_launch": dom["flops"] / max(dom["launches"], 1),
                 "seconds_per_step": dom["seconds"] / prof_steps}
    if ht_calls > 0 and ht_sec / args.steps > dom["seconds"] / prof_steps:
        fl = 16.0 / 3.0 * n_g ** 3
        ach = fl * ht_calls / ht_sec / 1e12
        rows = next(r for lim, r in ((128, 2), (256, 4), (512, 8), (768, 12), (832, 13), (1024, 16)) if n_g <= lim)
        roof = {"kernel": f"k_hetrd<{rows}> (Householder tridiagonalisation, n = {n_g}, one grid barrier per column)",
                "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP64_TFLOPS, "traffic": measured_traffic(args.workload, "k_hetrd", n_g),
                "launches_per_step": ht_calls / args.steps, "avg_launch_us": 1e6 * ht_sec / ht_calls,
                "algorithmic_flops_per_launch": fl, "seconds_per_step": ht_sec / args.steps,
                "dominant_gemm": gemm_roof}
    else:
        roof = gemm_roof
    roof["all_gemm_launches_per_step"] = g_all["launches"] / prof_steps
    roof["all_gemm_achieved_tflops"] = ach_all
    roof["transform_launches_achieved_tflops"] = ach_tr
    result = {
        "metric": "CASPT2 energies/s + p99 latency, split-prec vs FP64 GPU at |\u0394E2|<1e-8 Eh",
        "value": args.steps / total,
        "unit": "energies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (complex128)",
        "data": "synthetic (seeded Kramers-paired integrals, in HBM before timing)",
        "config": {"workload": f"{args.workload}: ninact {spec.ninact}, nact {spec.nact}, nsec {spec.nsec}, "
                               f"nelec {spec.nelec}, seed {spec.seed}, selectroot {spec.selectroot}",
                   "global_batch": 1, "parallelism": "single GPU"},
        "stage_seconds_last_step": {"casci": stage[0], "fock_fockdiag_tracic": stage[1], "densities": stage[2],
                                    "transforms": stage[3],
                                    **{f"subspace_{X}": stage[4 + k] for k, X in enumerate("ABCDEFGH")}},
        "energies_last_step": {"total": out.total, "e2": out.e2},
        "roofline": roof,
        "eig_fallbacks": ctx.eig_fallbacks(),
        "gemm_autotune": tuned,
        "casci_ndet_roots_davidson_iterations": ctx.casci_info(),
    }
    if args.split:
        result["split"] = split_steps(ctx, spec, args, out)
        s_e2, s_total = result["split"].pop("_e2"), result["split"].pop("_total")
        gold = os.path.join(HERE, "tests", "golden", f"synthetic_{args.workload}.json")
        if os.path.exists(gold):     # the oracle's energies of this seeded workload
            g = json.load(open(gold))
            result["vs_oracle"] = {
                "fp64_gpu_abs_dE2": {X: abs(out.e2[X] - g["e2"][X]) for X in "ABCDEFGH"},
                "split_abs_dE2": {X: abs(s_e2[X] - g["e2"][X]) for X in "ABCDEFGH"},
                "fp64_gpu_abs_dtotal": abs(out.total - g["total"]),
                "split_abs_dtotal": abs(s_total - g["total"]),
                "oracle_seconds_8_cores": g.get("oracle_seconds")}
    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)
    if args.served:
        idx = sorted({0, args.served // 3, (2 * args.served) // 3, args.served - 1})
        pts = plain_points(ctx, spec, sp, args.served, idx)
        ctx.close()
        srv = served_run(spec, sp, args.served, args.slots, args.slot_lanes, pts, args.eig_batch, args.eig_window,
                         args.split)
        srv["speedup_vs_plain_fp64"] = srv["energies_per_s"] / result["value"]
        result["served"] = srv
    else:
        ctx.close()
    if not args.no_cpu_baseline:
        ref, cb = cpu_baseline(sp)
        result["cpu_baseline"] = cb
        result["parity_vs_oracle_max_abs_e2"] = max(abs(out.e2[X] - ref.e2[X]) for X in "ABCDEFGH")
        result["parity_vs_oracle_total"] = abs(out.total - ref.total)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
