#!/usr/bin/env python3
"""Throughput of the KDTree hot path on one MI355X.

Workload (default): N = 256^3 = 16,777,216 clustered particles (half uniform
background, half in 2000 NFW halos; periodic box L=100), k = 32 nearest
neighbours + cubic-spline SPH density and smoothing length (CalcDensity
semantics), and 3-D FOF at b = 0.2 x mean spacing.  Inputs are generated
on the host from a seed, copied to HBM once, and the tree is built before the
timed region.  One step = one kNN + SPH density pass over all N particles
(neighbour IDs, rho and h written to HBM).

Prints ONE JSON line: metric/value = Mqueries/s of kNN + SPH density for the
headline form (mixed precision when available, else plain FP64), plus the
plain-FP64 rate, FOF Mparticles/s, build time, the roofline of the dominant
kernel (k_knn, HIP events on its launch stream) and the CPU baseline (the
oracle's OpenMP restatement of the reference, timed on a bounded sample).

The same line also carries, so that one default run checks every target:
the differentiation form at N=1M (forward/backward over a stored list and
with the list recomputed per block of tree slots), the 6-D phase-space
workload at N=16.8M (k=64, precision and plain forms, two steps each), the
host CPU model, and the serial CalcDensity rate of the oracle on a sample.

Other measured workloads (--workload), each printing its own JSON line:
  phase6d  N=16.8M clustered particles with velocities, metric (x/sigma_x,
           v/sigma_v): 6-D phase-space tree, k=64 kNN + 6-D SPH density,
           plus 3-D FOF at b=0.2 mean spacing on the same particles and
           FOFCriterion with FOF6d on the phase-space tree.
  adjoint  N=1M clustered, lognormal masses: SPH density over a frozen k=32
           list and its hand-written adjoint (d sum log rho / d x, d m), in
           input (ID) order and in tree order; list memory vs recompute time.
  large    N=134M clustered, lognormal masses: build, mixed kNN + density
           (k=32), the density adjoint with the list recomputed per block of
           tree slots (and, for comparison, over a stored list), and FOF.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMD16 x 2.4 GHz, one wave64 VALU instruction
# per SIMD every 4 cycles = 614.4 G wave-instructions/s (x 64 lanes x 2 FLOP
# per FMA x 2 packed = the 157.3 TFLOP/s FP32-vector peak)
PEAK_VALU_GINST = 614.4


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256 ** 3)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fof", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=2, help="time every s-th tree slot on the CPU")
    ap.add_argument("--forms", default="plain,mixed", help="forms to time (plain, mixed)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates rocprof byte counters) and exit")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per launch of the dominant kernel (from rocprofv3 --pmc runs)")
    ap.add_argument("--workload", default="knn3d", choices=["knn3d", "phase6d", "adjoint", "large"])
    ap.add_argument("--no-extras", action="store_true",
                    help="headline workload only (skip the adjoint and 6-D sections of the default line)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe, "launches": 3}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU sharding is designed but not built in this tree", file=sys.stderr)
        sys.exit(2)

    if args.workload != "knn3d":
        return {"phase6d": run_phase6d, "adjoint": run_adjoint, "large": run_large}[args.workload](args)

    import torch

    import nbodylib_amd as nb
    from nbodylib_amd import datagen
    from nbodylib_amd.device import DeviceTree

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, k, L = args.n, args.k, 100.0
    log(f"generating clustered N={n} seed={args.seed}")
    pos, _ = datagen.clustered(n, L, nhalo=2000, seed=args.seed)
    b = 0.2 * datagen.mean_spacing(n, L)
    pos_d = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
    torch.cuda.synchronize()

    def build(form):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = DeviceTree(pos_d, bucket_size=16, kerntype=nb.KSPH, period=L, form=form)
        torch.cuda.synchronize()
        return t, (time.perf_counter() - t0) * 1e3

    nn = torch.empty((n, k), dtype=torch.int32, device=dev)
    rho = torch.empty(n, dtype=torch.float64, device=dev)
    hh = torch.empty(n, dtype=torch.float64, device=dev)

    def run_form(form, label):
        log(f"{label}: building tree")
        tree, build_ms = build(form)
        tree, build_ms = build(form)  # second build: allocator warm
        log(f"{label}: build {build_ms:.1f} ms; warmup {args.warmup}")
        for _ in range(args.warmup):
            tree.knn_density(k, nn, rho, hh)
        torch.cuda.synchronize()
        nb.prof_reset()
        nb.prof_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tree.knn_density(k, nn, rho, hh)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        nb.prof_enable(False)
        kname = "k_knn_mixed" if form == nb.FORM_MIXED else "k_knn_fp64"
        kms, klaunch = nb.prof_get(kname)
        fbms, _ = nb.prof_get("k_knn_fallback")
        log(f"{label}: {wall / args.steps * 1e3:.2f} ms/step, {kname} {kms / max(klaunch, 1):.2f} ms/launch, "
            f"fallback {fbms / args.steps:.2f} ms/step, fallbacks {tree.last_fallbacks()}")
        return tree, build_ms, wall, kms, klaunch

    forms = args.forms.split(",")
    plain_tree, plain_build_ms, plain_wall, plain_kms, plain_kl = run_form(nb.FORM_PLAIN, "plain-fp64")
    rho_plain = rho.clone()
    headline = dict(form="plain-fp64", wall=plain_wall, kms=plain_kms, kl=plain_kl, build_ms=plain_build_ms)

    mixed = None
    try:
        if "mixed" not in forms:
            raise nb.NbxError(-3, "mixed form not requested")
        mixed_tree, mixed_build_ms, mixed_wall, mixed_kms, mixed_kl = run_form(nb.FORM_MIXED, "mixed")
        relerr = float(torch.max(torch.abs(rho / rho_plain - 1.0)))
        mixed = dict(form="mixed", wall=mixed_wall, kms=mixed_kms, kl=mixed_kl, build_ms=mixed_build_ms,
                     rel=relerr, fallbacks=mixed_tree.last_fallbacks())
        headline = mixed
        del mixed_tree
    except nb.NbxError as e:
        log(f"mixed form unavailable: {e}")

    fof_rate = None
    ngroups = None
    fof_ms = None
    if not args.no_fof:
        lab = torch.empty(n, dtype=torch.int32, device=dev)
        plain_tree.fof(b, 8, lab)  # warm
        torch.cuda.synchronize()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            ngroups = plain_tree.fof(b, 8, lab)
        torch.cuda.synchronize()
        fof_ms = (time.perf_counter() - t0) / reps * 1e3
        fof_rate = n / (fof_ms / 1e3) / 1e6
        log(f"FOF b={b:.5f}: {fof_ms:.1f} ms, {ngroups} groups")

    ms_per_step = headline["wall"] / args.steps * 1e3
    value = n / (ms_per_step / 1e3) / 1e6
    plain_rate = n / (plain_wall / args.steps) / 1e6
    # roofline of the dominant kernel: algorithmic bytes per query (xyz 24 +
    # m 8 + 4k IDs + rho 8 + h 8) over the k_knn launch time (HIP events)
    bytes_per_q = 48 + 4 * k
    kern_ms = headline["kms"] / max(headline["kl"], 1)
    achieved = bytes_per_q * n / (kern_ms / 1e3) / 1e9
    kname = "k_knn_mixed" if headline["form"] == "mixed" else "k_knn_fp64"
    traffic = None
    valu = None
    if os.path.exists(args.traffic):
        tj = json.load(open(args.traffic))
        ent = tj.get(kname)
        if ent and ent.get("n") == n and ent.get("k") == k:
            traffic = ent["bytes_per_launch"]
            # VALU wave-instructions per launch (rocprofv3 SQ_INSTS_VALU on the
            # same seeded workload) over this run's launch time
            vi = ent.get("valu_insts")
            if vi:
                vach = vi / (kern_ms / 1e3) / 1e9
                valu = {"insts_per_launch": vi, "achieved": round(vach, 1), "peak": PEAK_VALU_GINST,
                        "unit": "G VALU wave-instr/s", "frac": round(vach / PEAK_VALU_GINST, 4)}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic,
                "kernel": kname, "kernel_ms": round(kern_ms, 3), "bytes_per_query": bytes_per_q, "valu": valu}

    cpu = None
    if not args.no_cpu:
        import oracle

        stride = max(1, args.cpu_stride)
        log("cpu baseline: building the oracle tree (reference median-split restatement)")
        t0 = time.perf_counter()
        ot = oracle.Tree(pos, b=16, kerntype=oracle.KSPH, period=L)
        cpu_build = time.perf_counter() - t0
        slots = np.arange(0, n, stride, dtype=np.int64)
        nth = oracle.num_threads()
        log(f"cpu baseline: kNN+density on {len(slots)} sampled queries, {nth} threads")
        t0 = time.perf_counter()
        rho_cpu = ot.knn_density_sample(k, slots, nth)
        cpu_t = time.perf_counter() - t0
        model, ncpu, naff = cpu_info()
        cpu = {"value": round(len(slots) / cpu_t / 1e6, 4), "unit": "Mqueries/s", "cores": nth, "kind": "port",
               "sample": f"kNN k={k} + gather SPH density for every {stride}-th tree slot "
                         f"({len(slots)} of {n} queries) of the same clustered set, OpenMP schedule(guided); "
                         f"oracle tree build {cpu_build:.1f} s not included",
               "cpu_model": model, "host_cpus": ncpu, "cpus_available": naff, "oracle_flags": "-O3 -fopenmp"}
        # the same sample against the GPU's gather density (CalcDensityParticle for every particle)
        ids = ot.order()[slots]
        rho_g = torch.empty(n, dtype=torch.float64, device=dev)
        plain_tree.density(k, rho_g, variant=nb.DENS_GATHER)
        rg = rho_g.cpu().numpy()[ids]
        cpu["sample_rho_max_rel_vs_gpu_gather"] = float(np.max(np.abs(rg / rho_cpu - 1.0)))
        # serial as shipped: the library's CalcDensity has no OpenMP
        # (KDCalcSmoothQuantities.cxx:203-305); the same per-query work on one core
        sslots = np.arange(0, n, 64, dtype=np.int64)
        log(f"cpu baseline: serial kNN+density on {len(sslots)} sampled queries, 1 thread")
        t0 = time.perf_counter()
        ot.knn_density_sample(k, sslots, 1)
        ser_t = time.perf_counter() - t0
        cpu["serial_as_shipped"] = {"value": round(len(sslots) / ser_t / 1e6, 4), "unit": "Mqueries/s", "cores": 1,
                                    "sample": f"every 64-th tree slot ({len(sslots)} queries), one thread"}
        log(f"cpu baseline: {cpu['value']} Mq/s on {nth} threads, serial {cpu['serial_as_shipped']['value']} Mq/s, "
            f"sample rho vs GPU gather max rel {cpu['sample_rho_max_rel_vs_gpu_gather']:.2e}")
        del ot

    out = {
        "metric": "Mqueries/s kNN+SPH density (k=32, N=16.8M) mixed vs FP64 GPU; FOF Mparticles/s",
        "value": round(value, 3),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if headline["form"] == "mixed" else "f64",
        "data": "synthetic (seeded NFW-clustered particles generated on the host, resident in HBM)",
        "config": {"workload": f"clustered NFW 3-D periodic box L=100, N={n}, k={k} kNN + cubic-spline SPH density "
                               f"and smoothing length; FOF b=0.2 mean spacing", "form": headline["form"],
                   "n": n, "k": k, "fof_b": b, "parallelism": "single GPU"},
        "plain_fp64_mqps": round(plain_rate, 3),
        "mixed_mqps": round(value, 3) if mixed else None,
        "mixed_vs_plain": round(value / plain_rate, 3) if mixed else None,
        "mixed_density_max_rel_vs_fp64": mixed["rel"] if mixed else None,
        "mixed_fp64_fallback_queries": mixed["fallbacks"] if mixed else None,
        "build_ms": round(headline["build_ms"], 2),
        "fof_mparticles_s": round(fof_rate, 3) if fof_rate else None,
        "fof_ms": round(fof_ms, 2) if fof_ms else None,
        "fof_groups": ngroups,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    if not args.no_extras:
        del plain_tree
        torch.cuda.empty_cache()
        out["c3_adjoint"] = adjoint_section(args, steps=max(args.steps, 3))
        torch.cuda.empty_cache()
        out["c2_phase6d"] = phase6d_section(args, steps=2, warmup=1)
    print(json.dumps(out), flush=True)


def cpu_info():
    """(model name, logical CPUs of the host, CPUs this process may use)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        naff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        naff = None
    return model, os.cpu_count(), naff


class Timer:
    """HIP events on the current stream (the stream the library launches on)."""

    def __init__(self, torch):
        self.torch = torch
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        self.a.record()
        return self

    def __exit__(self, *exc):
        self.b.record()
        self.b.synchronize()
        self.ms = self.a.elapsed_time(self.b)


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def _line(metric, value, unit, args, ms, dtype, workload, extra, higher=True):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": higher, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded NFW-clustered particles generated on the host, resident in HBM)",
           "config": dict(workload=workload, parallelism="single GPU")}
    out.update(extra)
    print(json.dumps(out), flush=True)


def phase6d_section(args, steps, warmup):
    """6-D phase-space workload (N=16.8M, k=64): precision and plain forms,
    3-D FOF and FOFCriterion(FOF6d) on the same particles."""
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd import datagen
    from nbodylib_amd.device import DeviceTree

    dev = torch.device("cuda:0")
    n = args.n
    k = 64 if args.k == 32 else args.k
    L = 100.0
    log(f"phase6d: generating clustered N={n} with velocities")
    pos, vel, _ = datagen.clustered(n, L, nhalo=2000, seed=args.seed + 2, with_vel=True)
    sx, sv = pos.std(axis=0), vel.std(axis=0)
    xs = torch.from_numpy(np.ascontiguousarray((pos / sx).T)).to(dev)
    vs = torch.from_numpy(np.ascontiguousarray((vel / sv).T)).to(dev)
    del vel
    rho = torch.empty(n, dtype=torch.float64, device=dev)
    hh = torch.empty(n, dtype=torch.float64, device=dev)

    def timed(form, label):
        builds = []
        for _ in range(2):
            torch.cuda.synchronize()
            with Timer(torch) as tb:
                t6 = DeviceTree(xs, vel=vs, bucket_size=16, treetype=nb.TPHS, kerntype=nb.KSPH, aniso=-1, form=form)
            builds.append(tb.ms)
        log(f"phase6d {label}: 6-D tree build {builds[-1]:.1f} ms; warmup {warmup}")
        for _ in range(warmup):
            t6.density(k, rho, hh)
        nb.prof_reset()
        nb.prof_enable(True)
        times = []
        for _ in range(steps):
            with Timer(torch) as tm:
                t6.density(k, rho, hh)
            times.append(tm.ms)
        nb.prof_enable(False)
        kname = "k_knn_mixed" if form == nb.FORM_MIXED else "k_knn_fp64"
        kms, kl = nb.prof_get(kname)
        ms = _median(times)
        fb = t6.last_fallbacks()
        log(f"phase6d {label}: 6-D k={k} kNN + density {ms:.2f} ms ({kname} {kms / max(kl, 1):.2f} ms/launch, "
            f"fallbacks {fb})")
        del t6
        return ms, kms / max(kl, 1), builds[-1], kname, fb

    plain_ms, _, _, _, _ = timed(nb.FORM_PLAIN, "plain-fp64")
    rho_plain = rho.clone()
    ms, kern_ms, build_ms, kname, fb = timed(nb.FORM_MIXED, "mixed")
    rel = float(torch.max(torch.abs(rho / rho_plain - 1.0)))
    del rho_plain
    # 3-D FOF at b = 0.2 mean spacing on the same particles
    pd = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
    t3 = DeviceTree(pd, bucket_size=16, kerntype=nb.KSPH, period=L)
    b = 0.2 * datagen.mean_spacing(n, L)
    lab = torch.empty(n, dtype=torch.int32, device=dev)
    t3.fof(b, 8, lab)
    fts = []
    for _ in range(3):
        with Timer(torch) as tf:
            ng = t3.fof(b, 8, lab)
        fts.append(tf.ms)
    fms = _median(fts)
    log(f"phase6d: 3-D FOF b={b:.5f}: {fms:.1f} ms, {ng} groups")
    del t3, pd
    # FOFCriterion with FOF6d on the scaled phase-space tree: 0.2 mean spacing
    # in the scaled positions, 0.1 in the unit-variance velocities
    t6 = DeviceTree(xs, vel=vs, bucket_size=16, treetype=nb.TPHS, kerntype=nb.KSPH, aniso=-1)
    s6 = (0.2 * datagen.mean_spacing(n, L) / float(np.mean(sx))) ** 2
    prm6 = [nb.TPHS, s6, 0.01, 0, 0, 0, s6, 0.01]
    t6.fof_criterion(nb.FOF_CMP_6D, prm6, 8, lab)
    f6 = []
    for _ in range(3):
        with Timer(torch) as tf:
            ng6 = t6.fof_criterion(nb.FOF_CMP_6D, prm6, 8, lab)
        f6.append(tf.ms)
    f6ms = _median(f6)
    log(f"phase6d: FOF6d (FOFCriterion) params[6]={s6:.3e} params[7]=0.01: {f6ms:.1f} ms, {ng6} groups")
    del t6
    bpq = 48 + 8 + 4 * k + 16
    ach = bpq * n / (kern_ms / 1e3) / 1e9
    return {"workload": f"6-D phase space (x/sigma_x, v/sigma_v), clustered NFW N={n}, k={k} CalcDensity on a "
                        f"TPHS tree; 3-D FOF b=0.2 mean spacing; steps={steps}",
            "n": n, "k": k, "mixed_mqps": round(n / (ms / 1e3) / 1e6, 3), "mixed_ms": round(ms, 2),
            "plain_fp64_mqps": round(n / (plain_ms / 1e3) / 1e6, 3), "plain_fp64_ms": round(plain_ms, 2),
            "mixed_vs_plain": round(plain_ms / ms, 3), "mixed_density_max_rel_vs_fp64": rel,
            "mixed_fp64_fallback_queries": fb, "build_ms": round(build_ms, 2), "fof_ms": round(fms, 2),
            "fof_mparticles_s": round(n / fms / 1e3, 3), "fof_groups": ng, "fof6d_criterion_ms": round(f6ms, 2),
            "fof6d_criterion_groups": ng6,
            "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(ach / PEAK_HBM_GBS, 5), "kernel": kname, "kernel_ms": round(kern_ms, 3),
                         "bytes_per_query": bpq}}


def run_phase6d(args):
    r = phase6d_section(args, args.steps, args.warmup)
    _line(f"Mqueries/s 6-D kNN (k={r['k']}) + 6-D SPH density, N={r['n']}, precision form", r["mixed_mqps"],
          "Mqueries/s", args, r["mixed_ms"], "f32+f64", r["workload"], dict(r, cpu_baseline=None))


def _adjoint_pass(torch, nb, pos, mass, nbr, reps):
    """Times forward (density over the list) and backward (adjoint) with
    g = d(sum log rho)/d rho = 1/rho."""
    from nbodylib_amd.device import sph_density, _ptr, _stream, _check

    n, k = nbr.shape
    rho, _ = sph_density(pos, mass, nbr)
    g = (1.0 / rho).contiguous()
    dpos = torch.empty_like(pos)
    dmass = torch.empty_like(mass)
    fw, bw = [], []
    for _ in range(reps):
        with Timer(torch) as tf:
            rho, _ = sph_density(pos, mass, nbr)
        fw.append(tf.ms)
        with Timer(torch) as tb:  # the entry point zeroes dpos/dmass itself
            _check(nb.lib().nbx_sph_density_adjoint_dev(_ptr(pos), _ptr(mass), _ptr(nbr), n, k, _ptr(g), _ptr(dpos),
                                                        _ptr(dmass), _stream()))
        bw.append(tb.ms)
    return _median(fw), _median(bw)


def adjoint_section(args, steps):
    """Differentiation form at N=1M (clustered, lognormal masses, k=32):
    forward / backward over a stored list in input (ID) and tree order, and
    with the list recomputed per block of tree slots."""
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd import datagen
    from nbodylib_amd.device import DeviceTree

    dev = torch.device("cuda:0")
    n = 1 << 20 if args.n == 256 ** 3 else args.n
    k = args.k
    L = 100.0
    log(f"adjoint: generating clustered N={n}, lognormal masses")
    pos, _ = datagen.clustered(n, L, nhalo=2000, seed=args.seed + 3)
    m = datagen.lognormal_masses(n, seed=args.seed + 3)
    pd = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
    md = torch.from_numpy(m).to(dev)
    tree = DeviceTree(pd, mass=md, bucket_size=16, kerntype=nb.KSPH, form=nb.FORM_MIXED)
    # frozen list in input (ID) order: rows and entries are IDs, ascending
    nn_id = torch.empty((n, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((n, k), dtype=torch.float64, device=dev)
    tree.knn(k, nn_id, d2, variant=nb.KNN_NONPERIODIC, layout=nb.LAYOUT_BY_ID)
    del d2
    with Timer(torch) as tk:
        tree.knn(k, torch.empty_like(nn_id), variant=nb.KNN_NONPERIODIC, layout=nb.LAYOUT_BY_ID)
    knn_ms = tk.ms
    fw_id, bw_id = _adjoint_pass(torch, nb, pd, md, nn_id, steps)
    # the same list in tree order: positions/masses permuted by the tree
    order = tree.order().long()
    inv = torch.empty_like(order)
    inv[order] = torch.arange(n, device=dev)
    pt = pd[:, order].contiguous()
    mt = md[order].contiguous()
    nn_t = inv[nn_id[order].long()].to(torch.int32).contiguous()
    fw_t, bw_t = _adjoint_pass(torch, nb, pt, mt, nn_t, steps)
    del pt, mt, nn_t
    # the list recomputed per block of tree slots (no list of all n kept)
    block = 1 << 18
    rho = torch.empty(n, dtype=torch.float64, device=dev)
    hh = torch.empty(n, dtype=torch.float64, device=dev)
    tree.sph_density(k, rho, hh, block_rows=block)
    g = (1.0 / rho).contiguous()
    dp = torch.empty_like(pd)
    dm = torch.empty_like(md)
    rf, rb = [], []
    for _ in range(steps):
        with Timer(torch) as tf:
            tree.sph_density(k, rho, hh, block_rows=block)
        rf.append(tf.ms)
        with Timer(torch) as tb:
            tree.sph_density_adjoint(k, g, dp, dm, block_rows=block)
        rb.append(tb.ms)
    fw_r, bw_r = _median(rf), _median(rb)
    log(f"adjoint: ID order fwd {fw_id:.2f} / bwd {bw_id:.2f} ms; tree order fwd {fw_t:.2f} / bwd {bw_t:.2f} ms; "
        f"recomputed per {block}-slot block fwd {fw_r:.2f} / bwd {bw_r:.2f} ms; kNN for the list {knn_ms:.2f} ms")
    bpp = 24 + 8 + 8 + 8 + 4 * k + 32
    ach = bpp * n / (bw_t / 1e3) / 1e9
    return {"workload": f"clustered NFW N={n}, lognormal masses, frozen k={k} list; backward = hand-written "
                        f"adjoint of sum log rho", "n": n, "k": k,
            "fwd_ms_tree_order": round(fw_t, 3), "bwd_ms_tree_order": round(bw_t, 3),
            "bwd_over_fwd_tree_order": round(bw_t / fw_t, 3),
            "fwd_ms_id_order": round(fw_id, 3), "bwd_ms_id_order": round(bw_id, 3),
            "bwd_over_fwd_id_order": round(bw_id / fw_id, 3),
            "recompute_block_rows": block, "fwd_ms_recompute": round(fw_r, 3), "bwd_ms_recompute": round(bw_r, 3),
            "bwd_over_fwd_recompute": round(bw_r / fw_r, 3),
            "list_bytes_stored": n * k * 4, "list_bytes_recompute": block * k * 4,
            "list_recompute_ms": round(knn_ms, 3),
            "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(ach / PEAK_HBM_GBS, 5), "kernel": "k_sph_adj", "kernel_ms": round(bw_t, 3),
                         "bytes_per_particle": bpp}}


def run_adjoint(args):
    r = adjoint_section(args, max(args.steps, 3))
    _line(f"Mparticles/s SPH density adjoint over a frozen k={r['k']} list, N={r['n']}",
          r["n"] / (r["bwd_ms_tree_order"] / 1e3) / 1e6, "Mparticles/s", args, r["bwd_ms_tree_order"], "f64",
          r["workload"], dict(r, cpu_baseline=None))


def run_large(args):
    import torch

    import nbodylib_amd as nb
    from nbodylib_amd import datagen
    from nbodylib_amd.device import DeviceTree

    dev = torch.device("cuda:0")
    n = 512 ** 3 if args.n == 256 ** 3 else args.n
    k = args.k
    L = 100.0
    log(f"generating clustered N={n} (host)")
    pos, _ = datagen.clustered(n, L, nhalo=2000, seed=args.seed + 4)
    m = datagen.lognormal_masses(n, seed=args.seed + 4)
    pd = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)
    del pos
    md = torch.from_numpy(m).to(dev)
    del m
    log("inputs resident in HBM")
    torch.cuda.synchronize()
    with Timer(torch) as tb:
        tree = DeviceTree(pd, mass=md, bucket_size=16, kerntype=nb.KSPH, period=L, form=nb.FORM_MIXED)
    log(f"build {tb.ms:.1f} ms")
    nn = torch.empty((n, k), dtype=torch.int32, device=dev)
    rho = torch.empty(n, dtype=torch.float64, device=dev)
    hh = torch.empty(n, dtype=torch.float64, device=dev)
    tree.knn_density(k, nn, rho, hh)
    times = []
    for _ in range(max(args.steps, 1)):
        with Timer(torch) as tk:
            tree.knn_density(k, nn, rho, hh)
        times.append(tk.ms)
        log(f"kNN+density {tk.ms:.1f} ms")
    kms = _median(times)
    fb = tree.last_fallbacks()
    # adjoint over the stored list, in tree order (the list is by slot)
    order = tree.order().long()
    pt = pd[:, order].contiguous()
    mt = md[order].contiguous()
    del order
    fw, bw = _adjoint_pass(torch, nb, pt, mt, nn, 2)
    log(f"adjoint over the stored list: fwd {fw:.1f} ms bwd {bw:.1f} ms")
    del pt, mt, nn
    torch.cuda.empty_cache()
    # adjoint with the list recomputed per block of tree slots: no list of all n
    block = 1 << 22
    g = (1.0 / rho).contiguous()
    dp = torch.empty_like(pd)
    dm = torch.empty_like(md)
    with Timer(torch) as trf:
        tree.sph_density(k, rho, hh, block_rows=block)
    with Timer(torch) as trb:
        tree.sph_density_adjoint(k, g, dp, dm, block_rows=block)
    log(f"adjoint with the list recomputed per {block}-slot block: fwd {trf.ms:.1f} ms bwd {trb.ms:.1f} ms")
    del g, dp, dm
    b = 0.2 * datagen.mean_spacing(n, L)
    lab = torch.empty(n, dtype=torch.int32, device=dev)
    with Timer(torch) as tf:
        ng = tree.fof(b, 8, lab)
    log(f"FOF {tf.ms:.1f} ms, {ng} groups")
    mem = torch.cuda.max_memory_allocated() / 2 ** 30
    _line(f"Mqueries/s kNN+SPH density (k={k}, N={n}) mixed", n / (kms / 1e3) / 1e6, "Mqueries/s", args, kms,
          "f32+f64", f"clustered NFW N={n}, lognormal masses, periodic L=100: build, mixed kNN+density k={k}, "
                     f"adjoint with the list recomputed per block (and over the stored list), FOF b=0.2 mean "
                     f"spacing", {
              "build_ms": round(tb.ms, 2), "mixed_fp64_fallback_queries": fb,
              "adjoint_stored_fwd_ms": round(fw, 2), "adjoint_stored_bwd_ms": round(bw, 2),
              "adjoint_recompute_block_rows": block, "adjoint_recompute_fwd_ms": round(trf.ms, 2),
              "adjoint_recompute_bwd_ms": round(trb.ms, 2),
              "list_bytes_stored": n * k * 4, "list_bytes_recompute": block * k * 4,
              "fof_ms": round(tf.ms, 2), "fof_groups": ng, "torch_peak_gib_after_list_freed": round(mem, 2),
              "cpu_baseline": None})


if __name__ == "__main__":
    main()
