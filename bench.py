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
    if os.path.exists(args.traffic):
        tj = json.load(open(args.traffic))
        ent = tj.get(kname)
        if ent and ent.get("n") == n and ent.get("k") == k:
            traffic = ent["bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic,
                "kernel": kname, "kernel_ms": round(kern_ms, 3), "bytes_per_query": bytes_per_q}

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
        ot.knn_density_sample(k, slots)
        cpu_t = time.perf_counter() - t0
        cpu = {"value": round(len(slots) / cpu_t / 1e6, 4), "unit": "Mqueries/s", "cores": nth, "kind": "port",
               "sample": f"kNN k={k} + gather SPH density for every {stride}-th tree slot "
                         f"({len(slots)} of {n} queries) of the same clustered set, OpenMP schedule(guided); "
                         f"oracle tree build {cpu_build:.1f} s not included"}
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
