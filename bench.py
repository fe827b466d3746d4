#!/usr/bin/env python
"""Headline benchmark: accepted spheres/s to maximal packing on one MI355X.

A step is one proceed(seed) of the reference's voxel method (method 1) from an
empty periodic box to a maximal packing: 3D monodisperse r=0.005 in the unit
cube, seed 0, nbshots_singledraw 6000 (~734k spheres).  Inputs are synthetic
and generated on the device from the seed; the timed region is proceed() only
(domain and buffers are set up before it).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "accepted spheres/sec to maximal packing (3D, 1 MI355X) learned vs plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(dim=3, L=1.0, r=0.005, seed=0, nbshots=6000)


def run_gpu_step(R, profile_level=1):
    w = WORKLOAD
    dom = R.Domain([0.0] * w["dim"], [w["L"]] * w["dim"], w["r"])
    gen = R.RadiusGenerator([(w["r"], 1.0, 0)], dom.get_total_volume(), 0.0)
    algo = R.Algo(dom, gen, w["nbshots"])
    if profile_level:
        dom.profile(profile_level)
    t0 = time.perf_counter()
    st = algo.proceed(w["seed"])
    dt = time.perf_counter() - t0
    return dom, st, dt


def profile_figure(key):
    """A `key: value` line of profiles/k_cover_3d_r0.005.md (rocprofv3 PMC summary), or None."""
    path = os.path.join(ROOT, "profiles", "k_cover_3d_r0.005.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def measured_traffic():
    """HBM bytes per k_cover launch on this workload, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes summarised in profiles/ (each counter scaled by what it reads for
    a plain copy of known size); None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "k_cover_3d_r0.005.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(side=0.8):
    """The CPU oracle (restatement of the reference's single-rank path) on all
    host cores, on a bounded sample of the same workload: the same radius and
    seed in a periodic cube of side `side`, run to maximal packing."""
    import _oracle as O
    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O.build()
    t0 = time.perf_counter()
    o = O.run(3, [0.0] * 3, [side] * 3, [(WORKLOAD["r"], 1.0, 0)], seed=WORKLOAD["seed"])
    dt = time.perf_counter() - t0
    return dict(value=o["n"] / dt, unit="spheres/s", cores=threads, kind="port",
                sample=f"3D r={WORKLOAD['r']} seed {WORKLOAD['seed']} in the periodic cube of side {side} "
                       f"({side ** 3:.0%} of the workload volume) to maximal packing: {o['n']} spheres in {dt:.2f} s")


def batched(R, workers):
    """The batched workload: 256 independent 3D realizations (box edge L ~ U[0.2, 1],
    r ~ U[0.004, 0.008], seeds 0..255, each to maximal packing), timed once with
    `workers` realizations at a time and once one after another (the plain form)."""
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "batch_3d_256.json")))
    real = fx["realizations"]
    L = [c["L"] for c in real]
    r = [c["r"] for c in real]
    s = [c["seed"] for c in real]
    R.batch_run(3, L[:8], r[:8], s[:8], workers=workers)  # warm-up: contexts, pools, threads
    want = np.array([c["n"] for c in fx["oracle"]], dtype=np.uint64)
    res = {}
    for name, w in (("batched", workers), ("plain", 1)):
        t0 = time.perf_counter()
        out = R.batch_run(3, L, r, s, workers=w)
        dt = time.perf_counter() - t0
        res[name] = dict(seconds=round(dt, 3), value=round(float(out["counts"].sum()) / dt, 1),
                         counts_match_oracle=bool(np.array_equal(out["counts"], want)))
    return {"workload": "256 independent 3D monodisperse realizations, periodic box [0,L]^3, L~U[0.2,1], "
                        "r~U[0.004,0.008], seeds 0..255, each to maximal packing",
            "realizations": len(real), "spheres": int(want.sum()), "workers": workers, "unit": "spheres/s",
            "value": res["batched"]["value"], "plain_value": res["plain"]["value"],
            "speedup_vs_plain": round(res["batched"]["value"] / res["plain"]["value"], 3), **{
                k + "_" + kk: v for k, d in res.items() for kk, v in d.items() if kk != "value"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-side", type=float, default=0.8)
    ap.add_argument("--batch-workers", type=int, default=6,
                    help="realizations run at once in the batched measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched measurement")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (reference reads for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: the library runs one rank's whole domain on a single MI355X, "
                 "and sharding a realization over several GPUs is designed but not implemented.")

    import rsa_mpi_amd as R
    R.lib()
    for _ in range(args.warmup):
        run_gpu_step(R, profile_level=0)

    total_n, total_t = 0, 0.0
    prof = {}
    stats = None
    for _ in range(args.steps):
        dom, st, dt = run_gpu_step(R, profile_level=1)
        total_n += st["accepted"]
        total_t += dt
        stats = st
        for k, v in dom.kernel_profile().items():
            a = prof.setdefault(k, dict(launches=0, seconds=0.0, bytes=0.0))
            for f in a:
                a[f] += v[f]
    value = total_n / total_t

    # one extra, untimed step with every per-draw kernel bracketed by events
    dom2, _, dt2 = run_gpu_step(R, profile_level=2)
    detail = {kk: dict(launches=v["launches"], ms=round(1e3 * v["seconds"], 3),
                       us_per_launch=round(1e6 * v["seconds"] / max(1, v["launches"]), 2))
              for kk, v in dom2.kernel_profile().items()}
    del dom2

    parity = None
    gold = os.path.join(ROOT, "tests", "golden", "3d_r0.005_seed0.json")
    if os.path.exists(gold):
        from _common import set_hash
        g = json.load(open(gold))
        pos, rad, _ = dom.extract_arrays()
        parity = "sha256 match" if set_hash(pos, rad) == g["sha256"] and len(rad) == g["n"] else "MISMATCH"

    k = prof.get("k_cover")
    roofline = None
    if k and k["seconds"] > 0:
        achieved = k["bytes"] / k["seconds"] / 1e9
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(),
                        kernel="k_cover", launches=k["launches"] // args.steps,
                        avg_launch_ms=round(1e3 * k["seconds"] / k["launches"], 3),
                        algorithmic_bytes_per_launch=round(k["bytes"] / k["launches"]))
        # k_cover issues FP64 VALU work: its VALU instruction count (SQ_INSTS_VALU, from the
        # committed PMC summary) at 4 cycles per wave64 instruction on 1024 SIMDs at 2.4 GHz,
        # over the measured kernel time
        valu = profile_figure("valu_wave_instructions_per_run")
        if valu:
            busy_s = valu * 4 / (1024 * 2.4e9)
            roofline["valu"] = dict(wave_instructions_per_run=valu, issue_seconds_per_run=round(busy_s, 6),
                                    busy_frac=round(busy_s / (k["seconds"] / args.steps), 3))

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "spheres/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * total_t / args.steps, 2), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D monodisperse RSA, r=0.005, periodic unit cube, seed 0, voxel method to "
                               "maximal packing (plain GPU form)", "spheres": stats["accepted"],
                   "draws": stats["draws"], "refines": stats["refines"], "nvox_max": stats["nvox_max"],
                   "mt_words": stats["words"], "parity_vs_oracle": parity},
        "roofline": roofline,
        "phases_ms": {"voxel_refine": round(1e3 * stats["voxel_seconds"], 2),
                      "draws": round(1e3 * (stats["seconds"] - stats["voxel_seconds"]), 2),
                      "rounds": stats["rounds"], "rounds_max": stats["rounds_max"],
                      # the dart phase is latency-bound: its rate per draw
                      "us_per_draw": round(1e6 * (stats["seconds"] - stats["voxel_seconds"]) / stats["draws"], 2),
                      "draws_per_s": round(stats["draws"] / (stats["seconds"] - stats["voxel_seconds"]), 1)},
        # per kernel (event-timed); GB/s = algorithmic bytes over kernel time where bytes are counted
        "profile": {kk: dict(launches=v["launches"] // args.steps, ms=round(1e3 * v["seconds"] / args.steps, 3),
                             **({"GBps": round(v["bytes"] / v["seconds"] / 1e9, 1)} if v["bytes"] > 0 else {}))
                    for kk, v in prof.items()},
        "kernels_all_bracketed": dict(step_ms=round(1e3 * dt2, 1), **detail),
        "batched": None if args.no_batch else batched(R, args.batch_workers),
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(args.cpu_side),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
