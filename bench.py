#!/usr/bin/env python
"""Headline benchmark: accepted spheres/s to maximal packing on one MI355X.

A step is one proceed(seed) of the reference's voxel method (method 1) from an
empty periodic box to a maximal packing: 3D monodisperse r=0.002 in the unit
cube, seed 0, nbshots_singledraw 6000 (11 463 041 spheres, 772 M voxels at the
peak).  Inputs are synthetic and generated from the seed; the timed region is
proceed() only (domain and buffers are set up before it).  The placed set of
the last timed step is checked by SHA-256 against the oracle's fixture.  A
secondary line runs the r=0.005 cube the same way, and the learned form (an
MLP trained on the GPU orders the work; the placed set is unchanged) runs the
headline beside the plain one.  Prints one JSON line.
"""
import argparse
import hashlib
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

HEADLINE = dict(name="3D monodisperse RSA, r=0.002, periodic unit cube, seed 0, voxel method to maximal packing",
                dim=3, L=1.0, r=0.002, seed=0, nbshots=6000, golden="3d_r0.002_seed0.json",
                profile="k_cover_3d_r0.002.md")
SECONDARY = dict(name="3D monodisperse RSA, r=0.005, periodic unit cube, seed 0, voxel method to maximal packing",
                 dim=3, L=1.0, r=0.005, seed=0, nbshots=6000, golden="3d_r0.005_seed0.json")


def source_sha16():
    """SHA-256 (first 16 hex digits) of the HIP library's sources: rocprofv3 figures in
    profiles/ are used only when they were measured on this exact source."""
    h = hashlib.sha256()
    d = os.path.join(ROOT, "rsa_mpi_amd", "csrc")
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".hpp", ".cpp", ".h", ".inl")):
            h.update(f.encode())
            h.update(open(os.path.join(d, f), "rb").read())
    return h.hexdigest()[:16]


def profile_figures(name):
    """`key: value` lines of a rocprofv3 summary in profiles/, or {} when the summary is
    missing or was measured on other sources than the ones built here."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith(("|", "#", " ", "`")):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("source_sha16") != source_sha16():
        return {}
    out = {}
    for k, v in vals.items():
        try:
            out[k] = float(v)
        except ValueError:
            pass
    return out


def run_gpu_step(R, w, profile_level=1, learn=None):
    dom = R.Domain([0.0] * w["dim"], [w["L"]] * w["dim"], w["r"])
    gen = R.RadiusGenerator([(w["r"], 1.0, 0)], dom.get_total_volume(), 0.0)
    algo = R.Algo(dom, gen, w["nbshots"])
    if profile_level:
        dom.profile(profile_level)
    if learn:
        dom.learn_mode(**learn)
    t0 = time.perf_counter()
    st = algo.proceed(w["seed"])
    dt = time.perf_counter() - t0
    return dom, st, dt


def parity(dom, w):
    from _common import set_hash
    g = json.load(open(os.path.join(ROOT, "tests", "golden", w["golden"])))
    pos, rad, _ = dom.extract_arrays()
    ok = len(rad) == g["n"] and set_hash(pos, rad) == g["sha256"]
    return "sha256 match" if ok else "MISMATCH"


def measure(R, w, steps, warmup, learn=None):
    """warmup untimed steps, then `steps` timed ones; per-kernel event profile (refine
    kernels) of the timed steps, the parity of the last one and, in a learned mode, its
    predictor counters."""
    for _ in range(warmup):
        run_gpu_step(R, w, profile_level=0, learn=learn)
    total_n, total_t, prof, stats, dom = 0, 0.0, {}, None, None
    for s in range(steps):
        dom, st, dt = run_gpu_step(R, w, profile_level=1, learn=learn)
        print(f"#   step {s}: {st['accepted']} spheres in {dt:.3f} s", file=sys.stderr, flush=True)
        total_n += st["accepted"]
        total_t += dt
        stats = st
        for k, v in dom.kernel_profile().items():
            a = prof.setdefault(k, dict(launches=0, seconds=0.0, bytes=0.0))
            for f in a:
                a[f] += v[f]
    par = parity(dom, w)
    ls = dom.learn_stats() if learn else None
    del dom
    return dict(value=total_n / total_t, ms=1e3 * total_t / steps, prof=prof, stats=stats, parity=par, learn=ls)


# FLOP of one sample through the 16 -> 64 -> 64 -> 2 MLP (forward); forward + backward is 3x
MLP_FLOP = 2 * (16 * 64 + 64 * 64 + 64 * 2)
BF16_PEAK_TFLOPS = 2500.0
TRACE_RADII = (0.01, 0.007, 0.005)


def learned(R, steps, plain):
    """The learned form on the headline workload: the model is trained on the GPU from the
    traces of three smaller runs (r = 0.01, 0.007, 0.005 in the unit cube), then the headline
    runs with the model ordering the coverage walk and the draw queue.  Reported beside the
    plain form: spheres/s, the parity of the placed set, the misprediction rates (and the
    constant predictor's), the uniform-wave fractions, k_predict's matrix-core rate."""
    R.learn_reset(5)
    t0 = time.perf_counter()
    for r in TRACE_RADII:
        run_gpu_step(R, dict(HEADLINE, r=r, golden=None), profile_level=0, learn=dict(trace=True))
    t_trace = time.perf_counter() - t0
    nsamp = R.learn_samples()
    train_steps, batch = 600, 2048
    t0 = time.perf_counter()
    first, last, _ = R.learn_train(steps=train_steps, batch=batch, lr=3e-3, seed=1)
    t_train = time.perf_counter() - t0
    print(f"#   trained on {nsamp} samples: loss {first:.4f} -> {last:.4f} in {t_train:.2f} s",
          file=sys.stderr, flush=True)
    base = measure(R, HEADLINE, 1, 0, learn=dict(measure=True))["learn"]
    m = measure(R, HEADLINE, steps, 0, learn=dict(use=True))
    ls, prof = m["learn"], m["prof"]
    kp = prof.get("k_predict", {})
    vp = ls["voxels_predicted"]
    out = {"value": round(m["value"], 1), "unit": "spheres/s", "ms_per_step": round(m["ms"], 2),
           "vs_plain": round(m["value"] / plain, 3) if plain else None, "parity_vs_oracle": m["parity"],
           "trained_on": f"{nsamp} samples traced from r={list(TRACE_RADII)} ({t_trace:.2f} s)",
           "train": dict(steps=train_steps, batch=batch, seconds=round(t_train, 3), loss_first=round(first, 4),
                         loss_last=round(last, 4),
                         tflops=round(3 * MLP_FLOP * train_steps * batch / t_train / 1e12, 3)),
           "voxel_misprediction": round(ls["voxels_mispredicted"] / max(1, vp), 4),
           "constant_predictor_misprediction": round((vp - ls["voxels_covered_predicted"]) / max(1, vp), 4),
           "draw_misprediction": round(ls["draws_mispredicted"] / max(1, ls["draws_predicted"]), 4),
           "refines": ls["refines"], "refines_predicted": ls["refines_predicted"],
           "speculation_skipped": ls["speculation_skipped"],
           "uniform_waves_plain": round(base["waves_uniform"] / max(1, base["waves"]), 4),
           "uniform_waves_learned": round(ls["waves_uniform"] / max(1, ls["waves"]), 4),
           "profile": per_kernel(prof, steps)}
    mlp = profile_figures("mlp_kernels_3d_r0.002.md")
    if mlp:
        out["mfma_profile"] = {k: v for k, v in mlp.items() if k.endswith(("_mfma_busy_frac", "_ms"))}
    if kp.get("seconds"):
        per_step = kp["seconds"] / steps  # every step predicts the same voxels
        tf = vp * MLP_FLOP / per_step / 1e12
        out["k_predict"] = dict(ms=round(1e3 * per_step, 3), voxels=vp, flop_per_voxel=MLP_FLOP,
                                tflops=round(tf, 2), mfma_frac=round(tf / BF16_PEAK_TFLOPS, 4))
    return out


def phases(stats):
    draw_s = stats["seconds"] - stats["voxel_seconds"]
    return {"voxel_refine": round(1e3 * stats["voxel_seconds"], 2), "draws": round(1e3 * draw_s, 2),
            "rounds": stats["rounds"], "rounds_max": stats["rounds_max"],
            # the dart phase is latency-bound: its rate per draw
            "us_per_draw": round(1e6 * draw_s / stats["draws"], 2),
            "draws_per_s": round(stats["draws"] / draw_s, 1)}


def per_kernel(prof, steps):
    """per kernel (event-timed on the engine's stream); GB/s = algorithmic bytes over kernel time"""
    return {kk: dict(launches=v["launches"] // steps, ms=round(1e3 * v["seconds"] / steps, 3),
                     **({"GBps": round(v["bytes"] / v["seconds"] / 1e9, 1)} if v["bytes"] > 0 else {}))
            for kk, v in prof.items()}


def roofline(prof, steps, figures):
    """k_cover (the dominant refine kernel): algorithmic bytes per launch over its
    event-timed average launch duration, against the 8 TB/s HBM peak; traffic and the
    FP64 VALU issue fraction from the rocprofv3 summary of the same sources."""
    k = prof.get("k_cover")
    if not k or k["seconds"] <= 0:
        return None
    achieved = k["bytes"] / k["seconds"] / 1e9
    out = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(achieved / HBM_PEAK_GBS, 4), traffic=figures.get("traffic_bytes_per_launch"),
               kernel="k_cover", launches=k["launches"] // steps,
               avg_launch_ms=round(1e3 * k["seconds"] / k["launches"], 3),
               algorithmic_bytes_per_launch=round(k["bytes"] / k["launches"]))
    # FP64 VALU issue: SQ_INSTS_VALU wave-instructions at 4 cycles per wave64 instruction on
    # 1024 SIMDs at 2.4 GHz, over the measured kernel time
    valu = figures.get("valu_wave_instructions_per_run")
    if valu:
        busy_s = valu * 4 / (1024 * 2.4e9)
        out["valu"] = dict(wave_instructions_per_run=valu, issue_seconds_per_run=round(busy_s, 6),
                           busy_frac=round(busy_s / (k["seconds"] / steps), 3))
    # FP64 share of that issue, and the FP64 rate against the 78.6 TFLOP/s vector peak
    # (1024 SIMDs x 16 lanes x 2 FLOP x 2.4 GHz): FP64 wave-instructions x 64 lanes x the
    # measured active-lane fraction, one FLOP each (the build contracts no FMA)
    f64 = figures.get("fp64_wave_instructions_per_run")
    util = figures.get("valu_lane_utilization")
    if f64 and util:
        t = k["seconds"] / steps
        flops = f64 * 64 * util / t
        out["fp64_valu"] = dict(wave_instructions_per_run=f64, issue_frac=round(f64 * 4 / (1024 * 2.4e9) / t, 3),
                                active_lane_frac=round(util, 3), tflops=round(flops / 1e12, 2), peak_tflops=78.6,
                                frac=round(flops / 78.6e12, 4))
    return out


def host_cores():
    """The host cores this process may run on (the box allots a share of its CPUs per GPU;
    OMP_NUM_THREADS, when set, is that share)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    n = min(aff, int(env)) if env and env.isdigit() and int(env) > 0 else aff
    return n, aff


def cpu_baseline(side):
    """The CPU oracle (restatement of the reference's single-rank path, OpenMP loops as in
    the reference) on every core of this process's CPU share, on a bounded sample of the
    headline workload: the same radius and seed in a periodic cube of side `side`, run to
    maximal packing."""
    import _oracle as O
    threads, aff = host_cores()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O.build()
    w = HEADLINE
    t0 = time.perf_counter()
    o = O.run(3, [0.0] * 3, [side] * 3, [(w["r"], 1.0, 0)], seed=w["seed"])
    dt = time.perf_counter() - t0
    return dict(value=o["n"] / dt, unit="spheres/s", cores=threads, kind="port",
                host_cpus=os.cpu_count(), affinity_cpus=aff,
                sample=f"3D r={w['r']} seed {w['seed']} in the periodic cube of side {side} "
                       f"({side ** 3:.1%} of the workload volume) to maximal packing: {o['n']} spheres in {dt:.2f} s")


def batched(R, active):
    """The batched workload: 256 independent 3D realizations (box edge L ~ U[0.2, 1],
    r ~ U[0.004, 0.008], seeds 0..255, each to maximal packing), timed two ways: as
    the device work queue (`active` realizations in flight on one stream, one launch
    per step for all of them) and one realization after another through the
    single-context API (uniform_generate + extraction per realization).  Every placed
    set (extracted inside the timed region) is checked against the oracle's SHA-256."""
    from _common import set_hash
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "batch_3d_256.json")))
    real = fx["realizations"]
    L = [c["L"] for c in real]
    r = [c["r"] for c in real]
    s = [c["seed"] for c in real]
    # warm-up: the pools' device and pinned blocks for `active` realizations in flight, threads
    R.batch_queue(3, L[:active], r[:active], s[:active], active=active)

    def one_at_a_time():
        out = []
        for k in range(len(real)):
            dom = R.Domain([0.0] * 3, [L[k]] * 3, r[k])
            R.uniform_generate(dom, r[k], 6000, 10, s[k])
            p_, r_, _ = dom.extract_arrays()
            out.append((p_, r_))
            del dom
        return {"spheres": out, "counts": np.array([len(x[1]) for x in out], dtype=np.uint64)}

    res = {}
    for name, fn in (("queue", lambda: R.batch_queue(3, L, r, s, active=active, extract=True)),
                     ("plain", one_at_a_time)):
        print(f"#   batched: {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            out = fn()
        except RuntimeError as e:  # a failed device check or extraction checksum: reported, not a value
            res[name] = dict(seconds=None, value=None, sets_match_oracle="error", error=str(e)[:300])
            continue
        dt = time.perf_counter() - t0
        bad = [k for k, (p, rr) in enumerate(out["spheres"])
               if len(rr) != fx["oracle"][k]["n"] or set_hash(p, rr) != fx["oracle"][k]["sha256"]]
        res[name] = dict(seconds=round(dt, 3), value=round(float(out["counts"].sum()) / dt, 1),
                         sets_match_oracle=f"{256 - len(bad)}/256", mismatched=bad[:8])
        if bad:
            res[name]["value"] = None  # a wrong set is not a result
        if "stats" in out:
            res[name]["stats"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in out["stats"].items()}
    q, p = res["queue"]["value"], res["plain"]["value"]
    return {"workload": "256 independent 3D monodisperse realizations, periodic box [0,L]^3, L~U[0.2,1], "
                        "r~U[0.004,0.008], seeds 0..255, each to maximal packing (extraction in the timed region)",
            "realizations": len(real), "spheres": int(sum(c["n"] for c in fx["oracle"])), "active": active,
            "unit": "spheres/s", "value": q, "plain_value": p,
            "speedup_vs_plain": round(q / p, 3) if q and p else None, **{
                k + "_" + kk: v for k, d in res.items() for kk, v in d.items() if kk != "value"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-side", type=float, default=0.4)
    ap.add_argument("--batch-active", type=int, default=64,
                    help="realizations in flight in the device work queue form of the batched measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched measurement")
    ap.add_argument("--no-secondary", action="store_true", help="skip the r=0.005 line")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned form of the headline")
    ap.add_argument("--only", choices=["headline", "secondary", "learned", "batched"], default=None,
                    help="run one workload's timed steps and nothing else (for profiling)")
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
    if args.only == "learned":
        print(json.dumps(learned(R, args.steps, None)), flush=True)
        return
    if args.only == "batched":
        print(json.dumps(batched(R, args.batch_active)), flush=True)
        return
    if args.only:
        w = HEADLINE if args.only == "headline" else SECONDARY
        m = measure(R, w, args.steps, args.warmup)
        print(json.dumps({"workload": w["name"], "value": round(m["value"], 1), "ms_per_step": round(m["ms"], 2),
                          "parity_vs_oracle": m["parity"], "phases_ms": phases(m["stats"]),
                          "profile": per_kernel(m["prof"], args.steps)}),
              flush=True)
        return

    print(f"# headline: {args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr, flush=True)
    head = measure(R, HEADLINE, args.steps, args.warmup)
    stats = head["stats"]
    figures = profile_figures(HEADLINE["profile"])
    sec = None
    if not args.no_secondary:
        print("# secondary", file=sys.stderr, flush=True)
        m = measure(R, SECONDARY, args.steps, args.warmup)
        sec = {"workload": SECONDARY["name"], "value": round(m["value"], 1), "unit": "spheres/s",
               "ms_per_step": round(m["ms"], 2), "spheres": m["stats"]["accepted"], "draws": m["stats"]["draws"],
               "refines": m["stats"]["refines"], "parity_vs_oracle": m["parity"], "phases_ms": phases(m["stats"]),
               "roofline_k_cover_frac": (roofline(m["prof"], args.steps, {}) or {}).get("frac"),
               "profile": per_kernel(m["prof"], args.steps)}
    lrn = None
    if not args.no_learned:
        print("# learned", file=sys.stderr, flush=True)
        lrn = learned(R, args.steps, head["value"])
    bat = None
    if not args.no_batch:
        print("# batched", file=sys.stderr, flush=True)
        bat = batched(R, args.batch_active)
    cpu = None
    if not args.no_cpu_baseline:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_side)

    out = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "spheres/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(head["ms"], 2), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": HEADLINE["name"] + " (plain GPU form)", "spheres": stats["accepted"],
                   "draws": stats["draws"], "refines": stats["refines"], "nvox_max": stats["nvox_max"],
                   "mt_words": stats["words"], "parity_vs_oracle": head["parity"]},
        "roofline": roofline(head["prof"], args.steps, figures),
        "source_sha16": source_sha16(),
        "phases_ms": phases(stats),
        "profile": per_kernel(head["prof"], args.steps),
        "secondary": sec,
        "learned": lrn,
        "batched": bat,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
