"""Benchmark of the gpusmoldyn timestep on one MI355X.

  python bench.py [--steps K] [--warmup W] [--config c2|c3|c4|c1|c5]

A step is one full timestep (diffuse + unimolecular reactions + boundaries +
hash + sort + cell table + pair search + bimolecular reactions + compaction)
of every particle of the workload, with the particle state resident in HBM.
`value` is particle-steps/s = (sum over timed steps of the particles entering
the step) / wall time.  Prints one JSON line.

c5 is the serving workload: jobs arriving as a Poisson stream are admitted
into replica slots and stepped together; `value` is then the aggregate
particle-steps/s of the whole run and `serving` holds p50/p99 job latency.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpusmoldyn_amd as smd  # noqa: E402
from gpusmoldyn_amd import configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/sec (diffuse+sort+react) 1 GPU, batched vs plain; p99 job latency"


def stage_bytes(name, n, S, nb, passes):
    """Algorithmic bytes of one launch of each stage for n particles, S RNG
    streams, nb buckets (compulsory reads + writes, each datum once)."""
    if name == "producer":
        # record (x, y, z, type) 16 in; record 16, key 4, birth type 4 out; stream state
        return 40 * n + 172 * S
    if name == "sort":
        # per digit pass: count (key 4) + scatter (key, value 8 in, 8 out); first pass values implicit;
        # the last pass also copies the packed record of each particle into sorted order (16 + 16)
        return 20 * n * passes - 4 * n + 32 * n
    if name == "first":
        return 4 * (nb + 1) + 4 * n
    if name == "candidates":
        return 24 * n
    if name == "resolve":
        return 4 * n
    if name == "scatter":
        return 44 * n
    return 0


def build(cfg):
    if cfg == "c1":
        reps = [configs.c1()]
        steps = configs.C1_STEPS
        desc = dict(workload="c1: A+B->C, 8192 A + 8192 B, 20 um reflective cube, 1000 steps")
    elif cfg == "c2":
        reps = [configs.c2()]
        steps = configs.C2_STEPS
        desc = dict(workload="c2: 1,048,576 particles, A+B<->C (kf 5e-3, kr 0.1) + C->0 (0.01/s), periodic 100 um",
                    particles=1 << 20, streams=131072, grid="2^7 per axis")
    elif cfg == "c3":
        reps = [configs.c3()]
        steps = configs.C3_STEPS
        desc = dict(workload="c3: 4,194,304 particles, 64 Gaussian blobs, A+B<->C, reflective sphere r=60, "
                             "absorbing z=0, 200 um", particles=1 << 22, streams=262144, grid="256x256x128 x 0.1 um")
    elif cfg == "c4":
        reps = configs.c4()
        steps = configs.C4_STEPS
        desc = dict(workload="c4: 512 replicas of c1, n log-uniform 2k-64k, kf log-uniform 1e-4..1e-2",
                    replicas=len(reps), particles=int(sum(len(r[2]) for r in reps)))
    else:
        raise SystemExit(f"unknown config {cfg}")
    return reps, steps, desc


def make_engine(reps, radix_bits=0):
    e = smd.Engine(0, radix_bits)
    for spec, pos, types in reps:
        e.add_replica(spec, pos, types)
    e.finalize()
    return e


def timed(e, steps):
    e.sync()
    ps0 = e.particle_steps()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    t1 = time.perf_counter()
    ps1 = e.particle_steps()
    return ps1 - ps0, t1 - t0


# stages that are one kernel each, and that kernel's name in the profiles
SINGLE_KERNEL = {"producer": "k_producer", "first": "k_first", "candidates": "k_candidates", "resolve": "k_resolve",
                 "scatter": "k_scatter"}
PROFILE = {"c2": "step_kernels_c2_1M.md", "c3": "step_kernels_c3_4M.md", "c4": "step_kernels_c4_9M.md"}


def profiled_traffic(cfg, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size), with
    that summary's average duration; (None, None) without a summary."""
    name = PROFILE.get(cfg)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, None, None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == kernel and cells[6] not in ("None", "traffic_bytes"):
            return int(cells[6]), float(cells[2]), "profiles/" + name
    return None, None, None


def stage_graph_bytes(e, steps, ps):
    """Compulsory HBM bytes of one step of the reference's stage graph, each
    stage reading its inputs and writing its outputs once:
    N (140 + 16 P) + 12 C + 336 S with P = ceil(K / 4) passes of 4-bit digits
    over K key bits, C buckets and S streams (N = mean particles per step)."""
    lay = e.layout()
    n = ps / max(steps, 1)
    P = (lay["key_bits"] + 3) // 4
    return n * (140 + 16 * P) + 12 * lay["total_buckets"] + 336 * lay["total_streams"]


def roofline(e, reps, cfg, nprof=5):
    prof = e.profile(nprof)
    lay = e.layout()
    n = int(e.counts().sum())
    S = lay["total_streams"]
    rb = e.radix_bits()
    passes = (lay["key_bits"] + rb - 1) // rb
    rows = {}
    for k, us in prof.items():
        b = stage_bytes(k, n, S, lay["total_buckets"], passes)
        rows[k] = dict(us=us, bytes=b, gbs=(b / (us * 1e-6) / 1e9) if us > 0 else 0.0)
    # the dominant kernel: the longest stage that is a single kernel
    dom = max((k for k in rows if k in SINGLE_KERNEL), key=lambda k: rows[k]["us"])
    step_us = sum(r["us"] for r in rows.values())
    step_bytes = sum(r["bytes"] for r in rows.values())
    ach = rows[dom]["gbs"]
    traffic, prof_us, prof_file = profiled_traffic(cfg, SINGLE_KERNEL[dom])
    rl = {"bound": "hbm", "kernel": SINGLE_KERNEL[dom], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
          "bytes_per_launch": rows[dom]["bytes"], "avg_us": round(rows[dom]["us"], 2),
          "profile": prof_file, "profile_avg_us": prof_us}
    step = {"us": round(step_us, 2), "bytes": step_bytes,
            "gbs": round(step_bytes / (step_us * 1e-6) / 1e9, 1) if step_us > 0 else 0.0,
            "stages_us": {k: round(v["us"], 2) for k, v in rows.items()}}
    return rl, step


def cpu_baseline(reps, budget_s=20.0, threads=0):
    """The oracle (C++ -O3) stepping the first steps of the workload for about
    budget_s seconds, on `threads` OpenMP threads (0: all host cores)."""
    import oracle

    L = oracle.lib()
    all_threads = L.orc_num_threads()
    L.orc_set_threads(threads or all_threads)
    spec, pos, types = reps[0]
    sim = oracle.sim_for(smd, spec, pos, types, cap=4 * len(types) + 1024)
    g = smd.gauss_table()
    done_ps, t_used, k = 0, 0.0, 0
    while t_used < budget_s and k < 1000:
        n = sim.n
        t0 = time.perf_counter()
        sim.step(g, 1)
        t_used += time.perf_counter() - t0
        done_ps += n
        k += 1
    cores = L.orc_num_threads()
    L.orc_set_threads(all_threads)
    return {"value": done_ps / t_used, "unit": "particle-steps/s", "cores": cores,
            "kind": "port",
            "sample": f"CPU oracle (C++ -O3, OpenMP producer/pair search, {cores} threads), first {k} steps of "
                      f"the same workload ({len(types)} particles), {t_used:.1f} s"}


def serve(args):
    from gpusmoldyn_amd import serving

    jobs = serving.c5_jobs(n_jobs=args.jobs)
    srv = serving.Server(n_slots=32, slot_bucket_bits=18, slot_streams=32768, capacity=1 << 24)
    # warm-up: one small job through the same slots
    srv.run([serving.make_job(-1, 0.0, 4096, 20, 3, 1)], realtime=False)
    results, st = srv.run(jobs, realtime=True)
    srv.close()
    n_ps = st["particle_steps"]
    return {
        "metric": METRIC,
        "value": st["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": st["engine_steps"],
        "warmup": 1,
        "ms_per_step": st["wall_s"] * 1e3 / max(st["engine_steps"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"c5: {len(jobs)} jobs, Poisson 20 jobs/s, n log-uniform 1k-256k, steps log-uniform "
                               "100-10k, 2-4 species, random A+B<->C / decay tables, one stream per 2 particles; "
                               "32 slots of 2^18 buckets",
                   "jobs": len(jobs), "particle_steps": n_ps},
        "serving": {"p50_latency_s": st["p50_s"], "p99_latency_s": st["p99_s"], "wall_s": st["wall_s"],
                    "jobs": st["jobs"], "last_arrival_s": jobs[-1].arrival},
        "roofline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--radix-bits", type=int, default=0, help="sort digit width (4, 8, 11; 0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="c4: skip the one-replica-per-launch comparison")
    ap.add_argument("--no-graphs", action="store_true", help="one launch sequence per step instead of graphs")
    ap.add_argument("--sort", type=int, default=1, help="1: count/scan/scatter (default), 0: single-sweep radix sort")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--jobs", type=int, default=200, help="c5: number of jobs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also time a plain copy kernel over BYTES (reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); replicas and slabs would shard "
              "across GPUs (a replica set or a spatial slab per GPU), but this build runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        print(json.dumps(serve(args)), flush=True)
        return
    probe = None
    if args.copy_probe:
        us = smd.copy_probe(args.copy_probe, reps=10)
        probe = {"bytes": args.copy_probe, "avg_us": us, "gbs": 2 * args.copy_probe / (us * 1e-6) / 1e9}
    reps, steps, desc = build(args.config)
    steps = args.steps if args.steps is not None else steps
    e = make_engine(reps, args.radix_bits)
    if args.no_graphs:
        e.set_graphs(False)
    if args.sort != 1:
        e.set_sort(args.sort)
    e.step(args.warmup)
    ps, secs = timed(e, steps)
    rl, step = roofline(e, reps, args.config)
    out = {
        "metric": METRIC,
        "value": ps / secs,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": desc,
        "roofline": rl,
        "step_roofline": step,
    }
    b_step = stage_graph_bytes(e, steps, ps)
    gbs = b_step / (secs / max(steps, 1)) / 1e9
    out["step_stage_graph"] = {"bytes": int(b_step), "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if probe:
        out["copy_probe"] = probe
    if args.config == "c4" and not args.no_plain:
        # the same 512-replica workload, one simulation per launch sequence
        pps, psecs = 0, 0.0
        for k, rep in enumerate(reps):
            p = make_engine([rep], args.radix_bits)
            p.set_graphs(False)
            p.step(args.warmup)
            a, b = timed(p, steps)
            pps += a
            psecs += b
            p.close()
            if k % 64 == 63:
                print(f"[bench] plain replicas done: {k + 1}/{len(reps)}", file=sys.stderr, flush=True)
        plain = pps / psecs
        out["batched_vs_plain"] = {"batched": out["value"], "plain": plain, "ratio": out["value"] / plain,
                                   "plain_seconds": psecs, "batched_seconds": secs}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(reps, args.cpu_budget)
        out["cpu_baseline_single_thread"] = cpu_baseline(reps, args.cpu_budget / 2, threads=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
