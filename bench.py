"""Benchmark of the gpusmoldyn timestep on one MI355X.

  python bench.py [--steps K] [--warmup W] [--config all|c1|c2|c3|c4|c5|c5mix]

A step is one full timestep (diffuse + unimolecular reactions + boundaries +
hash + sort + cell table + pair search + bimolecular reactions + compaction)
of every particle of the workload, with the particle state resident in HBM.
particle-steps/s = (sum over timed steps of the particles entering the step)
/ wall time.  Prints one JSON line.

The default (`--config all`) measures what the headline metric names
("batched vs plain; p99 job latency"):
  value            c4, the 512-replica batched ensemble (two-step hipGraphs,
                   grids per replica size class and launch shapes autotuned
                   before the timed region), particle-steps/s over K timed
                   steps after W warm-up steps;
  batched_vs_plain the same 512 replicas run one simulation per launch
                   sequence, eager and graph-captured, K steps each;
  roofline         the dominant single kernel of the c4 step;
  c3_4M            the 4,194,304-particle clustered step (the 60%-of-peak
                   target is stated on it): its stage-graph fraction of peak;
  serving          c5, the Poisson job stream through the slot engine: p50 and
                   p99 job latency and aggregate particle-steps/s;
  cpu_baseline     the CPU oracle on a bounded sample of c4.
`--config cN` runs one workload alone (profiling runs use these; c5mix is the
serving engine at the fixed resident mix its roofline is measured on).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpusmoldyn_amd as smd  # noqa: E402
from gpusmoldyn_amd import configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/sec (diffuse+sort+react) 1 GPU, batched vs plain; p99 job latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Sorted records carry their particle's index (no sorted values written) while
# every index fits in 27 bits.
PACKED_INDEX_MAX = 1 << 27


def stage_bytes(name, n, S, nb, passes, packed=True):
    """Algorithmic bytes of one launch of each stage for n particles, S RNG
    streams, nb buckets (compulsory reads + writes, each datum once)."""
    if name == "producer":
        # record (x, y, z, type) 16 in; record 16, key 4 out (birth type and position only at
        # births); stream state
        return 36 * n + 172 * S
    if name == "sort":
        # per digit pass: count (key 4) + scatter (key, value 8 in, 8 out); first pass values implicit;
        # the last pass also copies the packed record of each particle into sorted order (16 + 16),
        # with the particle's index in the record instead of a sorted value written (4 less)
        return 20 * n * passes - 4 * n + 32 * n - (4 * n if packed else 0)
    if name == "first":
        return 4 * (nb + 1) + 4 * n
    if name == "candidates":
        return 24 * n
    if name == "resolve":
        return 4 * n
    if name == "scatter":
        # record 16 + fourth component 4 in, the same out (birth types only in tiles with births)
        return 40 * n
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
        desc = dict(workload="c4: 512 replicas of c1's network (A+B->C, reflective cube at 2.048 per um^3), "
                             "n log-uniform 2k-64k, kf log-uniform 1e-4..1e-2; grids and particles per random "
                             "stream autotuned per replica size class (before tuning: about one cell and 1/8 "
                             "stream per particle)",
                    replicas=len(reps), particles=int(sum(len(r[2]) for r in reps)), parallelism="one GPU")
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
PROFILE = {"c2": "step_kernels_c2_1M.md", "c3": "step_kernels_c3_4M.md", "c4": "step_kernels_c4_9M.md",
           "c5": "step_kernels_c5_mix8.md"}
SERVING_RESIDENT = 8  # jobs resident in the serving engine's measured mix


def profiled_traffic(cfg, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size), with
    that summary's average duration; (None, None, None) without a summary."""
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


def roofline(e, cfg, nprof=5, streams=None, buckets=None):
    """Per-stage HIP-event times of `nprof` eager steps (on the engine's
    stream) and the dominant single kernel's roofline.  streams / buckets:
    the streams and buckets in use when they are fewer than the engine's
    layout (a slot engine's idle slots)."""
    prof = e.profile(nprof)
    lay = e.layout()
    n = int(e.counts().sum())
    S = streams if streams is not None else lay["total_streams"]
    nb = buckets if buckets is not None else lay["total_buckets"]
    passes = e.sort_plan()[0]  # two for an ensemble sorted replica by replica
    packed = lay["capacity"] <= PACKED_INDEX_MAX
    rows = {}
    for k, us in prof.items():
        b = stage_bytes(k, n, S, nb, passes, packed)
        rows[k] = dict(us=us, bytes=b, gbs=(b / (us * 1e-6) / 1e9) if us > 0 else 0.0)
    dom = max((k for k in rows if k in SINGLE_KERNEL), key=lambda k: rows[k]["us"])
    step_us = sum(r["us"] for r in rows.values())
    step_bytes = sum(r["bytes"] for r in rows.values())
    ach = rows[dom]["gbs"]
    # the kernels with a launch-shape knob are templates on their block size
    kname = SINGLE_KERNEL[dom]
    if dom == "producer":
        kname += f"<{e.get_knob(smd.KNOB_PRODUCER_BLOCK)}>"
    elif dom == "candidates":
        kname += f"<{e.get_knob(smd.KNOB_PAIR_BLOCK)}>"
    traffic, prof_us, prof_file = profiled_traffic(cfg, kname)
    rl = {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
          "bytes_per_launch": rows[dom]["bytes"], "avg_us": round(rows[dom]["us"], 2),
          "profile": prof_file, "profile_avg_us": prof_us}
    step = {"us": round(step_us, 2), "bytes": step_bytes,
            "gbs": round(step_bytes / (step_us * 1e-6) / 1e9, 1) if step_us > 0 else 0.0,
            "stages_us": {k: round(v["us"], 2) for k, v in rows.items()},
            "stages_frac": {k: round(v["gbs"] / HBM_PEAK_GBS, 4) for k, v in rows.items() if v["bytes"]}}
    return rl, step


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def load_oracle():
    """The oracle built with -march=native on this host (a few seconds of
    g++), or the prebuilt x86-64-v3 one if that build fails."""
    native_dir = os.path.join(ROOT, "oracle", "_build_native")
    flags = "-O3 -march=native"
    try:
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "native", f"NATIVE_DIR={native_dir}"],
                       check=True, capture_output=True, timeout=120)
        os.environ["SMD_ORACLE_LIB"] = os.path.join(native_dir, "libsmdoracle.so")
    except (subprocess.SubprocessError, OSError) as ex:
        log(f"native oracle build failed ({ex}); using the prebuilt -march=x86-64-v3 oracle")
        flags = "-O3 -march=x86-64-v3"
    import oracle

    oracle.lib()
    return oracle, flags


def cpu_baseline(reps, steps, budget_s, threads):
    """The oracle on a bounded sample of the workload: one replica stepped on
    `threads` OpenMP threads until budget_s seconds or `steps` steps; or, for
    an ensemble, its replicas in order for `steps` steps each, `threads` of
    them at a time on one thread each, as many as the budget allows (oracle
    set-up outside the timed region).  Returns the report dict."""
    from concurrent.futures import ThreadPoolExecutor

    oracle, flags = load_oracle()
    L = oracle.lib()
    g = smd.gauss_table()
    model = cpu_model()
    if len(reps) == 1:
        L.orc_set_threads(threads)
        spec, pos, types = reps[0]
        sim = oracle.sim_for(smd, spec, pos, types, cap=4 * len(types) + 1024)
        done_ps, k = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s and k < steps:
            done_ps += sim.n
            sim.step(g, 1)
            k += 1
        wall = time.perf_counter() - t0
        sample = f"first {k} steps of the workload ({len(types)} particles) on {threads} OpenMP threads"
    else:
        def make(rep):
            spec, pos, types = rep
            return oracle.sim_for(smd, spec, pos, types, cap=4 * len(types) + 1024)

        def run(sim):
            L.orc_set_threads(1)
            ps = 0
            for _ in range(steps):
                ps += sim.n
                sim.step(g, 1)
            return ps

        # estimate the replicas the budget holds from the first `threads` of them
        sims = [make(r) for r in reps[:threads]]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            done_ps = sum(ex.map(run, sims))
        first = time.perf_counter() - t0
        n_first = sum(len(r[2]) for r in reps[:threads])
        used = threads
        if first < budget_s:
            per_particle = first / max(n_first, 1) / threads  # wall seconds per particle, `threads` in flight
            more, acc = [], 0.0
            for r in reps[threads:]:
                acc += len(r[2]) * per_particle
                if first + acc > budget_s:
                    break
                more.append(r)
            sims = [make(r) for r in more]
            t1 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=threads) as ex:
                done_ps += sum(ex.map(run, sims))
            first += time.perf_counter() - t1
            used += len(more)
        wall = first
        sample = (f"replicas 0..{used - 1} of the {len(reps)} ({sum(len(r[2]) for r in reps[:used])} particles), "
                  f"{steps} steps each, {threads} at a time on one thread each")
    return {"value": done_ps / wall, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C++ {flags}, {model}, {host_cores()} host cores): {sample}, {wall:.1f} s",
            "cpu_model": model}


def autotune_ensemble(reps, path=None):
    """Grid refinement and particles per random stream per replica size class
    (the c4 workload's autotuned cell size), untimed; returns the tuned
    replicas and the report.  path: a JSON file of an earlier tuning to reuse
    (written when absent), so profiling runs step the same workload."""
    from gpusmoldyn_amd import autotune

    t0 = time.perf_counter()
    if path and os.path.exists(path):
        saved = json.load(open(path))
        choice = {int(c): (d, p) for c, (d, p) in saved["choice"].items()}
        table = {}
    else:
        choice, table = autotune.tune_ensemble(reps, deltas=(-1, 0, 1), per_stream=(8, 16, 32))
        if path:
            json.dump({"choice": {str(c): list(v) for c, v in choice.items()}}, open(path, "w"))
    tuned = autotune.apply_ensemble(reps, choice)
    rep = {"grid_delta_and_particles_per_stream_by_size_class": {f"2^{c}": list(v) for c, v in sorted(choice.items())},
           "ns_per_particle_step": {k: round(v, 4) for k, v in table.items()},
           "seconds": round(time.perf_counter() - t0, 2)}
    return tuned, rep


def run_steps(cfg, args, reps=None, tune=False, knob_file=None):
    """Engine over the workload (with tuned launch shapes when `tune`), W
    warm-up steps, K timed steps; returns (engine, replicas, steps,
    particle-steps, seconds, description, knob report)."""
    r, default_steps, desc = build(cfg)
    reps = reps or r
    steps = args.steps if args.steps is not None else default_steps
    e = make_engine(reps, args.radix_bits)
    if args.no_graphs:
        e.set_graphs(False)
    if args.sort != 1:
        e.set_sort(args.sort)
    e.step(args.warmup)
    knobs = None
    if tune:
        from gpusmoldyn_amd import autotune

        t0 = time.perf_counter()
        if knob_file and os.path.exists(knob_file):
            saved = json.load(open(knob_file))
            best = {smd.KNOB_PRODUCER_BLOCK: saved["producer_block"], smd.KNOB_PAIR_BLOCK: saved["pair_block"]}
            for k, v in best.items():
                e.set_knob(k, v)
            table = {}
        else:
            best, table = autotune.tune_knobs(e, steps=10)
        knobs = {"producer_block": best[smd.KNOB_PRODUCER_BLOCK], "pair_block": best[smd.KNOB_PAIR_BLOCK],
                 "us_per_step": {f"{a}/{b}": round(v, 1) for (a, b), v in sorted(table.items())},
                 "seconds": round(time.perf_counter() - t0, 2)}
        if knob_file and not os.path.exists(knob_file):
            json.dump({"producer_block": knobs["producer_block"], "pair_block": knobs["pair_block"]},
                      open(knob_file, "w"))
        e.step(2)
    ps, secs = timed(e, steps)
    return e, reps, steps, ps, secs, desc, knobs


def plain_forms(reps, steps, warmup, radix_bits, knobs=None):
    """The batched workload run one simulation per launch sequence: each
    replica in its own engine (with the batched run's launch shapes), W
    warm-up steps, then K eager steps and K graph-replayed steps timed
    separately.  Returns the totals."""
    eager_ps = eager_s = graph_ps = graph_s = 0.0
    for k, rep in enumerate(reps):
        p = make_engine([rep], radix_bits)
        if knobs:
            p.set_knob(smd.KNOB_PRODUCER_BLOCK, knobs["producer_block"])
            p.set_knob(smd.KNOB_PAIR_BLOCK, knobs["pair_block"])
        p.set_graphs(False)
        p.step(warmup)
        a, b = timed(p, steps)
        eager_ps += a
        eager_s += b
        p.set_graphs(True)
        p.step(2)  # capture outside the timed region, as in the batched run's warm-up
        a, b = timed(p, steps)
        graph_ps += a
        graph_s += b
        p.close()
        if k % 128 == 127:
            log(f"plain replicas done: {k + 1}/{len(reps)}")
    return eager_ps, eager_s, graph_ps, graph_s


def serve(args):
    from gpusmoldyn_amd import serving

    jobs = serving.c5_jobs(n_jobs=args.jobs)
    srv = serving.Server(n_slots=32, slot_bucket_bits=18, slot_streams=32768, capacity=1 << 24,
                         graphs=not args.no_graphs)
    # warm-up: one small job through the same slots
    srv.run([serving.make_job(-1, 0.0, 4096, 20, 3, 1)], realtime=False)
    results, st = srv.run(jobs, realtime=True)
    srv.close()
    return jobs, st


def serving_mix(jobs, resident=SERVING_RESIDENT, graphs=True):
    """The serving engine (32 slots of 2^18 buckets) with a typical resident
    mix: the stream's first `resident` jobs admitted together and stepped past
    their entry step.  Returns (server, streams in use, buckets in use)."""
    from gpusmoldyn_amd import serving

    srv = serving.Server(n_slots=32, slot_bucket_bits=18, slot_streams=32768, capacity=1 << 24, graphs=graphs)
    e = srv.e
    mix = jobs[:resident]
    for k, j in enumerate(mix):
        e.admit(k, j.spec, j.pos, j.types)
    e.step(4)
    streams = sum(int(j.spec.num_streams) for j in mix)
    buckets = sum(1 << sum(int(j.spec.grid_bits[d]) for d in range(3)) for j in mix)
    return srv, streams, buckets


def serving_roofline(jobs, nprof=5):
    """Stage times (HIP events) and the dominant kernel's roofline of the
    serving engine at its typical resident mix, `nprof` eager steps timed."""
    srv, streams, buckets = serving_mix(jobs)
    e = srv.e
    rl, step = roofline(e, "c5", nprof, streams=streams, buckets=buckets)
    step["resident_particles"] = int(e.counts().sum())
    srv.close()
    return rl, step


def c5mix_line(args):
    """The serving engine at its typical resident mix (the default stream's
    first jobs), K timed steps after W warm-up steps: the workload of the
    serving roofline, run alone so that a profiler's byte counters see it."""
    from gpusmoldyn_amd import serving

    jobs = serving.c5_jobs()
    srv, streams, buckets = serving_mix(jobs, graphs=not args.no_graphs)
    e = srv.e
    e.step(args.warmup)
    steps = args.steps if args.steps is not None else 20
    ps, secs = timed(e, steps)
    rl, step = roofline(e, "c5", 5, streams=streams, buckets=buckets)
    out = {
        "metric": METRIC, "value": ps / secs, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": secs * 1e3 / max(steps, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c5 mix: the serving engine (32 slots of 2^18 buckets) with the default stream's "
                               f"first {SERVING_RESIDENT} jobs resident", "resident_particles": int(e.counts().sum())},
        "roofline": rl, "step_roofline": step,
    }
    if args.copy_probe:
        us = smd.copy_probe(args.copy_probe, reps=10)
        out["copy_probe"] = {"bytes": args.copy_probe, "avg_us": us, "gbs": 2 * args.copy_probe / (us * 1e-6) / 1e9}
    srv.close()
    return out


def serving_report(jobs, st):
    return {"p50_latency_s": round(st["p50_s"], 4), "p99_latency_s": round(st["p99_s"], 4),
            "particle_steps_per_s": st["particle_steps_per_s"], "wall_s": round(st["wall_s"], 2),
            "jobs": st["jobs"], "engine_steps": st["engine_steps"], "last_arrival_s": round(jobs[-1].arrival, 2),
            "workload": f"c5: {len(jobs)} jobs, Poisson 20 jobs/s, n log-uniform 1k-256k, steps log-uniform "
                        "100-10k, 2-4 species, random A+B<->C / decay tables, one stream per 2 particles; "
                        "32 slots of 2^18 buckets"}


def c5_line(args):
    jobs, st = serve(args)
    out = {
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
        "config": {"workload": serving_report(jobs, st)["workload"], "jobs": len(jobs),
                   "particle_steps": st["particle_steps"]},
        "serving": serving_report(jobs, st),
    }
    out["roofline"], out["step_roofline"] = serving_roofline(jobs)
    return out


def base_line(cfg, args, e, steps, ps, secs, desc):
    rl, step = roofline(e, cfg)
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
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="all")
    ap.add_argument("--radix-bits", type=int, default=0, help="sort digit width (4, 8, 11; 0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="c4: skip the one-replica-per-launch comparison")
    ap.add_argument("--no-serving", action="store_true", help="all: skip the c5 serving run")
    ap.add_argument("--no-c3", action="store_true", help="all: skip the c3 step")
    ap.add_argument("--no-graphs", action="store_true", help="one launch sequence per step instead of graphs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c4: the default grids and launch shapes instead of the autotuned ones")
    ap.add_argument("--tuning", default=None, metavar="PREFIX",
                    help="c4: reuse the tuning saved at PREFIX.grid.json / PREFIX.launch.json (saved there "
                         "when absent), so that profiling runs step the benchmark's workload")
    ap.add_argument("--sort", type=int, default=1, help="1: count/scan/scatter (default), 0: single-sweep radix sort")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--jobs", type=int, default=200, help="c5: number of jobs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also time a plain copy kernel over BYTES (reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); replicas and slabs would shard "
              "across GPUs (a replica set or a spatial slab per GPU), but this build runs on a single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        print(json.dumps(c5_line(args)), flush=True)
        return
    if args.config == "c5mix":
        print(json.dumps(c5mix_line(args)), flush=True)
        return
    probe = None
    if args.copy_probe:
        us = smd.copy_probe(args.copy_probe, reps=10)
        probe = {"bytes": args.copy_probe, "avg_us": us, "gbs": 2 * args.copy_probe / (us * 1e-6) / 1e9}
    cfg = "c4" if args.config == "all" else args.config
    if args.config == "all" and args.steps is None:
        args.steps = 20
    log(f"{cfg}: building the workload")
    tuned, tune_rep = None, None
    tune = cfg == "c4" and not args.no_autotune
    if tune:
        tuned, tune_rep = autotune_ensemble(build(cfg)[0], args.tuning + ".grid.json" if args.tuning else None)
        log(f"{cfg}: tuned {tune_rep['grid_delta_and_particles_per_stream_by_size_class']} in "
            f"{tune_rep['seconds']} s")
    e, reps, steps, ps, secs, desc, knobs = run_steps(cfg, args, reps=tuned, tune=tune,
                                                      knob_file=args.tuning + ".launch.json" if args.tuning else None)
    log(f"{cfg}: {steps} steps in {secs * 1e3:.2f} ms")
    out = base_line(cfg, args, e, steps, ps, secs, desc)
    e.close()
    if tune:
        out["autotune"] = dict(tune_rep, launch=knobs)
    if probe:
        out["copy_probe"] = probe
    if cfg == "c4" and not args.no_plain:
        log("c4: plain forms (each replica alone)")
        eps, es, gps, gs = plain_forms(reps, steps, args.warmup, args.radix_bits, knobs)
        out["batched_vs_plain"] = {
            "batched": out["value"], "plain_eager": eps / es, "plain_graph": gps / gs,
            "ratio_vs_eager": out["value"] / (eps / es), "ratio_vs_graph": out["value"] / (gps / gs),
            "batched_seconds": secs, "plain_eager_seconds": es, "plain_graph_seconds": gs,
            "note": "plain: each replica in its own engine, one launch sequence per step (eager) or two-step "
                    "hipGraphs (graph), K steps each after the same warm-up"}
    if args.config == "all" and not args.no_c3:
        log("c3: 4M-particle clustered step")
        e3, _, s3, ps3, secs3, d3, _ = run_steps("c3", args)
        l3 = base_line("c3", args, e3, s3, ps3, secs3, d3)
        e3.close()
        out["c3_4M"] = {"workload": d3["workload"], "steps": s3, "ms_per_step": l3["ms_per_step"],
                        "particle_steps_per_s": l3["value"], "stage_graph": l3["step_stage_graph"],
                        "roofline": l3["roofline"], "stages_us": l3["step_roofline"]["stages_us"]}
    if args.config == "all" and not args.no_serving:
        log("c5: serving stream")
        jobs, st = serve(args)
        out["serving"] = serving_report(jobs, st)
        rl5, st5 = serving_roofline(jobs)
        out["serving"]["roofline"] = rl5
        out["serving"]["stages_us"] = st5["stages_us"]
    if not args.no_cpu:
        log("CPU baseline")
        threads = host_cores()
        out["cpu_baseline"] = cpu_baseline(reps, steps, args.cpu_budget, threads)
        out["cpu_baseline_single_thread"] = cpu_baseline(reps, steps, args.cpu_budget / 2, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
