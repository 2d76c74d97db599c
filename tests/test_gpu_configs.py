"""GPU parity at the workloads' real sizes: the HIP path through the C-ABI
against the CPU oracle, bit for bit (positions included: both sides evaluate
the same FP32 operations, so the tolerance is 0).

  c2   1,048,576 particles, periodic 100 um cube, A + B <-> C, C -> 0
  c3   4,194,304 particles in 64 Gaussian blobs, reflective sphere r = 60,
       absorbing plane z = 0, 200 um cube (coordinates up to 2000 cells wide
       in a 256-cell wrapped grid: the pair search's face-reach margins at
       large coordinates)
  c4   all 512 replicas of the batched ensemble
  c5   jobs of the serving stream (2, 3 and 4 species) through the slot
       engine at the serving layout
  the reference's own scenarios test1 (Lotka-Volterra) and test2
  (Michaelis-Menten), particleSystem.cpp:319-511

A few steps each, so that the oracle finishes in seconds; the bench runs the
same engines for the full step counts.
"""
import numpy as np
import pytest

from gpusmoldyn_amd import configs

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu-test] {msg}", flush=True)


def _same(gp, gt, rp, rt, what):
    assert len(gt) == len(rt), f"{what}: count {len(gt)} vs {len(rt)}"
    assert np.array_equal(gt, rt), f"{what}: types"
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), f"{what}: positions"


def _one(smd, orc, spec, pos, types, steps, what, check_every=None, cap=None):
    g = smd.gauss_table()
    e = smd.Engine(0)
    e.add_replica(spec, pos, types)
    e.finalize()
    sim = orc.sim_for(smd, spec, pos, types, cap=cap or 8 * len(types) + 1024)
    every = check_every or steps
    done = 0
    while done < steps:
        k = min(every, steps - done)
        e.step(k)
        sim.step(g, k)
        done += k
        gp, gt = e.get(0)
        rp, rt = sim.state()
        _same(gp, gt, rp, rt, f"{what} after {done} steps")
    assert np.array_equal(e.get_streams(0), sim.mt), f"{what}: stream states"
    e.close()
    return gt


def test_c2_full_size(smd, orc, gpu):
    spec, pos, types = configs.c2()
    banner(f"c2 full size n={len(types)} 6 steps")
    _one(smd, orc, spec, pos, types, 6, "c2", check_every=2)


def test_c3_full_size(smd, orc, gpu):
    spec, pos, types = configs.c3()
    banner(f"c3 full size n={len(types)} 4 steps")
    gt = _one(smd, orc, spec, pos, types, 4, "c3", check_every=2)
    # the absorbing plane and the sphere both acted (not vacuous)
    assert len(gt) < len(types) * 1.01


def test_c4_all_replicas(smd, orc, gpu):
    reps = configs.c4()
    steps = 6
    banner(f"c4 all {len(reps)} replicas, {sum(len(t) for _, _, t in reps)} particles, {steps} steps")
    e = smd.Engine(0)
    for spec, pos, types in reps:
        e.add_replica(spec, pos, types)
    e.finalize()
    e.step(steps)
    g = smd.gauss_table()
    reacted = 0
    for r, (spec, pos, types) in enumerate(reps):
        sim = orc.sim_for(smd, spec, pos, types)
        sim.step(g, steps)
        gp, gt = e.get(r)
        rp, rt = sim.state()
        _same(gp, gt, rp, rt, f"c4 replica {r}")
        reacted += int((gt == 3).sum())
    assert reacted > 0
    e.close()


def test_c4_segmented_equals_global(smd, gpu):
    """The full c4 ensemble in the benchmark's tuned layout (16 particles per
    stream) stepped with the segmented sort (two passes) and with one sort
    over the 9.25M global keys (three passes of 4096-key tiles): the same
    particles, bit for bit."""
    from gpusmoldyn_amd import autotune

    reps = autotune.apply_ensemble(configs.c4(), {c: (0, 16) for c in range(10, 17)})
    steps = 4
    res = []
    for seg in (1, 0):
        banner(f"c4 tuned layout, segmented sort {seg}, {steps} steps")
        e = smd.Engine(0)
        for spec, pos, types in reps:
            e.add_replica(spec, pos, types)
        e.finalize()
        e.set_knob(smd.KNOB_SEGMENTED_SORT, seg)
        assert e.sort_plan() == ((2, True) if seg else (3, False))
        e.step(steps)
        res.append([e.get(r) for r in range(len(reps))])
        e.close()
    for r, ((p1, t1), (p0, t0)) in enumerate(zip(*res)):
        _same(p1, t1, p0, t0, f"c4 replica {r}")


def test_c5_jobs_match_oracle(smd, orc, gpu):
    """Jobs of the serving stream itself (of each species count 2, 3 and 4,
    its first three below 1e7 particle-steps), served together through the
    slot engine at the serving layout (32 slots of 2^18 buckets), each equal
    to the oracle running the job alone."""
    from gpusmoldyn_amd import serving

    small = [j for j in serving.c5_jobs() if len(j.types) * j.steps <= 1e7]
    jobs = [j for sp in (2, 3, 4) for j in [x for x in small if x.spec.num_types - 1 == sp][:3]]
    jobs.sort(key=lambda j: j.arrival)
    assert {j.spec.num_types - 1 for j in jobs} == {2, 3, 4} and len(jobs) == 9
    banner(f"c5 jobs {[(j.jid, len(j.types), j.steps, j.spec.num_types - 1) for j in jobs]}")
    srv = serving.Server(n_slots=32, slot_bucket_bits=18, slot_streams=32768, capacity=1 << 22,
                         inject_cap=1 << 20)
    results, stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == len(jobs)
    by_id = {r.jid: r for r in results}
    g = smd.gauss_table()
    for j in jobs:
        sim = orc.sim_for(smd, j.spec, j.pos, j.types, cap=8 * len(j.types) + 1024)
        sim.step(g, j.steps)
        rp, rt = sim.state()
        r = by_id[j.jid]
        _same(r.pos, r.types, rp, rt, f"c5 job {j.jid}")


def test_reference_test1_lotka_volterra(smd, orc, gpu):
    """A -> 2A, A + B -> 2B, B -> 0 in the reference world: the prey doubles
    most steps, so the engine grows its buffers on the way."""
    spec, pos, types = configs.ref_test1()
    banner("reference test1 (Lotka-Volterra) 10 steps")
    gt = _one(smd, orc, spec, pos, types, 10, "test1", check_every=2, cap=1 << 21)
    assert (gt == 1).sum() > 16 * (types == 1).sum() and (gt == 2).sum() > 0


def test_reference_test2_michaelis_menten(smd, orc, gpu):
    spec, pos, types = configs.ref_test2()
    banner("reference test2 (Michaelis-Menten) 300 steps")
    gt = _one(smd, orc, spec, pos, types, 300, "test2", check_every=100)
    assert (gt == 3).sum() + (gt == 4).sum() > 0
