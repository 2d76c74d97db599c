"""GPU parity: the HIP path through the C-ABI against the CPU oracle and the
golden vectors.  Integer, index and RNG results must match bit for bit; so
must positions, because both sides evaluate the same FP32 operations with
explicit fused multiply-adds (the tolerance for positions is therefore 0)."""
import os

import numpy as np
import pytest

from gpusmoldyn_amd import BC_ABSORB, BC_CLAMP, BC_PERIODIC, BC_REFLECT, configs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def banner(msg):
    print(f"\n[gpu-test] {msg}", flush=True)


def test_mt_matches_reference_dcmt(smd, gpu):
    f = np.load(os.path.join(GOLD, "mt_streams.npz"))
    streams, seeds, outs = f["streams"], f["seeds"], f["outputs"]
    # stream k of smd_mt_generate uses record k % 32768: build 32768 streams
    # whose seeds put the fixture seeds at the fixture stream ids.
    S = 32768
    all_seeds = np.arange(S, dtype=np.uint32) * np.uint32(2654435761)
    all_seeds[streams] = seeds
    for batch3 in (False, True):
        banner(f"mt_generate S={S} count=512 batch3={batch3}")
        got = smd.mt_generate(all_seeds, 512, device=gpu, batch3=batch3)
        assert np.array_equal(got[streams], outs[:, :512])
    banner("mt_generate 7 streams, 4096 draws")
    got = smd.mt_generate(seeds, 4096, device=gpu)
    # streams 0..6 use records 0..6: the fixture rows of records 0, 1, 2 line up
    for row in range(3):
        assert np.array_equal(got[row], outs[row])


@pytest.mark.parametrize("radix_bits", [4, 8, 11])
@pytest.mark.parametrize("n,bits", [(0, 8), (1, 8), (4095, 12), (4096, 12), (4097, 20), (100003, 24), (1 << 20, 21),
                                    (3000000, 25)])
def test_sort_pairs_stable(smd, gpu, n, bits, radix_bits):
    # up to 6M keys the sort runs 2048-key tiles, beyond that 4096-key tiles
    banner(f"sort n={n} bits={bits} radix={radix_bits}")
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[: n // 3] = keys[0]  # long runs of equal keys
    vals = np.arange(n, dtype=np.uint32)
    k, v = smd.sort_pairs(keys, vals, bits, radix_bits, device=gpu)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_sort_pairs_large_tiles(smd, gpu):
    # above 6M keys: 4096-key tiles, ragged last tile (the engine's own cases,
    # where the expected count and the bound pick different tile sizes, run in
    # the full-size c3 and c4 tests)
    n, bits = 13_000_001, 23
    banner(f"sort n={n} bits={bits} radix=8")
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
    keys[1000:200000] = 12345
    vals = np.arange(n, dtype=np.uint32)
    k, v = smd.sort_pairs(keys, vals, bits, 8, device=gpu)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


# ---------------------------------------------------------------- workloads
def w_c1_small():
    spec, pos, types = configs.c1(nA=3000, nB=3000, kf=5.0, n_streams=2048, grid_bits=(5, 5, 5))
    return spec, pos, types


def w_c1_ref():
    # the CPU-runnable workload itself, reference stream mapping (S = 32768)
    return configs.c1()


def w_periodic_reversible():
    spec = configs.make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=1e-3,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=1.0),
                   dict(order=1, reactants=[3], products=[1, 2], rate=50.0, pgem=0.2),
                   dict(order=1, reactants=[3], products=[], rate=5.0)],
        box_lo=[0.0] * 3, box_hi=[5.0] * 3, faces=[BC_PERIODIC] * 6, grid_bits=(4, 4, 4),
        num_streams=1000, unbind_offsets=1)
    rng = np.random.default_rng(11)
    n = 8000
    pos = configs.uniform_box(rng, n, 0.0, 5.0)
    types = rng.integers(1, 4, size=n).astype(np.int32)
    return spec, pos, types


def w_offsets():
    # products placed at rotated unbinding offsets (reversible pair with a
    # reachable geminate probability): exercises sincos/theta bit parity
    spec = configs.make_spec(
        D=[0.0, 0.01, 0.01, 0.005], dt=1e-1,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=0.05),
                   dict(order=1, reactants=[3], products=[1, 2], rate=3.0, pgem=0.2)],
        box_lo=[0.0] * 3, box_hi=[4.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(4, 4, 4),
        num_streams=700, unbind_offsets=1)
    rng = np.random.default_rng(12)
    n = 6000
    pos = configs.uniform_box(rng, n, 0.0, 4.0)
    types = rng.integers(1, 4, size=n).astype(np.int32)
    return spec, pos, types


def w_c3_small():
    spec, pos, types = configs.c3(n=60000, n_blobs=8, sigma=1.0, n_streams=4096, grid_bits=(5, 5, 5), cell=0.15,
                                  L=20.0)
    spec.sphere[0][3] = 6.0
    spec.rxn[0].rate = 2.0
    return spec, pos, types


def w_clamp():
    spec = configs.make_spec(
        D=[0.0, 0.3, 0.1], dt=1e-2, reactions=[], box_lo=[-1.0] * 3, box_hi=[1.0] * 3, faces=[BC_CLAMP] * 6,
        grid_bits=(6, 6, 6), particle_radius=1.0 / 64.0)
    rng = np.random.default_rng(13)
    n = 20000
    pos = configs.uniform_box(rng, n, -1.0, 1.0)
    types = np.where(np.arange(n) < 3 * n // 4, 1, 2).astype(np.int32)
    return spec, pos, types


def w_zeroth(seed=4357):
    # order-0 creation (one mean below 12, one split into pieces) feeding
    # A + B -> C and C -> 0
    V = 125.0
    spec = configs.make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=1e-2,
        reactions=[dict(order=0, products=[1], rate=4.0 / (1e-2 * V)),
                   dict(order=2, reactants=[1, 2], products=[3], rate=2.0),
                   dict(order=0, products=[2], rate=40.0 / (1e-2 * V)),
                   dict(order=1, reactants=[3], products=[], rate=2.0)],
        box_lo=[0.0] * 3, box_hi=[5.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(4, 4, 4), num_streams=300,
        seed=seed, volume=V)
    rng = np.random.default_rng(14)
    n = 3000
    pos = configs.uniform_box(rng, n, 0.0, 5.0)
    types = rng.integers(1, 4, size=n).astype(np.int32)
    return spec, pos, types


WORKLOADS = {
    "c1_small": w_c1_small,
    "zeroth": w_zeroth,
    "periodic_reversible": w_periodic_reversible,
    "offsets": w_offsets,
    "c3_small": w_c3_small,
    "clamp_reference_test3": w_clamp,
}


def _engine(smd, spec, pos, types, radix_bits=0):
    e = smd.Engine(0, radix_bits)
    e.add_replica(spec, pos, types)
    e.finalize()
    return e


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_cell_table(smd, gpu, name):
    # first[h] = lower bound of key h in the sorted keys (empty buckets
    # included; key total buckets marks removed particles), then the count
    spec, pos, types = WORKLOADS[name]()
    banner(f"cell table workload {name}")
    e = _engine(smd, spec, pos, types)
    got = e.debug_stage(2, len(types))
    first = e.debug_first()
    lay = e.layout()
    e.close()
    keys = got["keys"][: got["n_total"]]
    dead = lay["total_buckets"]
    assert len(first) == dead + 2
    want = np.searchsorted(keys, np.arange(dead + 1, dtype=np.uint64), side="left")
    assert np.array_equal(first[: dead + 1], want)
    assert first[dead] == got["n_live"]
    assert first[dead + 1] == got["n_total"]


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_stage_parity(smd, orc, gpu, name):
    spec, pos, types = WORKLOADS[name]()
    g = smd.gauss_table()
    n = len(types)
    has_bim = any(spec.rxn[k].order == 2 for k in range(spec.num_reactions))
    for stage in ((1, 2, 3, 4) if has_bim else (1,)):
        banner(f"stage {stage} workload {name} n={n}")
        e = _engine(smd, spec, pos, types)
        got = e.debug_stage(stage, n)
        e.close()
        sim = orc.sim_for(smd, spec, pos, types)
        ref = sim.stage(g, stage)
        assert np.array_equal(got["type"], ref["type"]), "types"
        assert np.array_equal(got["pos"].view(np.uint32), ref["pos"].view(np.uint32)), "positions"
        assert np.array_equal(got["keys"], ref["keys"]), "keys"
        if stage >= 2:
            assert got["n_live"] == ref["n_live"]
            assert np.array_equal(got["vals"], ref["vals"]), "sorted index"
        if stage >= 3:
            assert np.array_equal(got["cand"], ref["cand"]), "pair candidates"


@pytest.mark.parametrize("name,steps", [("c1_small", 60), ("zeroth", 40), ("periodic_reversible", 40), ("offsets", 40),
                                        ("c3_small", 20), ("clamp_reference_test3", 30)])
def test_multistep_parity(smd, orc, gpu, name, steps):
    spec, pos, types = WORKLOADS[name]()
    banner(f"multistep {name} steps={steps}")
    g = smd.gauss_table()
    e = _engine(smd, spec, pos, types)
    sim = orc.sim_for(smd, spec, pos, types, cap=8 * len(types) + 1024)
    for k in range(steps):
        e.step(1)
        sim.step(g, 1)
        if k % 10 == 9 or k == steps - 1:
            gp, gt = e.get(0)
            rp, rt = sim.state()
            assert len(gt) == len(rt), f"count at step {k}"
            assert np.array_equal(gt, rt), f"types at step {k}"
            assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), f"positions at step {k}"
    assert np.array_equal(e.get_streams(0), sim.mt), "stream states"
    e.close()


def test_c1_full_run_matches_oracle(smd, orc, gpu):
    """The CPU-runnable workload end to end (1000 steps, S = 32768 so that
    particle i uses stream i exactly as the reference does for N <= 32768)."""
    spec, pos, types = w_c1_ref()
    banner("c1 full 1000 steps")
    g = smd.gauss_table()
    e = _engine(smd, spec, pos, types)
    e.step(configs.C1_STEPS)
    sim = orc.sim_for(smd, spec, pos, types)
    sim.step(g, configs.C1_STEPS)
    gp, gt = e.get(0)
    rp, rt = sim.state()
    assert np.array_equal(gt, rt)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    # A + B -> C happened (k nA nB / V dt ~ 8.4/s)
    assert (gt == 3).sum() > 0


def test_ensemble_equals_alone(smd, gpu):
    reps = [w_c1_small(), w_zeroth(11), w_periodic_reversible(), w_offsets(), w_zeroth(12), w_clamp()]
    banner("ensemble of 6 vs alone")
    e = smd.Engine(0)
    for spec, pos, types in reps:
        e.add_replica(spec, pos, types)
    e.finalize()
    e.step(25)
    for r, (spec, pos, types) in enumerate(reps):
        a = _engine(smd, spec, pos, types)
        a.step(25)
        ap, at = a.get(0)
        bp, bt = e.get(r)
        assert np.array_equal(at, bt), f"replica {r} types"
        assert np.array_equal(ap.view(np.uint32), bp.view(np.uint32)), f"replica {r} positions"
        a.close()
    e.close()


def test_digit_widths_agree(smd, gpu):
    spec, pos, types = w_periodic_reversible()
    banner("radix 4 vs 8 vs 11")
    res = []
    for rb in (4, 8, 11):
        e = _engine(smd, spec, pos, types, radix_bits=rb)
        assert e.radix_bits() == rb
        e.step(15)
        res.append(e.get(0))
        e.close()
    for p, t in res[1:]:
        assert np.array_equal(res[0][1], t) and np.array_equal(res[0][0].view(np.uint32), p.view(np.uint32))


def test_checkpoint_resume(smd, gpu):
    spec, pos, types = w_periodic_reversible()
    banner("checkpoint/resume")
    a = _engine(smd, spec, pos, types)
    a.step(10)
    p10, t10 = a.get(0)
    st10 = a.get_streams(0)
    a.step(10)
    p20, t20 = a.get(0)
    b = _engine(smd, spec, p10, t10)
    b.set_streams(0, st10)
    b.step(10)
    q20, u20 = b.get(0)
    assert np.array_equal(t20, u20) and np.array_equal(p20.view(np.uint32), q20.view(np.uint32))


def test_growth_beyond_capacity(smd, orc, gpu):
    # A -> 2A doubles the population every few steps: the engine must grow its
    # buffers without losing or reordering particles
    spec = configs.make_spec(D=[0.0, 0.2], dt=1e-2, reactions=[dict(order=1, reactants=[1], products=[1, 1],
                                                                     rate=40.0)],
                             box_lo=[0.0] * 3, box_hi=[10.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(3, 3, 3),
                             num_streams=256)
    pos = configs.uniform_box(np.random.default_rng(21), 500, 0.0, 10.0)
    types = np.ones(500, np.int32)
    banner("growth")
    e = smd.Engine(0)
    e.add_replica(spec, pos, types)
    e.finalize(capacity=1024)
    sim = orc.sim_for(smd, spec, pos, types, cap=1 << 20)
    g = smd.gauss_table()
    e.step(12)
    sim.step(g, 12)
    gp, gt = e.get(0)
    rp, rt = sim.state()
    assert len(gt) > 1024
    assert np.array_equal(gt, rt) and np.array_equal(gp.view(np.uint32), rp.view(np.uint32))


def test_ensemble_growth_beyond_capacity(smd, orc, gpu):
    """Three replicas whose populations double every few steps (A -> 2A) in
    an engine of small capacity: growth reallocates the particle buffers and
    the segmented sort's scratch mid-run (graphs on); every replica equals
    the oracle."""
    reps = []
    for k, (n, gb) in enumerate(((400, (3, 3, 3)), (700, (4, 3, 3)), (250, (3, 4, 3)))):
        spec = configs.make_spec(D=[0.0, 0.2], dt=1e-2,
                                 reactions=[dict(order=1, reactants=[1], products=[1, 1], rate=40.0)],
                                 box_lo=[0.0] * 3, box_hi=[10.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=gb,
                                 num_streams=128 + 64 * k, seed=500 + k)
        reps.append((spec, configs.uniform_box(np.random.default_rng(30 + k), n, 0.0, 10.0), np.ones(n, np.int32)))
    banner("ensemble growth")
    e = smd.Engine(0)
    for spec, pos, types in reps:
        e.add_replica(spec, pos, types)
    e.finalize(capacity=2048)
    assert e.sort_plan()[1]
    e.step(12)
    g = smd.gauss_table()
    total = 0
    for r, (spec, pos, types) in enumerate(reps):
        sim = orc.sim_for(smd, spec, pos, types, cap=1 << 20)
        sim.step(g, 12)
        gp, gt = e.get(r)
        rp, rt = sim.state()
        assert np.array_equal(gt, rt) and np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), r
        total += len(gt)
    assert total > 2048
    e.close()


def test_ensemble_zeroth_order(smd, orc, gpu):
    """Order-0 creation in an ensemble (segmented sort): particles created on
    the device in each replica join its segment at compaction; every replica
    equals the oracle."""
    reps = [w_zeroth(seed=4357 + 11 * k) for k in range(3)]
    banner("ensemble with order-0 creation")
    e = smd.Engine(0)
    for spec, pos, types in reps:
        e.add_replica(spec, pos, types)
    e.finalize()
    assert e.sort_plan()[1]
    e.step(15)
    g = smd.gauss_table()
    for r, (spec, pos, types) in enumerate(reps):
        sim = orc.sim_for(smd, spec, pos, types)
        sim.step(g, 15)
        gp, gt = e.get(r)
        rp, rt = sim.state()
        assert np.array_equal(gt, rt) and np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), r
    e.close()


@pytest.mark.parametrize("name", ["periodic_reversible", "offsets", "c3_small"])
def test_mixed_ensemble_matches_oracle(smd, orc, gpu, name):
    """An ensemble mixing a workload with c1 replicas of other grids (periodic
    boxes, unbinding offsets, sphere and absorbing plane beside plain ones),
    stepped with the segmented sort: every replica equals the oracle."""
    spec, pos, types = WORKLOADS[name]()
    reps = [(spec, pos, types),
            configs.c1(nA=900, nB=1100, kf=4.0, n_streams=200, seed=61, rng_seed=62, grid_bits=(4, 4, 4)),
            configs.c1(nA=1500, nB=1500, kf=8.0, n_streams=320, seed=63, rng_seed=64, grid_bits=(5, 4, 5))]
    banner(f"mixed ensemble with {name}")
    e = smd.Engine(0)
    for sp, p, t in reps:
        e.add_replica(sp, p, t)
    e.finalize()
    e.step(10)
    g = smd.gauss_table()
    for r, (sp, p, t) in enumerate(reps):
        sim = orc.sim_for(smd, sp, p, t)
        sim.step(g, 10)
        gp, gt = e.get(r)
        rp, rt = sim.state()
        assert np.array_equal(gt, rt) and np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), (name, r)
    e.close()


def test_extinction_and_empty(smd, gpu):
    spec = configs.make_spec(D=[0.0, 0.2], dt=1e-1, reactions=[dict(order=1, reactants=[1], products=[], rate=100.0)],
                             box_lo=[0.0] * 3, box_hi=[10.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(3, 3, 3),
                             num_streams=64)
    pos = configs.uniform_box(np.random.default_rng(22), 300, 0.0, 10.0)
    banner("extinction")
    e = _engine(smd, spec, pos, np.ones(300, np.int32))
    e.step(5)
    assert e.counts()[0] == 0
    e.step(3)  # stepping an empty system is a no-op
    assert e.counts()[0] == 0


def test_graph_replay_equals_eager(smd, gpu):
    """Pairs of steps replayed from a captured hipGraph give the eager result,
    replay after replay, with eager steps, syncs and host reads in between."""
    for name in ("periodic_reversible", "c1_small", "zeroth"):
        spec, pos, types = WORKLOADS[name]()
        banner(f"graph vs eager {name}")
        a = _engine(smd, spec, pos, types)
        b = _engine(smd, spec, pos, types)
        b.set_graphs(False)
        for k, n in enumerate((2, 4, 1, 6, 3, 2)):
            a.step(n)
            b.step(n)
            ap, at = a.get(0)
            bp, bt = b.get(0)
            assert np.array_equal(at, bt), f"{name} types after chunk {k}"
            assert np.array_equal(ap.view(np.uint32), bp.view(np.uint32)), f"{name} positions after chunk {k}"
        assert a.particle_steps() == b.particle_steps()
        a.close()
        b.close()


def test_autotune_picks_a_valid_fastest_config(smd, gpu):
    from gpusmoldyn_amd import autotune

    spec, pos, types = w_c1_small()
    banner("autotune")
    best, table = autotune.tune(spec, pos, types, steps=5, max_delta=1)
    assert best in table and table[best] == min(table.values())
    assert len(table) >= 2
    # the tuned grid still runs and conserves A + C and B + C
    e = smd.Engine(0, best[1])
    e.add_replica(autotune.apply(spec, best[0], best[2]), pos, types)
    e.finalize()
    e.step(20)
    c = e.species_counts(0)
    assert c[1] + c[3] == (types == 1).sum() and c[2] + c[3] == (types == 2).sum()


def test_sort_kinds_agree(smd, gpu):
    """The count/scan/scatter sort and the single-sweep sort give the same
    stable order, hence the same particles step after step."""
    spec, pos, types = w_periodic_reversible()
    banner("sort kinds")
    a = _engine(smd, spec, pos, types)
    b = _engine(smd, spec, pos, types)
    b.set_sort(0)
    a.step(15)
    b.step(15)
    ap, at = a.get(0)
    bp, bt = b.get(0)
    assert np.array_equal(at, bt) and np.array_equal(ap.view(np.uint32), bp.view(np.uint32))


def test_sorted_index_unpacked_above_2_27(smd, gpu):
    """Sorted records carry their particle's index above the type bits while
    every index fits in 27 bits; an engine with a larger capacity keeps the
    sorted values instead.  Both give the same particles, bit for bit."""
    spec, pos, types = w_periodic_reversible()
    banner("packed sorted index vs sorted values (capacity 2^27 + 4096)")
    a = _engine(smd, spec, pos, types)
    b = smd.Engine(0)
    b.add_replica(spec, pos, types)
    b.finalize((1 << 27) + 4096)
    try:
        a.step(15)
        b.step(15)
        ap, at = a.get(0)
        bp, bt = b.get(0)
        assert np.array_equal(at, bt) and np.array_equal(ap.view(np.uint32), bp.view(np.uint32))
        assert len(at) != len(types) or not np.array_equal(at, types)  # reactions happened
    finally:
        a.close()
        b.close()


def test_launch_shapes_agree(smd, gpu):
    """Every producer and pair-search workgroup size gives the same particles,
    bit for bit (the autotuner picks among them by time alone)."""
    spec, pos, types = w_c3_small()
    banner("launch shapes")
    res = []
    for pb, qb in ((256, 256), (128, 64), (512, 128), (512, 64)):
        e = _engine(smd, spec, pos, types)
        e.set_knob(smd.KNOB_PRODUCER_BLOCK, pb)
        e.set_knob(smd.KNOB_PAIR_BLOCK, qb)
        assert e.get_knob(smd.KNOB_PRODUCER_BLOCK) == pb and e.get_knob(smd.KNOB_PAIR_BLOCK) == qb
        e.step(6)
        res.append(e.get(0))
        e.close()
    for p, t in res[1:]:
        assert np.array_equal(res[0][1], t) and np.array_equal(res[0][0].view(np.uint32), p.view(np.uint32))
    with pytest.raises(smd.SmdError):
        e = _engine(smd, spec, pos, types)
        try:
            e.set_knob(smd.KNOB_PAIR_BLOCK, 96)
        finally:
            e.close()


@pytest.mark.parametrize("big", [(4, 4, 4), (6, 5, 5), (6, 6, 5), (6, 6, 6)])
def test_segmented_sort_agrees(smd, gpu, big):
    """An ensemble sorts replica by replica on local cells (digit widths 8+8,
    9+8 and 9+9 bits, from the largest grid: 2^12 to 2^17 buckets; 2^18 takes
    the global sort either way) or over all keys at once: the same particles,
    bit for bit, with births, bimolecular deaths, an absorbing face, an empty
    replica and one below a tile."""
    reps = []
    for k, (n, gb) in enumerate(((30000, big), (3000, (4, 4, 4)), (0, (3, 3, 3)), (200, (5, 4, 4)),
                                  (12000, (5, 5, 5)), (45000, big))):
        spec, pos, types = configs.c3(n=max(n, 1), n_blobs=4, sigma=1.0, n_streams=1024, grid_bits=gb, cell=0.15,
                                      L=20.0, seed=300 + k)
        spec.sphere[0][3] = 6.0
        spec.rxn[0].rate = 2.0
        reps.append((spec, pos[:n], types[:n]))
    banner(f"segmented sort, largest grid {big}")
    res = []
    for seg in (1, 0):
        e = smd.Engine(0)
        for spec, pos, types in reps:
            e.add_replica(spec, pos, types)
        e.finalize()
        e.set_knob(smd.KNOB_SEGMENTED_SORT, seg)
        assert e.get_knob(smd.KNOB_SEGMENTED_SORT) == seg
        passes, segmented = e.sort_plan()
        assert segmented == (seg == 1 and sum(big) <= 17)
        assert passes == (2 if segmented else -(-e.layout()["key_bits"] // 8))
        e.step(8)
        res.append([e.get(r) for r in range(len(reps))])
        e.close()
    for r, ((p1, t1), (p0, t0)) in enumerate(zip(*res)):
        assert np.array_equal(t1, t0) and np.array_equal(p1.view(np.uint32), p0.view(np.uint32)), r
    assert sum(len(t) for _, t in res[0]) > 0


def test_segmented_sort_long_segments(smd, gpu):
    """Replicas of more than 64 sort tiles (4096 keys each) have their tile
    counts scanned per digit instead of by one workgroup per replica: two such
    replicas (300k and 530k particles) between short ones give the global
    sort's particles bit for bit."""
    reps = []
    for k, n in enumerate((20000, 300000, 5000, 530000, 70000)):
        spec, pos, types = configs.c3(n=n, n_blobs=8, sigma=2.0, n_streams=4096, grid_bits=(6, 6, 5), cell=0.2,
                                      L=40.0, seed=700 + k)
        reps.append((spec, pos, types))
    banner("segmented sort, long segments")
    res = []
    for seg in (1, 0):
        e = smd.Engine(0)
        for spec, pos, types in reps:
            e.add_replica(spec, pos, types)
        e.finalize()
        e.set_knob(smd.KNOB_SEGMENTED_SORT, seg)
        assert e.sort_plan()[1] == (seg == 1)
        e.step(4)
        res.append([e.get(r) for r in range(len(reps))])
        e.close()
    for r, ((p1, t1), (p0, t0)) in enumerate(zip(*res)):
        assert np.array_equal(t1, t0) and np.array_equal(p1.view(np.uint32), p0.view(np.uint32)), r
    assert len(res[0][3][1]) > 64 * 4096


def test_segmented_sort_many_replicas(smd, gpu):
    """More replicas than one workgroup's lanes (1100 small replicas, some
    empty): the segmented sort's tile table and layout scans loop over chunks
    of replicas; results equal the global sort's bit for bit."""
    rng = np.random.default_rng(5)
    reps = []
    for k in range(1100):
        n = int(rng.integers(0, 160)) if k % 37 else 0
        spec, pos, types = configs.c1(nA=n // 2, nB=n - n // 2, kf=50.0, n_streams=64, seed=1000 + k,
                                      rng_seed=2000 + k, grid_bits=(3, 3, 3))
        reps.append((spec, pos, types))
    banner("segmented sort, 1100 replicas")
    res = []
    for seg in (1, 0):
        e = smd.Engine(0)
        for spec, pos, types in reps:
            e.add_replica(spec, pos, types)
        e.finalize()
        e.set_knob(smd.KNOB_SEGMENTED_SORT, seg)
        assert e.sort_plan()[1] == (seg == 1)
        e.step(6)
        res.append([e.get(r) for r in range(len(reps))])
        e.close()
    for r, ((p1, t1), (p0, t0)) in enumerate(zip(*res)):
        assert np.array_equal(t1, t0) and np.array_equal(p1.view(np.uint32), p0.view(np.uint32)), r
    assert sum(int((t == 3).sum()) for _, t in res[0]) > 0  # some pairs reacted


def test_tuned_ensemble_matches_oracle(smd, orc, gpu):
    """The per-size-class choices of the autotuner are valid simulations:
    replicas on refined and coarsened grids, with their own or fewer streams,
    equal the oracle on the same grids and streams."""
    from gpusmoldyn_amd import autotune

    reps = configs.c4(n_rep=12, nmin=2048, nmax=16384)
    deltas = {autotune.size_class(len(t)): ((-1, 16), (1, None))[k % 2] for k, (_, _, t) in enumerate(reps)}
    tuned = autotune.apply_ensemble(reps, deltas)
    banner(f"tuned ensemble deltas {deltas}")
    e = smd.Engine(0)
    for spec, pos, types in tuned:
        e.add_replica(spec, pos, types)
    e.finalize()
    e.step(8)
    g = smd.gauss_table()
    for r, (spec, pos, types) in enumerate(tuned):
        sim = orc.sim_for(smd, spec, pos, types)
        sim.step(g, 8)
        gp, gt = e.get(r)
        rp, rt = sim.state()
        assert np.array_equal(gt, rt) and np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), r
    e.close()
