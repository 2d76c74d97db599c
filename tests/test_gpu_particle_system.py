"""The ParticleSystem C++ class driven directly (tests/cpp/ps_direct.cpp, built
by __graft_entry__.build()): HIP's uint3/float3 in the constructor, getters
and randPos; the public reactionOrderInfo[3] and reactionList members; update()
and getArray() -- with the reference's Michaelis-Menten scenario (test2),
compared with the CPU oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from gpusmoldyn_amd import configs

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "ps_direct")
GOLD = os.path.join(HERE, "golden")


def _run(tmp_path, spec, pos, types, steps):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(bytes(spec))
        f.write(np.array([len(types), steps], np.uint32).tobytes())
        f.write(np.ascontiguousarray(pos, np.float32).tobytes())
        f.write(np.ascontiguousarray(types, np.int32).tobytes())
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return out.read_bytes()


class _Reader:
    def __init__(self, raw):
        self.raw, self.off = raw, 0

    def take(self, dtype, count):
        a = np.frombuffer(self.raw, dtype, count, self.off)
        self.off += a.nbytes
        return a


def test_particle_system_class_against_oracle(smd, orc, gpu, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/ps_direct not built (run __graft_entry__.build())"
    spec, pos, types = configs.ref_test2(n=5500)
    steps = 200
    print(f"\n[gpu-test] ParticleSystem class, test2, {len(types)} particles, {steps} updates", flush=True)
    rd = _Reader(_run(tmp_path, spec, pos, types, steps))
    m = int(rd.take(np.uint32, 1)[0])
    gp = rd.take(np.float32, 4 * m).reshape(m, 4)
    gt = rd.take(np.int32, m)
    sim = orc.sim_for(smd, spec, pos, types, cap=8 * len(types) + 1024)
    sim.step(smd.gauss_table(), steps)
    rp, rt = sim.state()
    assert len(rt) == m
    assert np.array_equal(gt, rt)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    assert (gt == 3).sum() + (gt == 4).sum() > 0

    # randPos: unirandOOF(-1, 1) x 3 per call from SFMT seeded with 0, the
    # reference's own sequence (golden vectors of its compiled SFMT.cpp)
    rp3 = rd.take(np.float32, 24)
    f = np.load(os.path.join(GOLD, "sfmt_uniform.npz"))
    assert int(f["seeds"][0]) == 0
    assert np.array_equal(rp3.view(np.uint32), f["values"][0][:24].view(np.uint32))

    # reactionOrderInfo (fillReactionOrderInfo, smolreact.cpp:173-271) of
    # test2: order 2: E + S -> ES (0); order 1: ES -> S + S (1), ES -> E + P (2)
    T = spec.num_types
    want = {
        0: dict(per_order=0, count=[0] * T, start=[0] * T, l2g=[], idx=[]),
        1: dict(per_order=2, count=[0, 0, 0, 2, 0], start=[0, 0, 0, 0, 2], l2g=[1, 2], idx=[1, 2]),
        2: dict(per_order=1, count=[0, 1, 1, 0, 0], start=[0, 0, 1, 2, 2], l2g=[0, 0], idx=[0]),
    }
    for o in range(3):
        per_order, size = (int(x) for x in rd.take(np.int32, 2))
        count = rd.take(np.int32, T).tolist()
        start = rd.take(np.int32, T).tolist()
        l2g = rd.take(np.int32, size).tolist()
        idx = rd.take(np.int32, per_order).tolist()
        w = want[o]
        assert (per_order, count, start, l2g, idx) == (w["per_order"], w["count"], w["start"], w["l2g"], w["idx"]), o
    der = smd.derive(spec)
    for r in range(spec.num_reactions):
        order, nprod, ra, rb, pa, pb = (int(x) for x in rd.take(np.int32, 6))
        b2, prob = rd.take(np.float32, 2)
        prodpos = rd.take(np.float32, 6)
        x = spec.rxn[r]
        assert (order, nprod) == (x.order, x.nprod)
        assert (ra, rb) == (x.reactants[0], x.reactants[1] if x.order == 2 else -1)
        assert (pa, pb) == (x.products[0] if x.nprod >= 1 else -1, x.products[1] if x.nprod >= 2 else -1)
        assert b2 == np.float32(der[r].bind_radius2) and prob == np.float32(der[r].prob)
        assert np.array_equal(prodpos, np.array(der[r].product_pos, np.float32))
    assert rd.off == len(rd.raw)
