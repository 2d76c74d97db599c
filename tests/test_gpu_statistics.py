"""Statistical checks of the GPU engine against the oracle and against theory.

  * Two-sample KS test (p > 0.01) of species-count trajectories on c1's
    network: 256 GPU replicas vs 256 oracle replicas with independent seeds
    and initial positions, at four time points.
  * Smoluchowski: A + B -> C measured rate constant equals the requested kf
    within 2% (the binding radius of rxnparam.cpp:bindingradius is the one
    whose steady-state rate is kf, reaction_kernel.cuh applies it).
  * Mass action: A <-> B steady-state fraction equals k2 / (k1 + k2) within
    2% (first-order probabilities of smolreact.cpp:rxnsetrate).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, B, C = 1, 2, 3


def _box_spec(smd, reactions, L, faces, seed, num_streams=64, D=(0.0, 1.0, 1.0, 0.5), grid_bits=(5, 5, 5)):
    from gpusmoldyn_amd import configs

    return configs.make_spec(D=list(D), dt=1e-3, reactions=reactions, box_lo=[0.0] * 3, box_hi=[L] * 3,
                             faces=[faces] * 6, grid_bits=grid_bits, num_streams=num_streams, seed=seed)


def _uniform(rng, n, L):
    p = np.empty((n, 4), np.float32)
    p[:, :3] = rng.uniform(0.0, L, size=(n, 3)).astype(np.float32)
    p[:, 3] = 1.0
    return p


def test_ks_species_trajectories_vs_oracle(smd, orc, gpu):
    """c1's network (A + B -> C; D = 1, 1, 0.5; reflective cube at c1's
    density of 2.048 per um^3), 1024 A + 1024 B per replica and kf = 0.1 so
    that about 40 reactions happen in 400 steps: the C count at steps 100,
    200, 300 and 400 of 256 GPU replicas against 256 oracle replicas with
    independent seeds and positions, two-sample KS p > 0.01 at every point."""
    from concurrent.futures import ThreadPoolExecutor

    from scipy.stats import ks_2samp

    R, nA, L, kf, chunk, points = 256, 1024, 10.0, 0.1, 100, 4
    rx = [dict(order=2, reactants=[A, B], products=[C], rate=kf)]
    types = np.array([A] * nA + [B] * nA, np.int32)
    n0 = 2 * nA

    def spec(seed):
        return _box_spec(smd, rx, L, smd.BC_REFLECT, seed=seed, num_streams=256, grid_bits=(4, 4, 4))

    e = smd.Engine(0)
    for k in range(R):
        e.add_replica(spec(4357 + k), _uniform(np.random.default_rng(k), n0, L), types)
    e.finalize()
    gpu_c = np.empty((points, R), np.int64)
    for t in range(points):
        e.step(chunk)
        gpu_c[t] = n0 - e.counts().astype(np.int64)  # each A + B -> C removes one particle
    e.close()

    g = smd.gauss_table()
    L_orc = orc.lib()

    def oracle_replica(k):
        L_orc.orc_set_threads(1)  # one replica per worker thread
        sim = orc.sim_for(smd, spec(91000 + k), _uniform(np.random.default_rng(50000 + k), n0, L), types)
        out = []
        for _ in range(points):
            sim.step(g, chunk)
            out.append(n0 - sim.n)
        return out

    with ThreadPoolExecutor(max_workers=16) as ex:
        cpu_c = np.array(list(ex.map(oracle_replica, range(R))), np.int64).T
    for t in range(points):
        res = ks_2samp(gpu_c[t], cpu_c[t])
        print(f"KS step {(t + 1) * chunk}: GPU mean C {gpu_c[t].mean():.2f}, oracle {cpu_c[t].mean():.2f}, "
              f"p = {res.pvalue:.3f}", flush=True)
        assert res.pvalue > 0.01
    assert gpu_c[-1].mean() > 20


def smoluchowski_k_hat(counts_fn, step_fn, nA0, n0, V, steps, skip, dt=1e-3):
    """Maximum-likelihood rate constant of A + B -> C with nA = nB:
    sum(events) / sum(nA nB dt / V) over steps >= skip.  Each reaction
    removes one particle, so events and nA follow from the particle counts."""
    events, exposure = 0, 0.0
    prev = counts_fn()
    for t in range(steps):
        step_fn()
        nxt = counts_fn()
        if t >= skip:
            nA = nA0 - (n0 - prev)
            events += int((prev - nxt).sum())
            exposure += float((nA * nA).sum()) * dt / V
        prev = nxt
    return events, events / exposure


def test_smoluchowski_rate(smd, gpu):
    """A + B -> C at kf = 0.1 um^3/s (binding radius 0.029 um, a tenth of the
    diffusion-limited rate, so the Smoluchowski transient is below 1% after
    20 steps) in a periodic 10 um box: measured rate constant = kf +- 2%."""
    kf, R, nA0, L, steps, skip = 0.1, 64, 8000, 10.0, 500, 20
    rx = [dict(order=2, reactants=[A, B], products=[C], rate=kf)]
    types = np.array([A] * nA0 + [B] * nA0, np.int32)
    e = smd.Engine(0)
    for k in range(R):
        spec = _box_spec(smd, rx, L, smd.BC_PERIODIC, seed=4357 + k, num_streams=512)
        e.add_replica(spec, _uniform(np.random.default_rng(k), 2 * nA0, L), types)
    e.finalize()
    e.set_graphs(False)
    events, k_hat = smoluchowski_k_hat(lambda: e.counts().astype(np.int64), lambda: e.step(1), nA0, 2 * nA0,
                                       L ** 3, steps, skip)
    print(f"Smoluchowski: events {events}, k_hat {k_hat:.5f} vs kf {kf}")
    assert events > 50000
    assert abs(k_hat / kf - 1.0) < 0.02


def test_mass_action_first_order(smd, gpu):
    k1, k2, R, n, L = 2.0, 1.0, 64, 2000, 10.0
    rx = [dict(order=1, reactants=[A], products=[B], rate=k1),
          dict(order=1, reactants=[B], products=[A], rate=k2)]
    types = np.full(n, A, np.int32)
    e = smd.Engine(0)
    for k in range(R):
        spec = _box_spec(smd, rx, L, smd.BC_REFLECT, seed=4357 + k, D=(0.0, 1.0, 1.0), num_streams=256)
        e.add_replica(spec, _uniform(np.random.default_rng(k), n, L), types)
    e.finalize()
    e.step(4000)  # 4 s = 12 relaxation times of 1 / (k1 + k2)
    fa = []
    for _ in range(20):
        e.step(200)
        fa.append(np.mean([e.species_counts(r)[A] for r in range(R)]) / n)
    frac = float(np.mean(fa))
    want = k2 / (k1 + k2)
    print(f"mass action: fraction A {frac:.4f} vs {want:.4f}")
    assert abs(frac / want - 1.0) < 0.02


def test_reversible_equilibrium(smd, gpu):
    """A + B <-> C mass-action steady state within 2%: nC / (nA nB) =
    k_b / (kr V) with k_b = numrxnrate(step, a, b) / dt, the association rate
    of rxnparam's reversible radial distribution for the binding radius a and
    unbinding radius b that smolreact.cpp:393, 459 derive."""
    import os
    import sys

    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle import reversible_equilibrium, reversible_spec

    kr, want, _ = reversible_equilibrium(smd)
    R = 256
    e = smd.Engine(0)
    types = np.array([A] * 250 + [B] * 250 + [C] * 250, np.int32)
    for k in range(R):
        e.add_replica(reversible_spec(kr, 4357 + k), _uniform(np.random.default_rng(k), 750, 10.0), types)
    e.finalize()
    e.step(1000)
    got = []
    for _ in range(20):
        e.step(250)
        got.append(1000 - e.counts().astype(np.int64))  # nA = nB = 500 - nC, so n = 1000 - nC
    mean = float(np.mean(got))
    print(f"equilibrium: mean nC {mean:.2f} vs mass action {want:.2f}")
    assert abs(mean / want - 1.0) < 0.02
