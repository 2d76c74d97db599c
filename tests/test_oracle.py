"""Pins of the CPU oracle itself: math kernels against libm, the cell-list pair
search against O(N^2) brute force, the sort against a stable argsort, and the
step's statistics against closed forms (diffusion moments, first-order decay
and branching, bimolecular mass action)."""
import ctypes
import math

import numpy as np
import pytest

from gpusmoldyn_amd import BC_ABSORB, BC_CLAMP, BC_PERIODIC, BC_REFLECT, configs


def test_sincos_and_theta_against_libm(orc):
    L = orc.lib()
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32)
    s, c = ctypes.c_float(), ctypes.c_float()
    err = 0.0
    for x in xs:
        L.orc_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
        err = max(err, abs(s.value - math.sin(float(x))), abs(c.value - math.cos(float(x))))
    assert err < 4e-7
    us = np.concatenate([np.linspace(0, 1, 20001), [1e-9, 1e-6, 0.25, 0.2500001, 0.9999999]]).astype(np.float32)
    errt = max(abs(L.orc_theta(float(u)) - math.acos(1.0 - 2.0 * float(u))) for u in us)
    assert errt < 2e-6


def _small_bim_spec(periodic, n_streams=128, L=4.0, k=2.0, grid=(3, 3, 3)):
    faces = [BC_PERIODIC] * 6 if periodic else [BC_REFLECT] * 6
    return configs.make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=1e-3,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=k)],
        box_lo=[0.0] * 3, box_hi=[L] * 3, faces=faces, grid_bits=grid, num_streams=n_streams)


@pytest.mark.parametrize("periodic", [False, True])
def test_cell_search_equals_bruteforce(orc, smd, periodic):
    spec = _small_bim_spec(periodic)
    rng = np.random.default_rng(1)
    n = 3000
    pos = configs.uniform_box(rng, n, 0.0, 4.0)
    types = rng.integers(1, 3, size=n).astype(np.int32)
    sim = orc.sim_for(smd, spec, pos, types)
    st = sim.stage(smd.gauss_table(), 3)
    assert st["n_live"] == n
    spt = np.zeros((n, 4), np.float32)
    spt[:, :3] = st["pos"][st["vals"], :3]
    spt[:, 3] = st["type"][st["vals"]].view(np.float32)
    bf = np.zeros(n, np.int32)
    orc.lib().orc_candidates_bruteforce(ctypes.byref(sim.s.net), spt.ctypes.data_as(ctypes.c_void_p), n,
                                        bf.ctypes.data_as(ctypes.c_void_p))
    assert (bf >= 0).sum() > 50  # the test is not vacuous
    assert np.array_equal(st["cand"], bf)


def test_sort_is_stable_by_bucket(orc, smd):
    spec = _small_bim_spec(False)
    rng = np.random.default_rng(2)
    n = 5000
    pos = configs.uniform_box(rng, n, 0.0, 4.0)
    types = rng.integers(1, 3, size=n).astype(np.int32)
    sim = orc.sim_for(smd, spec, pos, types)
    k1 = sim.stage(smd.gauss_table(), 1)["keys"]
    sim2 = orc.sim_for(smd, spec, pos, types)
    st = sim2.stage(smd.gauss_table(), 2)
    order = np.argsort(k1, kind="stable")
    assert np.array_equal(st["vals"], order)
    assert np.array_equal(st["keys"], k1[order])


def _diffusion_spec(D=1.0, dt=1e-3, n_streams=0, faces=None, L=1000.0):
    return configs.make_spec(D=[0.0, D], dt=dt, reactions=[], box_lo=[0.0] * 3, box_hi=[L] * 3,
                             faces=faces or [BC_REFLECT] * 6, grid_bits=(4, 4, 4), num_streams=n_streams)


def test_diffusion_moments(orc, smd):
    spec = _diffusion_spec(D=1.0, dt=1e-3)
    n = 200000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = 500.0
    pos[:, 3] = 1.0
    sim = orc.sim_for(smd, spec, pos, np.ones(n, np.int32))
    g = smd.gauss_table()
    sim.step(g, 1)
    p, _ = sim.state()
    d = p[:, :3].astype(np.float64) - 500.0
    rate = float(np.sqrt(np.float32(2.0) * np.float32(1.0) * np.float32(1e-3)))
    var = rate * rate * 0.9996794
    # every displacement is rate * a table entry (500 + x is exact to ~3e-5)
    assert abs(d.mean()) < 5 * math.sqrt(var / (3 * n))
    assert abs(d.var() / var - 1.0) < 5 * math.sqrt(2.0 / (3 * n))
    assert np.all(p[:, 3] == 1.0)


def test_first_order_decay_and_branching(orc, smd):
    k1, k2, dt = 30.0, 10.0, 1e-2
    spec = configs.make_spec(
        D=[0.0, 0.0, 0.0, 0.0], dt=dt,
        reactions=[dict(order=1, reactants=[1], products=[2], rate=k1),
                   dict(order=1, reactants=[1], products=[], rate=k2)],
        box_lo=[0.0] * 3, box_hi=[10.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(2, 2, 2), num_streams=4096)
    der = smd.derive(spec)
    p_tot = 1.0 - math.exp(-(k1 + k2) * dt)
    p1 = k1 / (k1 + k2) * p_tot
    assert abs(der[0].prob - p1) < 1e-6
    assert abs(der[1].prob - (k2 / (k1 + k2) * p_tot) / (1.0 - p1)) < 1e-6
    n = 100000
    pos = configs.uniform_box(np.random.default_rng(3), n, 1.0, 9.0)
    sim = orc.sim_for(smd, spec, pos, np.ones(n, np.int32))
    sim.step(smd.gauss_table(), 1)
    _, t = sim.state()
    nB = (t == 2).sum()
    nA = (t == 1).sum()
    ndead = n - len(t)
    for cnt, p in ((nB, p1), (ndead, k2 / (k1 + k2) * p_tot), (nA, 1 - p_tot)):
        assert abs(cnt - n * p) < 5 * math.sqrt(n * p * (1 - p))


def test_split_appends_births_in_parent_order(orc, smd):
    spec = configs.make_spec(
        D=[0.0, 0.1, 0.1, 0.1], dt=1e-2,
        reactions=[dict(order=1, reactants=[1], products=[2, 3], rate=20.0)],
        box_lo=[0.0] * 3, box_hi=[10.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=(2, 2, 2), num_streams=512)
    n = 4000
    pos = configs.uniform_box(np.random.default_rng(4), n, 1.0, 9.0)
    sim = orc.sim_for(smd, spec, pos, np.ones(n, np.int32))
    sim.step(smd.gauss_table(), 1)
    p, t = sim.state()
    nsplit = (t[:n] == 2).sum()
    assert len(t) == n + nsplit
    assert np.all(t[n:] == 3)
    # each birth sits at its parent's pre-boundary position (same point)
    parents = np.nonzero(t[:n] == 2)[0]
    assert np.allclose(p[n:, :3], p[parents, :3], atol=1e-5)


def test_bimolecular_mass_action(orc, smd):
    # well-mixed regime (rms step / binding radius ~ 10): the simulated
    # reaction count per step follows k nA nB / V dt
    k, dt, L = 2e-2, 1e-3, 10.0
    spec = configs.make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=dt,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=k)],
        box_lo=[0.0] * 3, box_hi=[L] * 3, faces=[BC_PERIODIC] * 6, grid_bits=(5, 5, 5), num_streams=32768)
    n = 40000
    rng = np.random.default_rng(5)
    pos = configs.uniform_box(rng, n, 0.0, L)
    types = np.array([1, 2] * (n // 2), np.int32)
    sim = orc.sim_for(smd, spec, pos, types)
    g = smd.gauss_table()
    nsteps = 20
    expected = 0.0
    nA = nB = n // 2
    for _ in range(nsteps):
        expected += k * nA * nB / L ** 3 * dt
        sim.step(g, 1)
        _, t = sim.state()
        nA, nB = (t == 1).sum(), (t == 2).sum()
    got = n // 2 - nA
    assert abs(got - expected) < 4 * math.sqrt(expected) + 0.03 * expected


def test_boundaries(orc, smd):
    # reflective faces keep particles inside; absorbing -z removes; periodic wraps
    for faces, check in (
        ([BC_REFLECT] * 6, "inside"),
        ([BC_PERIODIC] * 6, "inside"),
        ([BC_REFLECT] * 4 + [BC_ABSORB, BC_REFLECT], "absorb"),
        ([BC_CLAMP] * 6, "clamp"),
    ):
        spec = _diffusion_spec(D=5.0, dt=1e-2, faces=faces, L=2.0, n_streams=1024)  # max step 1.16 < box
        spec.particle_radius = 1.0 / 64.0
        n = 20000
        pos = configs.uniform_box(np.random.default_rng(6), n, 0.0, 2.0)
        sim = orc.sim_for(smd, spec, pos, np.ones(n, np.int32))
        sim.step(smd.gauss_table(), 3)
        p, t = sim.state()
        if check == "inside":
            assert len(t) == n and p[:, :3].min() >= 0.0 and p[:, :3].max() < 2.0
        elif check == "absorb":
            assert len(t) < n and p[:, 2].min() >= 0.0 and p[:, :3].max() <= 2.0
        else:
            r = np.float32(1.0 / 64.0)
            assert p[:, :3].min() >= np.float32(0.0) + r and p[:, :3].max() <= np.float32(2.0) - r


def test_sphere_reflects_both_ways(orc, smd):
    spec = _diffusion_spec(D=20.0, dt=1e-2, L=10.0, n_streams=2048)
    spec.num_spheres = 1
    for q, v in enumerate((5.0, 5.0, 5.0, 2.0)):
        spec.sphere[0][q] = v
    rng = np.random.default_rng(7)
    n = 20000
    pos = configs.uniform_box(rng, n, 0.5, 9.5)
    r0 = np.linalg.norm(pos[:, :3] - 5.0, axis=1)
    keep = np.abs(r0 - 2.0) > 1e-3
    pos = pos[keep]
    inside0 = r0[keep] < 2.0
    sim = orc.sim_for(smd, spec, pos, np.ones(len(pos), np.int32))
    sim.step(smd.gauss_table(), 5)
    p, _ = sim.state()
    inside = np.linalg.norm(p[:, :3].astype(np.float64) - 5.0, axis=1) < 2.0
    assert np.array_equal(inside, inside0)


def test_smoluchowski_rate(orc, smd):
    """A + B -> C at kf = 0.1 um^3/s in a periodic box: the maximum-likelihood
    rate constant (events / sum nA nB dt / V) after a 20-step transient equals
    the kf that rxnparam.cpp:bindingradius turned into the binding radius,
    within 2% (about 35000 events over 16 replicas: 0.5% standard error)."""
    kf, R, nA0, L, steps, skip, dt = 0.1, 16, 8000, 10.0, 500, 20, 1e-3
    g = smd.gauss_table()
    types = np.array([1] * nA0 + [2] * nA0, np.int32)
    events, exposure = 0, 0.0
    for k in range(R):
        spec = configs.make_spec(
            D=[0.0, 1.0, 1.0, 0.5], dt=dt, reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=kf)],
            box_lo=[0.0] * 3, box_hi=[L] * 3, faces=[BC_PERIODIC] * 6, grid_bits=(5, 5, 5), num_streams=512,
            seed=4357 + k)
        sim = orc.sim_for(smd, spec, configs.uniform_box(np.random.default_rng(k), 2 * nA0, 0.0, L), types)
        for t in range(steps):
            cur = sim.n
            sim.step(g, 1)
            if t >= skip:
                nA = nA0 - (2 * nA0 - cur)  # each reaction removes one particle
                events += cur - sim.n
                exposure += nA * nA * dt / L ** 3
    k_hat = events / exposure
    assert events > 20000
    assert abs(k_hat / kf - 1.0) < 0.02, k_hat


def reversible_spec(kr, seed, kf=0.5, L=10.0, dt=1e-3, num_streams=256):
    return configs.make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=dt,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=kf),
                   dict(order=1, reactants=[3], products=[1, 2], rate=kr, pgem=0.2)],
        box_lo=[0.0] * 3, box_hi=[L] * 3, faces=[BC_PERIODIC] * 6, grid_bits=(5, 5, 5), num_streams=num_streams,
        seed=seed, unbind_offsets=1)


def reversible_equilibrium(smd, kf=0.5, L=10.0, dt=1e-3, n_total_a=500, target=4e-3):
    """A + B <-> C with the forward binding radius of the irreversible
    calibration (smolreact.cpp:393) and products placed at the unbinding
    radius of pgem = 0.2 (smolreact.cpp:459).  Associations then happen at the
    rate of rxnparam's reversible radial distribution, k_b =
    numrxnrate(step, a, b) / dt = kf / (1 - pgem), so the simulated
    equilibrium is nC / (nA nB) = k_b / (kr V) (not kf / (kr V)).  Returns
    (kr giving nC / (nA nB) = target, expected mean nC)."""
    der = smd.derive(reversible_spec(1.0, 1, kf, L, dt))
    a, b = der[0].bind_radius, der[1].unbind_radius
    step = math.sqrt(2.0 * 2.0 * dt)
    kb = smd.numrxnrate(step, a, b) / dt
    kr = kb / (L ** 3 * target)
    # nC = target (n - nC)^2
    q = target
    nC = ((2 * n_total_a * q + 1) - math.sqrt((2 * n_total_a * q + 1) ** 2 - 4 * q * q * n_total_a ** 2)) / (2 * q)
    return kr, nC, kb


def test_reversible_equilibrium(orc, smd):
    """Mass-action steady state of A + B <-> C within 2% (16 replicas,
    sampled every 250 steps after 1000 steps from a near-equilibrium start)."""
    kr, want, kb = reversible_equilibrium(smd)
    assert abs(kb / 0.5 - 1.25) < 0.01  # kf / (1 - pgem)
    g = smd.gauss_table()
    got = []
    for k in range(16):
        pos = configs.uniform_box(np.random.default_rng(k), 750, 0.0, 10.0)
        types = np.array([1] * 250 + [2] * 250 + [3] * 250, np.int32)
        sim = orc.sim_for(smd, reversible_spec(kr, 4357 + k), pos, types)
        sim.step(g, 1000)
        for _ in range(20):
            sim.step(g, 250)
            got.append(1000 - sim.n)  # nA = nB = 500 - nC, so n = 1000 - nC
    mean = float(np.mean(got))
    assert abs(mean / want - 1.0) < 0.02, (mean, want)


def _creation_spec(means, k_decay=0.0, L=10.0, dt=1e-2, seed=4357):
    V = L ** 3
    rx = [dict(order=0, products=[1], rate=m / (dt * V)) for m in means]
    if k_decay:
        rx.append(dict(order=1, reactants=[1], products=[], rate=k_decay))
    return configs.make_spec(D=[0.0, 0.0], dt=dt, reactions=rx, box_lo=[0.0] * 3, box_hi=[L] * 3,
                             faces=[BC_REFLECT] * 6, grid_bits=(3, 3, 3), num_streams=64, seed=seed, volume=V)


@pytest.mark.parametrize("mean", [3.0, 30.0])
def test_zeroth_order_counts_are_poisson(orc, smd, mean):
    """Creations per step follow Poisson(k dt V): the multiplication method
    below a mean of 12 and the sum of pieces above it (chi-square over the
    count histogram, p > 0.001), at uniform positions inside the box."""
    from scipy import stats

    spec = _creation_spec([mean])
    der = smd.derive(spec)
    assert der[0].prob == np.float32(np.float32(spec.rxn[0].rate) * np.float32(spec.dt) * np.float32(spec.volume))
    steps = 3000
    sim = orc.sim_for(smd, spec, np.zeros((0, 4), np.float32), np.zeros(0, np.int32), cap=int(mean * steps * 2) + 64)
    g = smd.gauss_table()
    counts = []
    for _ in range(steps):
        n0 = sim.n
        sim.step(g, 1)
        counts.append(sim.n - n0)
    counts = np.array(counts)
    lo, hi = int(stats.poisson.ppf(0.001, mean)), int(stats.poisson.ppf(0.999, mean))
    edges = np.arange(lo, hi + 2)
    obs = np.array([(counts < lo).sum()] + [(counts == k).sum() for k in edges[:-1]] + [(counts > hi).sum()])
    p = np.concatenate([[stats.poisson.cdf(lo - 1, mean)], stats.poisson.pmf(edges[:-1], mean),
                        [stats.poisson.sf(hi, mean)]])
    keep = p * steps >= 5
    chi = stats.chisquare(obs[keep], p[keep] / p[keep].sum() * obs[keep].sum())
    assert chi.pvalue > 1e-3, (counts.mean(), counts.var(), chi)
    pos, t = sim.state()
    assert np.all(t == 1)
    # u = (v + 1/2) / 2^32 on the open interval, rounded to float: u = 1.0f
    # (x = hi) for the top 128 outputs, never u = 0
    assert np.all((pos[:, :3] > 0) & (pos[:, :3] <= 10.0)) and np.all(pos[:, 3] == 1.0)
    assert np.all(np.abs(pos[:, :3].mean(axis=0) - 5.0) < 0.1)


def test_zeroth_order_steady_state(orc, smd):
    """0 -> A (k0) with A -> 0 (k1): mass-action steady state. Decay acts
    before the step's creations, so n = (n + Poisson(k0 V dt)) thinned by
    exp(-k1 dt) gives mean k0 V dt / (1 - exp(-k1 dt)); within 2%."""
    lam, k1, dt = 5.0, 2.0, 1e-2
    spec = _creation_spec([lam], k_decay=k1, dt=dt)
    want = lam / (1.0 - math.exp(-k1 * dt))
    g = smd.gauss_table()
    got = []
    for rep in range(4):
        spec.seed = 4357 + rep
        sim = orc.sim_for(smd, spec, np.zeros((0, 4), np.float32), np.zeros(0, np.int32), cap=1 << 14)
        sim.step(g, 400)  # 8 relaxation times
        for _ in range(40):
            sim.step(g, 25)
            got.append(sim.n)
    assert abs(np.mean(got) / want - 1.0) < 0.02, (np.mean(got), want)
