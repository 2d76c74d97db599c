"""Synthetic workloads (species, reactions, geometry, initial particles).

All random inputs come from numpy's PCG64 seeded with 4357 (replica k of an
ensemble: 4357 + k); device RNG streams are seeded from the dcmt sequence of
each spec's `seed`.

  plain_c1   A + B -> C, 8192 A + 8192 B in a reflective 20 um cube
  c2         1,048,576 particles, A + B <-> C and C -> 0, periodic 100 um cube
  c3         4,194,304 particles in 64 Gaussian blobs, A + B <-> C, reflective
             sphere r = 60 um, absorbing plane z = 0, 200 um cube
  (c2, c3 and c4 enter in hash-bucket order, as a lattice initialisation would)
  c4         512 replicas of c1's network, n log-uniform in [2048, 65536],
             kf log-uniform in [1e-4, 1e-2]
  ref_test1  the reference's Lotka-Volterra scenario (test1)
  ref_test2  the reference's Michaelis-Menten scenario (test2)
"""
import math

import numpy as np

from . import BC_ABSORB, BC_CLAMP, BC_PERIODIC, BC_REFLECT, SimSpec

SEED = 4357
A, B, C = 1, 2, 3


def make_spec(D, dt, reactions, box_lo, box_hi, faces, grid_bits, cell_size=None, origin=None, num_streams=0,
              seed=SEED, spheres=(), unbind_offsets=0, particle_radius=0.0, volume=None):
    """Build a SimSpec.  D includes the dummy type 0; reactions are dicts with
    order, reactants, products, rate and optional pgem/prob."""
    s = SimSpec()
    s.num_types = len(D)
    for k, d in enumerate(D):
        s.D[k] = d
    s.dt = dt
    L = [box_hi[d] - box_lo[d] for d in range(3)]
    s.volume = volume if volume is not None else L[0] * L[1] * L[2]
    s.num_reactions = len(reactions)
    for k, r in enumerate(reactions):
        x = s.rxn[k]
        x.order = r["order"]
        rc = list(r.get("reactants", [])) + [-1, -1]
        pr = list(r.get("products", [])) + [-1, -1]
        x.reactants[0], x.reactants[1] = rc[0], rc[1]
        x.products[0], x.products[1] = pr[0], pr[1]
        x.nprod = len(r.get("products", []))
        x.rate = r["rate"]
        x.pgem = r.get("pgem", 0.0)
        x.prob = r.get("prob", 0.0)
    for d in range(3):
        s.box_lo[d] = box_lo[d]
        s.box_hi[d] = box_hi[d]
        s.grid_bits[d] = grid_bits[d]
        s.cell_size[d] = cell_size[d] if cell_size is not None else 0.0
        s.origin[d] = origin[d] if origin is not None else box_lo[d]
    for f in range(6):
        s.face[f] = faces[f]
    s.particle_radius = particle_radius
    s.num_spheres = len(spheres)
    for k, sp in enumerate(spheres):
        for q in range(4):
            s.sphere[k][q] = sp[q]
    s.unbind_offsets = unbind_offsets
    s.num_streams = num_streams
    s.seed = seed
    return s


def cell_order(spec, pos, types):
    """The particles reordered by hash-grid bucket (stable), the order a
    lattice initialisation (ParticleSystem::initGrid) produces: the step's
    gathers into cell order then read nearby memory."""
    from . import spec_cells

    cell, origin = spec_cells(spec)
    g = np.floor((pos[:, :3] - np.asarray(origin, np.float32)) / np.asarray(cell, np.float32)).astype(np.int64)
    bits = [int(spec.grid_bits[d]) for d in range(3)]
    key = ((g[:, 2] & ((1 << bits[2]) - 1)) << (bits[0] + bits[1])) | ((g[:, 1] & ((1 << bits[1]) - 1)) << bits[0]) \
        | (g[:, 0] & ((1 << bits[0]) - 1))
    order = np.argsort(key, kind="stable")
    return np.ascontiguousarray(pos[order]), np.ascontiguousarray(types[order])


def uniform_box(rng, n, lo, hi):
    p = np.empty((n, 4), np.float32)
    p[:, :3] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    p[:, 3] = 1.0
    return p


def c1_spec(kf=1e-3, n_streams=0, seed=SEED, grid_bits=(6, 6, 6)):
    return make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=1e-3,
        reactions=[dict(order=2, reactants=[A, B], products=[C], rate=kf)],
        box_lo=[0.0] * 3, box_hi=[20.0] * 3, faces=[BC_REFLECT] * 6, grid_bits=grid_bits,
        num_streams=n_streams, seed=seed)


def c1(nA=8192, nB=8192, kf=1e-3, n_streams=0, seed=SEED, rng_seed=SEED, grid_bits=(6, 6, 6)):
    rng = np.random.default_rng(rng_seed)
    pos = uniform_box(rng, nA + nB, 0.0, 20.0)
    types = np.array([A] * nA + [B] * nB, np.int32)
    return c1_spec(kf, n_streams, seed, grid_bits), pos, types


C1_STEPS = 1000


def c2_spec(n_streams=131072, grid_bits=(7, 7, 7), L=100.0, dt=1e-3, seed=SEED):
    return make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=dt,
        reactions=[
            dict(order=2, reactants=[A, B], products=[C], rate=5e-3),
            dict(order=1, reactants=[C], products=[A, B], rate=0.1, pgem=0.2),
            dict(order=1, reactants=[C], products=[], rate=0.01),
        ],
        box_lo=[0.0] * 3, box_hi=[L] * 3, faces=[BC_PERIODIC] * 6, grid_bits=grid_bits,
        num_streams=n_streams, seed=seed, unbind_offsets=1)


def c2(n=1 << 20, n_streams=131072, grid_bits=(7, 7, 7), L=100.0, dt=1e-3, seed=SEED, ordered=True):
    rng = np.random.default_rng(SEED)
    pos = uniform_box(rng, n, 0.0, L)
    types = np.empty(n, np.int32)
    types[: n // 2] = A
    types[n // 2:] = B
    spec = c2_spec(n_streams, grid_bits, L, dt, seed)
    if ordered:
        pos, types = cell_order(spec, pos, types)
    return spec, pos, types


C2_STEPS = 500


def c3_spec(n_streams=262144, grid_bits=(8, 8, 7), cell=0.1, L=200.0, seed=SEED):
    return make_spec(
        D=[0.0, 1.0, 1.0, 0.5], dt=1e-3,
        reactions=[
            dict(order=2, reactants=[A, B], products=[C], rate=5e-3),
            dict(order=1, reactants=[C], products=[A, B], rate=0.1, pgem=0.2),
        ],
        box_lo=[0.0] * 3, box_hi=[L] * 3,
        faces=[BC_REFLECT, BC_REFLECT, BC_REFLECT, BC_REFLECT, BC_ABSORB, BC_REFLECT],
        grid_bits=grid_bits, cell_size=[cell] * 3, num_streams=n_streams, seed=seed,
        spheres=[(L / 2, L / 2, L / 2, 60.0)], unbind_offsets=1)


def c3(n=1 << 22, n_blobs=64, sigma=2.0, n_streams=262144, grid_bits=(8, 8, 7), cell=0.1, L=200.0, seed=SEED,
       ordered=True):
    rng = np.random.default_rng(SEED)
    centres = rng.uniform(0.0, L, size=(n_blobs, 3))
    which = rng.integers(0, n_blobs, size=n)
    p = centres[which] + sigma * rng.standard_normal((n, 3))
    # mirror the few points outside the box back in
    p = np.where(p < 0.0, -p, p)
    p = np.where(p > L, 2 * L - p, p)
    p = np.clip(p, 1e-4, L - 1e-4)
    pos = np.empty((n, 4), np.float32)
    pos[:, :3] = p.astype(np.float32)
    pos[:, 3] = 1.0
    types = (np.arange(n) % 3 + 1).astype(np.int32)
    spec = c3_spec(n_streams, grid_bits, cell, L, seed)
    if ordered:
        pos, types = cell_order(spec, pos, types)
    return spec, pos, types


C3_STEPS = 200


def grid_bits_for(n):
    """Hash grid of about n buckets, at least 4 per axis."""
    tb = max(6, int(round(math.log2(max(n, 64)))))
    b = [tb // 3] * 3
    for k in range(tb - 3 * (tb // 3)):
        b[k] += 1
    return tuple(max(2, x) for x in b)


def streams_for(n, per=8):
    return max(64, int(math.ceil(n / per)))


def c4(n_rep=512, nmin=2048, nmax=65536, kmin=1e-4, kmax=1e-2, per_stream=8, ordered=True):
    """List of (spec, pos, types) for the batched ensemble."""
    rng = np.random.default_rng(SEED)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=n_rep)).astype(np.int64)
    ks = np.exp(rng.uniform(math.log(kmin), math.log(kmax), size=n_rep))
    out = []
    for i in range(n_rep):
        n = int(ns[i])
        nA = n // 2
        spec, pos, types = c1(nA=nA, nB=n - nA, kf=float(ks[i]), n_streams=streams_for(n, per_stream),
                              seed=SEED + i, rng_seed=SEED + i, grid_bits=grid_bits_for(n))
        if ordered:
            pos, types = cell_order(spec, pos, types)
        out.append((spec, pos, types))
    return out


C4_STEPS = 1000


# The reference's hard-coded scenarios (particleSystem.cpp:319-571), in its
# world [-1, 1]^3 with the clamp boundary (particles_kernel.cu:60-95),
# particle radius 1/64 and a 64^3 grid of 1/32 cells (particleSystem.cpp:59-85).
def _reference_world(D, dt, volume, reactions, unbind_offsets=0, grid_bits=(6, 6, 6), num_streams=32768,
                     cell=2.0 / 64.0):
    r = 1.0 / 64.0
    return make_spec(D=D, dt=dt, reactions=reactions, box_lo=[-1.0] * 3, box_hi=[1.0] * 3,
                     faces=[BC_CLAMP] * 6, grid_bits=grid_bits, cell_size=[cell] * 3, origin=[-1.0] * 3,
                     num_streams=num_streams, particle_radius=r, volume=volume, unbind_offsets=unbind_offsets)


def ref_test1(n=4096, k2=5e-3, rng_seed=SEED):
    """Lotka-Volterra (test1Reactions/test1Types, particleSystem.cpp:319-405):
    A -> 2A at 10/s, A + B -> 2B, B -> 0 at 10/s; D = 0.001, dt = 0.1,
    volume 10; the first half A, the rest B; products of a split at the
    reaction point (the reference's unbindRadius = 0).  The reference's
    k2 = 2000 maps to a binding radius of 8e4 (larger than the world); k2
    defaults to 5e-3, a radius of 0.21, on an 8^3 grid of 0.25 cells."""
    spec = _reference_world(
        D=[0.0, 0.001, 0.001], dt=0.1, volume=10.0,
        reactions=[dict(order=1, reactants=[A], products=[A, A], rate=10.0),
                   dict(order=2, reactants=[A, B], products=[B, B], rate=k2),
                   dict(order=1, reactants=[B], products=[], rate=10.0)],
        grid_bits=(3, 3, 3), cell=0.25)
    rng = np.random.default_rng(rng_seed)
    pos = uniform_box(rng, n, -1.0, 1.0)
    types = np.where(np.arange(n) < n // 2, A, B).astype(np.int32)
    return spec, pos, types


def ref_test2(n=11000, rng_seed=SEED):
    """Michaelis-Menten (test2Reactions/test2Types, particleSystem.cpp:408-511):
    E + S -> ES at 0.01, ES -> S + S at 1 (pgem 0.2; the reference writes
    products (2, 2) where its comment says E + S), ES -> E + P at 1;
    D = 0.05 for all four species, dt = 1e-3, volume 90; the first n/11
    particles E, the rest S."""
    spec = _reference_world(
        D=[0.0, 0.05, 0.05, 0.05, 0.05], dt=1e-3, volume=90.0,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=0.01),
                   dict(order=1, reactants=[3], products=[2, 2], rate=1.0, pgem=0.2),
                   dict(order=1, reactants=[3], products=[1, 4], rate=1.0)])
    rng = np.random.default_rng(rng_seed)
    pos = uniform_box(rng, n, -1.0, 1.0)
    types = np.where(np.arange(n) < n // 11, 1, 2).astype(np.int32)
    return spec, pos, types
