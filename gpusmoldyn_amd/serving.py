"""Serving loop: reaction-diffusion jobs of different sizes arrive over time,
are admitted into free replica slots of one engine by HBM footprint, and are
stepped together (continuous batching) until each has done its own number of
steps; then its particles are copied out and its slot is freed.

A job admitted before engine step t has its particles injected by step t's
compaction and takes its first timestep in step t + 1, so a job of K steps
is complete after K + 1 engine steps.  Its result does not depend on the
other resident jobs: its buckets, random streams and compaction segment are
its own.

The reference has no serving mode (its host loop, particles.cpp, runs one
simulation per process); this module only drives the C-ABI entry points
smd_engine_finalize_slots / admit / retire / step / get.
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import BC_REFLECT, Engine, SmdError
from .configs import SEED, grid_bits_for, make_spec, streams_for

# Device bytes held per particle of capacity: positions and types double
# buffered, birth staging, sorted copy, candidates, keys and values double
# buffered.
BYTES_PER_PARTICLE = 2 * 16 + 2 * 4 + 16 + 4 + 16 + 4 + 4 * 4


@dataclass
class Job:
    jid: int
    arrival: float  # seconds after the start of the run
    spec: object
    pos: np.ndarray
    types: np.ndarray
    steps: int
    reserve: int  # particle slots reserved (upper bound of the job's count)

    @property
    def footprint(self):
        """HBM bytes the job holds while resident."""
        return self.reserve * BYTES_PER_PARTICLE


@dataclass
class JobResult:
    jid: int
    arrival: float
    admitted: float
    completed: float
    slot: int
    n0: int
    steps: int
    pos: np.ndarray = field(repr=False)
    types: np.ndarray = field(repr=False)

    @property
    def latency(self):
        return self.completed - self.arrival


def random_network(rng, n_species):
    """Random A + B <-> C / decay tables over 2-4 species (rates log-uniform)."""
    kf = float(math.exp(rng.uniform(math.log(1e-4), math.log(1e-2))))
    kr = float(math.exp(rng.uniform(math.log(1e-2), math.log(1.0))))
    kd = float(math.exp(rng.uniform(math.log(1e-3), math.log(1e-1))))
    if n_species == 2:
        # A + B -> B (B consumes A), B -> A (conversion back)
        D = [0.0, 1.0, 1.0]
        rx = [dict(order=2, reactants=[1, 2], products=[2], rate=kf),
              dict(order=1, reactants=[2], products=[1], rate=kr)]
    elif n_species == 3:
        D = [0.0, 1.0, 1.0, 0.5]
        rx = [dict(order=2, reactants=[1, 2], products=[3], rate=kf),
              dict(order=1, reactants=[3], products=[1, 2], rate=kr, pgem=0.2),
              dict(order=1, reactants=[3], products=[], rate=kd)]
    else:
        D = [0.0, 1.0, 1.0, 0.5, 0.25]
        rx = [dict(order=2, reactants=[1, 2], products=[3], rate=kf),
              dict(order=1, reactants=[3], products=[1, 2], rate=kr, pgem=0.2),
              dict(order=1, reactants=[3], products=[4], rate=kd)]
    return D, rx


def make_job(jid, arrival, n, steps, n_species, seed, dt=1e-3, density=2.048, per_stream=8):
    """A job of n particles at c1's density (2.048 per um^3) in a reflective
    cube, A and B in equal numbers, one random stream per `per_stream`
    particles (at most 32768 streams)."""
    rng = np.random.default_rng(seed)
    D, rx = random_network(rng, n_species)
    L = (n / density) ** (1.0 / 3.0)
    gb = grid_bits_for(n)
    spec = make_spec(D=D, dt=dt, reactions=rx, box_lo=[0.0] * 3, box_hi=[L] * 3, faces=[BC_REFLECT] * 6,
                     grid_bits=gb, num_streams=min(streams_for(n, per_stream), 32768), seed=seed,
                     unbind_offsets=1)
    pos = np.empty((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, L, size=(n, 3)).astype(np.float32)
    pos[:, 3] = 1.0
    types = np.where(np.arange(n) < n // 2, 1, 2).astype(np.int32)
    return Job(jid, arrival, spec, pos, types, steps, reserve=n)


def c5_jobs(n_jobs=200, rate=20.0, nmin=1000, nmax=262144, smin=100, smax=10000, seed=SEED, per_stream=2):
    """Poisson arrivals at `rate` jobs/s; n and steps log-uniform; 2-4 species.
    A few hundred thousand particles are resident at a time, so each job gets
    one stream per 2 particles: the producer's parallelism is its stream
    count, and at one per 8 it ran as few waves as the GPU has SIMDs."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=n_jobs)).astype(np.int64)
    st = np.exp(rng.uniform(math.log(smin), math.log(smax), size=n_jobs)).astype(np.int64)
    sp = rng.integers(2, 5, size=n_jobs)
    return [make_job(k, float(t[k]), int(ns[k]), int(st[k]), int(sp[k]), SEED + 1000 + k, per_stream=per_stream)
            for k in range(n_jobs)]


class Server:
    """One engine of `n_slots` replica slots.  `capacity` is the particle
    capacity of the engine; admission keeps the HBM footprint of the resident
    jobs within that of `capacity / 2` particles (the other half is the
    headroom the engine keeps for births within a step)."""

    def __init__(self, n_slots=32, slot_bucket_bits=18, slot_streams=32768, capacity=1 << 24,
                 inject_cap=1 << 21, max_chunk=64, device=0, graphs=True):
        self.e = Engine(device)
        self.e.finalize_slots(n_slots, slot_bucket_bits, slot_streams, capacity, inject_cap)
        self.e.set_graphs(graphs)
        self.n_slots = n_slots
        self.slot_bucket_bits = slot_bucket_bits
        self.slot_streams = slot_streams
        self.budget = (capacity // 2) * BYTES_PER_PARTICLE
        self.inject_cap = inject_cap
        self.max_chunk = max_chunk
        self.steps_done = 0

    def close(self):
        self.e.close()

    def fits(self, job):
        gb = sum(job.spec.grid_bits[d] for d in range(3))
        S = job.spec.num_streams or 32768
        return gb <= self.slot_bucket_bits and S <= self.slot_streams and job.footprint <= self.budget

    def run(self, jobs, realtime=True, log=None):
        """Serve `jobs` (sorted by arrival).  With realtime, a job is admitted
        no earlier than its arrival time after the start; otherwise all jobs
        are available at once.  Returns (results, stats)."""
        for j in jobs:
            if not self.fits(j):
                raise SmdError(f"job {j.jid} does not fit a slot")
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        free = list(range(self.n_slots - 1, -1, -1))
        resident = {}  # slot -> [job, steps left, admitted time]
        reserved = 0  # HBM bytes of the resident jobs
        results = []
        ps0 = self.e.particle_steps()
        t0 = time.perf_counter()
        while pending or resident:
            now = time.perf_counter() - t0
            injected = 0
            while (pending and free and (not realtime or pending[0].arrival <= now)
                   and reserved + pending[0].footprint <= self.budget
                   and injected + len(pending[0].types) <= self.inject_cap):
                j = pending.popleft()
                slot = free.pop()
                self.e.admit(slot, j.spec, j.pos, j.types)
                injected += len(j.types)
                reserved += j.footprint
                resident[slot] = [j, j.steps + 1, time.perf_counter() - t0]
            if not resident:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                continue
            k = min(min(v[1] for v in resident.values()), self.max_chunk)
            self.e.step(k)
            self.steps_done += k
            done = []
            for slot, v in resident.items():
                v[1] -= k
                if v[1] == 0:
                    done.append(slot)
            for slot in done:
                j, _, adm = resident.pop(slot)
                pos, typ = self.e.get(slot)
                self.e.retire(slot)
                t = time.perf_counter() - t0
                results.append(JobResult(j.jid, j.arrival, adm, t, slot, len(j.types), j.steps, pos, typ))
                reserved -= j.footprint
                free.append(slot)
                if log is not None:
                    log(f"job {j.jid} n={len(j.types)} steps={j.steps} latency={t - j.arrival:.3f}s")
        # the last retirements take effect at the next step
        self.e.sync()
        wall = time.perf_counter() - t0
        # every job's particles also enter its retirement step, which removes them
        ps = self.e.particle_steps() - ps0 - sum(len(r.types) for r in results)
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        stats = dict(jobs=len(results), wall_s=wall, particle_steps=ps, particle_steps_per_s=ps / wall,
                     p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                     engine_steps=self.steps_done)
        return results, stats
