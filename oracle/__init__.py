"""Python access to the CPU oracle (oracle.cpp) -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker.  The product package never imports this module.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SMD_ORACLE_LIB selects another build of the same source (the benchmark's
# CPU baseline uses one built with -march=native on the host that runs it)
LIB = os.environ.get("SMD_ORACLE_LIB") or os.path.join(HERE, "_build", "libsmdoracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsmdref.so")

MAXT, MAXR, MAXS = 16, 32, 4


class OrcNet(ctypes.Structure):
    _fields_ = [
        ("num_types", ctypes.c_int), ("num_rxn", ctypes.c_int),
        ("rate", ctypes.c_float * MAXT), ("D", ctypes.c_float * MAXT),
        ("uni_count", ctypes.c_int * MAXT), ("uni_start", ctypes.c_int * MAXT), ("uni_list", ctypes.c_int * MAXR),
        ("prob", ctypes.c_float * MAXR), ("bind_r2", ctypes.c_float * MAXR),
        ("prod_x", ctypes.c_int * MAXR), ("prod_y", ctypes.c_int * MAXR),
        ("prod_pos", (ctypes.c_float * 6) * MAXR),
        ("pair_rxn", ctypes.c_int * (MAXT * MAXT)),
        ("has_uni", ctypes.c_int), ("has_bim", ctypes.c_int),
        ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
        ("face", ctypes.c_int * 6), ("particle_radius", ctypes.c_float),
        ("num_spheres", ctypes.c_int), ("sphere", (ctypes.c_float * 4) * MAXS),
        ("unbind_offsets", ctypes.c_int),
        ("origin", ctypes.c_float * 3), ("cell", ctypes.c_float * 3),
        ("gbits", ctypes.c_uint * 3), ("num_streams", ctypes.c_uint),
        ("num_zero", ctypes.c_int), ("zero_list", ctypes.c_int * MAXR), ("zero_pieces", ctypes.c_int * MAXR),
        ("zero_g", ctypes.c_float * MAXR),
    ]


class OrcSim(ctypes.Structure):
    _fields_ = [
        ("net", OrcNet), ("n", ctypes.c_uint), ("cap", ctypes.c_uint),
        ("pos", ctypes.c_void_p), ("type", ctypes.c_void_p),
        ("mt", ctypes.c_void_p), ("par", ctypes.c_void_p),
        ("zmt", ctypes.c_void_p), ("zpar", ctypes.c_void_p),
    ]


_lib = None


def build(quiet=True):
    subprocess.run(["make", "-C", HERE, "-s"], check=True, capture_output=quiet)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.orc_mt_init.argtypes = [vp, ctypes.c_uint32]
        L.orc_mt_draw.argtypes = [vp, vp]
        L.orc_mt_draw.restype = ctypes.c_uint32
        L.orc_sincos.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.orc_theta.argtypes = [ctypes.c_float]
        L.orc_theta.restype = ctypes.c_float
        L.orc_step.argtypes = [ctypes.POINTER(OrcSim), vp, ctypes.c_int]
        L.orc_step_stage.argtypes = [ctypes.POINTER(OrcSim), vp, ctypes.c_int, vp, vp, vp,
                                     ctypes.POINTER(ctypes.c_uint)]
        L.orc_candidates_bruteforce.argtypes = [ctypes.POINTER(OrcNet), vp, ctypes.c_uint, vp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def net_from_spec(spec, derived, cell, origin):
    """Oracle network from a SimSpec and the derived reaction parameters
    (prob, bindRadius^2, product offsets).  The superstructure (CSR by
    reactant, pair table) is rebuilt here independently of the product."""
    n = OrcNet()
    T, R = spec.num_types, spec.num_reactions
    n.num_types, n.num_rxn = T, R
    dt = np.float32(spec.dt)
    for t in range(T):
        n.D[t] = spec.D[t]
        n.rate[t] = float(np.sqrt(np.float32(2.0) * np.float32(spec.D[t]) * dt))
    uni = [[r for r in range(R) if spec.rxn[r].order == 1 and spec.rxn[r].reactants[0] == t] for t in range(T)]
    k = 0
    for t in range(T):
        n.uni_start[t] = k
        n.uni_count[t] = len(uni[t])
        for r in uni[t]:
            n.uni_list[k] = r
            k += 1
    for i in range(MAXT * MAXT):
        n.pair_rxn[i] = -1
    for r in range(R):
        x = spec.rxn[r]
        n.prob[r] = derived[r].prob
        n.bind_r2[r] = derived[r].bind_radius2
        n.prod_x[r] = x.products[0] if x.nprod >= 1 else -1
        n.prod_y[r] = x.products[1] if x.nprod >= 2 else -1
        for q in range(6):
            n.prod_pos[r][q] = derived[r].product_pos[q]
        if x.order == 1:
            n.has_uni = 1
        if x.order == 2:
            n.has_bim = 1
            n.pair_rxn[x.reactants[0] * MAXT + x.reactants[1]] = r
            n.pair_rxn[x.reactants[1] * MAXT + x.reactants[0]] = r
    for d in range(3):
        n.lo[d], n.hi[d] = spec.box_lo[d], spec.box_hi[d]
        n.origin[d], n.cell[d] = float(origin[d]), float(cell[d])
        n.gbits[d] = spec.grid_bits[d]
    for f in range(6):
        n.face[f] = spec.face[f]
    n.particle_radius = spec.particle_radius
    n.num_spheres = spec.num_spheres
    for s in range(spec.num_spheres):
        for q in range(4):
            n.sphere[s][q] = spec.sphere[s][q]
    n.unbind_offsets = spec.unbind_offsets
    n.num_streams = spec.num_streams or 32768
    # order-0: mean = k dt V per step (smolreact.cpp:295-303), V the spec's
    # volume or the box volume; Poisson by the multiplication method
    # (stats.cpp:234-253), means of 12 and above split into pieces below 12
    vol = np.float32(spec.volume)
    if not vol > 0:
        ext = [np.float32(spec.box_hi[d]) - np.float32(spec.box_lo[d]) for d in range(3)]
        vol = ext[0] * ext[1] * ext[2]
    k = 0
    for r in range(R):
        x = spec.rxn[r]
        if x.order != 0 or x.nprod < 1:
            continue
        mean = np.float32(x.rate) * dt * vol
        if not mean > 0:
            continue
        pieces = 1 if mean < np.float32(12.0) else int(np.ceil(mean / np.float32(8.0)))
        xm = mean if pieces == 1 else mean / np.float32(pieces)
        n.zero_list[k] = r
        n.zero_pieces[k] = pieces
        n.zero_g[k] = float(np.float32(math.exp(-float(xm))))
        n.prob[r] = float(mean)
        k += 1
    n.num_zero = k
    return n


class Sim:
    """One simulation on the CPU oracle."""

    def __init__(self, net, pos4, types, seeds, params, cap=None):
        L = lib()
        n = len(types)
        cap = cap or max(2 * n + 16, 64)
        self.pos = np.zeros((cap, 4), np.float32)
        self.type = np.zeros(cap, np.int32)
        self.pos[:n] = pos4
        self.type[:n] = types
        S = net.num_streams
        self.mt = np.zeros((S, 20), np.uint32)
        for k in range(S):
            L.orc_mt_init(_p(self.mt[k]), int(seeds[k]))
        self.par = np.ascontiguousarray(params, np.uint32).reshape(S, 3)
        self.s = OrcSim()
        self.s.net = net
        self.s.n, self.s.cap = n, cap
        self.s.pos, self.s.type = _p(self.pos), _p(self.type)
        self.s.mt, self.s.par = _p(self.mt), _p(self.par)
        self.zmt = np.zeros(20, np.uint32)
        self.zpar = np.zeros(3, np.uint32)
        self.s.zmt, self.s.zpar = _p(self.zmt), _p(self.zpar)

    def set_creation_stream(self, seed, params):
        """Order-0 creation stream: seed and 3 parameter words."""
        lib().orc_mt_init(_p(self.zmt), int(seed))
        self.zpar[:] = np.asarray(params, np.uint32)[:3]

    @property
    def n(self):
        return self.s.n

    def step(self, gauss, nsteps=1):
        rc = lib().orc_step(ctypes.byref(self.s), _p(gauss), nsteps)
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")

    def stage(self, gauss, stage):
        n = self.s.n
        keys = np.zeros(n, np.uint32)
        vals = np.zeros(n, np.uint32)
        cand = np.zeros(max(n, 1), np.int32)
        nl = ctypes.c_uint(0)
        lib().orc_step_stage(ctypes.byref(self.s), _p(gauss), stage, _p(keys), _p(vals), _p(cand), ctypes.byref(nl))
        return dict(pos=self.pos[:n].copy(), type=self.type[:n].copy(), keys=keys, vals=vals,
                    cand=cand[: nl.value].copy(), n_live=nl.value)

    def state(self):
        n = self.s.n
        return self.pos[:n].copy(), self.type[:n].copy()


def stream_params(S, dat_records):
    """(S, 3) parameters: stream k uses record k % 32768."""
    idx = np.arange(S) % 32768
    return np.ascontiguousarray(dat_records[idx, :3], np.uint32)


def load_dat(path):
    return np.fromfile(path, dtype="<u4").reshape(-1, 4)


def sim_for(smd, spec, pos, types, cap=None):
    """Oracle simulation matching what the GPU engine builds for `spec`."""
    der = smd.derive(spec)
    cell, origin = smd.spec_cells(spec)
    net = net_from_spec(spec, der, cell, origin)
    S = net.num_streams
    seeds = smd.stream_seeds(spec.seed, S + 1)
    dat = load_dat(os.path.join(os.path.dirname(smd.__file__), "data", "MersenneTwister.dat"))
    sim = Sim(net, pos, types, seeds[:S], stream_params(S, dat), cap=cap)
    if net.num_zero:
        # the stream after the S particle streams: seed S, record S mod 32768
        sim.set_creation_stream(seeds[S], dat[S % 32768, :3])
    return sim
