"""The reference's extern "C" layer (include/gpusmoldyn_legacy.h) driven the
way ParticleSystem::update() drives it (particleSystem.cpp:829-1001), and the
ParticleSystem class itself, checked bit for bit against the engine."""
import ctypes
import os

import numpy as np
import pytest

from gpusmoldyn_amd import BC_CLAMP, configs

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


class SimParams(ctypes.Structure):
    _fields_ = [("colliderPos", ctypes.c_float * 3), ("colliderRadius", ctypes.c_float),
                ("gravity", ctypes.c_float * 3), ("globalDamping", ctypes.c_float),
                ("particleRadius", ctypes.c_float), ("gridSize", ctypes.c_uint * 3), ("numCells", ctypes.c_uint),
                ("worldOrigin", ctypes.c_float * 3), ("cellSize", ctypes.c_float * 3),
                ("numBodies", ctypes.c_uint), ("maxParticlesPerCell", ctypes.c_uint),
                ("spring", ctypes.c_float), ("damping", ctypes.c_float), ("shear", ctypes.c_float),
                ("attraction", ctypes.c_float), ("boundaryDamping", ctypes.c_float)]


class Reaction(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("order", "bindRadiusSquared", "unbindRadius", "geminateProbability", "prob",
                                  "rate", "nprod", "products", "reactants", "productPositions")]


def legacy_spec(n_types_D, reactions, grid_bits=6):
    return configs.make_spec(D=n_types_D, dt=0.1, reactions=reactions, box_lo=[-1.0] * 3, box_hi=[1.0] * 3,
                             faces=[BC_CLAMP] * 6, grid_bits=(grid_bits,) * 3, cell_size=[1.0 / 32] * 3,
                             origin=[-1.0] * 3, num_streams=32768, particle_radius=1.0 / 64.0)


def network_test1_like():
    # particleSystem.cpp:319-406 (A -> 2A, A + B -> 2B, B -> 0) with rates that
    # keep the population bounded over the test
    reactions = [dict(order=1, reactants=[1], products=[1, 1], rate=0.4),
                 dict(order=2, reactants=[1, 2], products=[2, 2], rate=2e-4),
                 dict(order=1, reactants=[2], products=[], rate=0.3)]
    spec = legacy_spec([0.0, 0.001, 0.001], reactions)
    rng = np.random.default_rng(31)
    n = 20000
    pos = configs.uniform_box(rng, n, -1.0, 1.0)
    types = np.where(np.arange(n) < n // 2, 1, 2).astype(np.int32)
    return spec, pos, types


class LegacyDriver:
    """ParticleSystem::_initialize + update() written against the C layer."""

    def __init__(self, L, smd, spec, pos, types, cap):
        self.L = L
        self.spec = spec
        self.cap = cap
        self.n = len(types)
        T = spec.num_types
        R = spec.num_reactions
        assert L.cudaInit(0, None) == 0
        self.dev = {}

        def alloc(name, nbytes):
            p = vp()
            assert L.allocateArray(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 4))) == 0
            self.dev[name] = p
            return p

        def put(name, arr):
            arr = np.ascontiguousarray(arr)
            p = alloc(name, arr.nbytes)
            assert L.copyArrayToDevice(p, arr.ctypes.data_as(vp), 0, arr.nbytes) == 0
            return p

        for name, b in (("pos", 16), ("types", 4), ("addedPos", 16), ("addedTypes", 4), ("birth", 4),
                        ("death", 4), ("scan", 4), ("sortedPos", 16), ("hash", 4), ("index", 4), ("tk", 4),
                        ("tv", 4)):
            alloc(name, b * cap)
        self.upload(pos, types)
        D = np.array([spec.D[t] for t in range(T)], np.float32)
        rates = np.sqrt(np.float32(2.0) * D * np.float32(spec.dt)).astype(np.float32)
        put("D", D)
        put("rates", rates)
        der = smd.derive(spec)
        order = np.array([spec.rxn[r].order for r in range(R)], np.int32)
        prods = np.array([[spec.rxn[r].products[0] if spec.rxn[r].nprod >= 1 else -1,
                           spec.rxn[r].products[1] if spec.rxn[r].nprod >= 2 else -1] for r in range(R)], np.int32)
        reacts = np.array([[spec.rxn[r].reactants[0], spec.rxn[r].reactants[1]] for r in range(R)], np.int32)
        rl = Reaction()
        rl.order = put("r_order", order)
        rl.bindRadiusSquared = put("r_br2", np.array([d.bind_radius2 for d in der], np.float32))
        rl.unbindRadius = put("r_ub", np.array([d.unbind_radius for d in der], np.float32))
        rl.geminateProbability = put("r_gem", np.full(R, 0.2, np.float32))
        rl.prob = put("r_prob", np.array([d.prob for d in der], np.float32))
        rl.rate = put("r_rate", np.array([spec.rxn[r].rate for r in range(R)], np.float32))
        rl.nprod = put("r_nprod", np.array([spec.rxn[r].nprod for r in range(R)], np.int32))
        rl.products = put("r_prod", prods)
        rl.reactants = put("r_react", reacts)
        rl.productPositions = put("r_pp", np.array([list(d.product_pos) for d in der], np.float32))
        rlb = np.frombuffer(bytes(rl), np.uint8).copy()
        self.rl = put("rl", rlb)
        # order-1 superstructure (fillReactionOrderInfo) and the order-2 pair table
        uni = [[r for r in range(R) if order[r] == 1 and reacts[r, 0] == t] for t in range(T)]
        nrxn = np.array([len(u) for u in uni], np.int32)
        idx = np.concatenate([[0], np.cumsum(nrxn)[:-1]]).astype(np.int32)
        l2g = np.array([r for u in uni for r in u] or [0], np.int32)
        self.nrx1 = int((order == 1).sum())
        self.nrx2 = int((order == 2).sum())
        put("nrxn", nrxn)
        put("idx", idx)
        put("l2g", l2g)
        table = -np.ones((T, T), np.int32)
        for r in range(R):
            if order[r] == 2:
                table[reacts[r, 0], reacts[r, 1]] = r
                table[reacts[r, 1], reacts[r, 0]] = r
        put("table", table)
        self.numTypes = T
        self.params = SimParams()
        self.params.particleRadius = spec.particle_radius
        for d in range(3):
            self.params.gridSize[d] = 1 << spec.grid_bits[d]
            self.params.worldOrigin[d] = spec.origin[d]
            self.params.cellSize[d] = spec.cell_size[d]
        self.numCells = self.params.gridSize[0] * self.params.gridSize[1] * self.params.gridSize[2]
        self.params.numCells = self.numCells
        alloc("cellStart", 4 * self.numCells)
        alloc("cellEnd", 4 * self.numCells)
        alloc("rng", 84 * 32768)
        L.InitializeRandomNumbers()
        assert L.initializeMT(self.dev["rng"]) == 0
        g = smd.gauss_table()
        assert L.copyGaussianToConstant(g.ctypes.data_as(vp), 4096) == 0
        assert L.setParameters(ctypes.byref(self.params)) == 0
        self.gridSortBits = sum(spec.grid_bits[d] for d in range(3))

    def upload(self, pos, types):
        pos = np.ascontiguousarray(pos, np.float32)
        types = np.ascontiguousarray(types, np.int32)
        self.L.copyArrayToDevice(self.dev["pos"], pos.ctypes.data_as(vp), 0, pos.nbytes)
        self.L.copyArrayToDevice(self.dev["types"], types.ctypes.data_as(vp), 0, types.nbytes)
        self.n = len(types)

    def download(self, n=None):
        n = self.n if n is None else n
        pos = np.zeros((n, 4), np.float32)
        types = np.zeros(n, np.int32)
        if n:
            self.L.copyArrayFromDevice(pos.ctypes.data_as(vp), self.dev["pos"], 0, pos.nbytes)
            self.L.copyArrayFromDevice(types.ctypes.data_as(vp), self.dev["types"], 0, types.nbytes)
        return pos, types

    def update(self):
        L, d, n = self.L, self.dev, self.n
        assert L.setParameters(ctypes.byref(self.params)) == 0
        assert L.integrateSystem(self.rl, d["l2g"], d["idx"], d["nrxn"], self.nrx1, d["pos"], d["types"],
                                 d["addedPos"], d["addedTypes"], d["birth"], d["death"], d["rates"], d["rng"],
                                 ctypes.c_float(self.spec.dt), n, self.numTypes) == 0
        left, added = n, 0
        if self.nrx1 > 0:
            assert L.smdExclusiveScanInt(d["scan"], d["birth"], n) == 0
            added = L.compactPosAndType(d["scan"], d["birth"], d["addedPos"], d["addedTypes"], n)
            assert L.smdExclusiveScanInt(d["scan"], d["death"], n) == 0
            left = L.compactPosAndType(d["scan"], d["death"], d["pos"], d["types"], n)
            assert added >= 0 and left >= 0
        assert L.adjustPositions(d["pos"], d["types"], left, ctypes.c_float(self.spec.particle_radius)) == 0
        total = left
        if self.nrx2 > 0 and left > 0:
            assert L.calcHash(d["hash"], d["index"], d["pos"], left) == 0
            assert L.radixSort(d["hash"], d["index"], d["tk"], d["tv"], None, None, None, None, left,
                               self.gridSortBits, False) == 0
            assert L.reorderDataAndFindCellStart(d["cellStart"], d["cellEnd"], d["sortedPos"], d["hash"],
                                                 d["index"], d["pos"], left, self.numCells) == 0
            minus1 = np.full(left, -1, np.int32)
            L.copyArrayToDevice(d["death"], minus1.ctypes.data_as(vp), 0, minus1.nbytes)
            assert L.collide(d["pos"], d["sortedPos"], d["index"], d["types"], d["death"], d["D"], d["table"],
                             self.rl, d["cellStart"], d["cellEnd"], left, self.numCells, self.numTypes,
                             d["rng"]) == 0
            assert L.smdExclusiveScanInt(d["scan"], d["death"], left) == 0
            total = L.compactPosAndType(d["scan"], d["death"], d["pos"], d["types"], left)
            assert total >= 0
        # append the unimolecular births (particleSystem.cpp:985-986)
        if added:
            ap = np.zeros((added, 4), np.float32)
            at = np.zeros(added, np.int32)
            L.copyArrayFromDevice(ap.ctypes.data_as(vp), d["addedPos"], 0, ap.nbytes)
            L.copyArrayFromDevice(at.ctypes.data_as(vp), d["addedTypes"], 0, at.nbytes)
            L.copyArrayToDevice(d["pos"], ap.ctypes.data_as(vp), total * 16, ap.nbytes)
            L.copyArrayToDevice(d["types"], at.ctypes.data_as(vp), total * 4, at.nbytes)
        self.n = total + added
        assert self.n <= self.cap
        return self.n > 0

    def close(self):
        for p in self.dev.values():
            self.L.freeArray(p)


def _lib(smd):
    L = smd.lib()
    L.compactPosAndType.restype = ctypes.c_int
    L.mapGLBufferObject.restype = vp
    L.allocateArray.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.integrateSystem.argtypes = [vp] * 4 + [ctypes.c_int] + [vp] * 8 + [ctypes.c_float, ctypes.c_uint,
                                                                       ctypes.c_uint]
    L.adjustPositions.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_float]
    L.radixSort.argtypes = [vp] * 8 + [ctypes.c_uint, ctypes.c_uint, ctypes.c_bool]
    L.collide.argtypes = [vp] * 10 + [ctypes.c_uint] * 3 + [vp]
    return L


def test_legacy_update_loop_matches_engine(smd, gpu):
    L = _lib(smd)
    spec, pos, types = network_test1_like()
    print("\n[gpu-test] legacy update loop, test1-like network, 20000 particles", flush=True)
    drv = LegacyDriver(L, smd, spec, pos, types, cap=1 << 18)
    e = smd.Engine(0)
    e.add_replica(spec, pos, types)
    e.finalize()
    for k in range(12):
        drv.update()
        e.step(1)
        lp, lt = drv.download()
        ep, et = e.get(0)
        assert len(lt) == len(et), f"count at step {k}"
        assert np.array_equal(lt, et), f"types at step {k}"
        assert np.array_equal(lp.view(np.uint32), ep.view(np.uint32)), f"positions at step {k}"
    assert (lt == 2).sum() > 0
    drv.close()


def test_legacy_gl_entry_points_are_unsupported(smd, gpu):
    L = _lib(smd)
    assert L.cudaGLInit(0, None) == 7
    assert L.registerGLBufferObject(1) == 7
    assert L.mapGLBufferObject(1) is None
    host = np.zeros(4, np.float32)
    assert L.copyArrayFromDevice(host.ctypes.data_as(vp), None, 5, 16) == 7


def test_legacy_scan_and_sort(smd, gpu):
    L = _lib(smd)
    assert L.cudaInit(0, None) == 0
    rng = np.random.default_rng(41)
    for n in (1, 1000, 1024, 5000, 1 << 20):
        flags = rng.integers(0, 2, size=n).astype(np.int32)
        dp = vp()
        assert L.allocateArray(ctypes.byref(dp), n * 4) == 0
        L.copyArrayToDevice(dp, flags.ctypes.data_as(vp), 0, flags.nbytes)
        assert L.smdExclusiveScanInt(dp, dp, n) == 0
        out = np.zeros(n, np.int32)
        L.copyArrayFromDevice(out.ctypes.data_as(vp), dp, 0, out.nbytes)
        ref = np.concatenate([[0], np.cumsum(flags)[:-1]]).astype(np.int32)
        assert np.array_equal(out, ref)
        L.freeArray(dp)


def _reference_spec():
    # the reference constructor's scenario (test3Types/test3Reactions,
    # particleSystem.cpp:513-571) in its world (particleSystem.cpp:59-85)
    return configs.make_spec(D=[0.0, 0.3, 0.1], dt=0.01, reactions=[], box_lo=[-1.0] * 3, box_hi=[1.0] * 3,
                             faces=[BC_CLAMP] * 6, grid_bits=(6, 6, 6), cell_size=[1 / 32.0] * 3,
                             origin=[-1.0] * 3, num_streams=32768, particle_radius=1 / 64.0, volume=90.0)


def test_particle_system_reference_ctor(smd, orc, gpu):
    """ParticleSystem(numParticles, gridSize, false): the reference's active
    scenario (test3), CONFIG_GRID reset, then update() -- against the CPU
    oracle started from the same state."""
    L = smd.lib()
    L.smd_ps_create_reference.restype = vp
    L.smd_ps_create_reference.argtypes = [ctypes.c_uint] * 4
    for f in ("smd_ps_update", "smd_ps_reset", "smd_ps_num_particles", "smd_ps_destroy"):
        getattr(L, f).argtypes = [vp] + ([ctypes.c_int] if f == "smd_ps_reset" else [])
    L.smd_ps_get_positions.argtypes = [vp, vp, ctypes.c_uint]
    L.smd_ps_get_types.argtypes = [vp, vp, ctypes.c_uint]
    n = 16384
    print("\n[gpu-test] ParticleSystem reference constructor", flush=True)
    ps = L.smd_ps_create_reference(n, 64, 64, 64)
    assert ps
    assert L.smd_ps_reset(ps, 1) == 0
    p0 = np.zeros((n, 4), np.float32)
    t0 = np.zeros(n, np.int32)
    L.smd_ps_get_positions(ps, p0.ctypes.data_as(vp), n)
    L.smd_ps_get_types(ps, t0.ctypes.data_as(vp), n)
    assert (t0 == 1).sum() == 3 * (n // 4)
    # initGrid (particleSystem.cpp:1088-1117): s = ceil(n^(1/3)) points per
    # axis at spacing 2 / (s - 2) from -1 + r, jittered by at most r / 100
    # (the lattice runs past 1; update() clamps it)
    r = np.float32(1 / 64.0)
    s = int(np.ceil(np.float32(n) ** np.float32(1.0 / 3.0)))
    i = np.arange(n)
    lat = np.stack([i % s, (i // s) % s, i // (s * s)], axis=1).astype(np.float32)
    want = lat * np.float32(2.0 / (s - 2)) + r - np.float32(1.0)
    assert np.all(p0[:, 3] == 1.0)
    assert np.abs(p0[:, :3] - want).max() <= r * np.float32(0.01) * 1.001
    sim = orc.sim_for(smd, _reference_spec(), p0, t0)
    for _ in range(20):
        assert L.smd_ps_update(ps) == 1
    sim.step(smd.gauss_table(), 20)
    p1 = np.zeros((n, 4), np.float32)
    L.smd_ps_get_positions(ps, p1.ctypes.data_as(vp), n)
    rp, rt = sim.state()
    assert np.array_equal(p1.view(np.uint32), rp.view(np.uint32))
    r = np.float32(1 / 64.0)
    assert p1[:, :3].min() >= np.float32(-1.0) + r and p1[:, :3].max() <= np.float32(1.0) - r
    L.smd_ps_destroy(ps)


def test_particle_system_update_false_on_extinction(smd, gpu):
    L = smd.lib()
    L.smd_ps_create.restype = vp
    L.smd_ps_create.argtypes = [ctypes.POINTER(smd.SimSpec), vp, vp, ctypes.c_uint, ctypes.c_int]
    L.smd_ps_update.argtypes = [vp]
    L.smd_ps_destroy.argtypes = [vp]
    spec = legacy_spec([0.0, 0.001], [dict(order=1, reactants=[1], products=[], rate=50.0)])
    pos = configs.uniform_box(np.random.default_rng(5), 100, -0.9, 0.9)
    types = np.ones(100, np.int32)
    ps = L.smd_ps_create(ctypes.byref(spec), pos.ctypes.data_as(vp), types.ctypes.data_as(vp), 100, 0)
    assert ps
    res = [L.smd_ps_update(ps) for _ in range(10)]
    assert set(res) <= {0, 1} and res[-1] == 0  # false once no particle is left
    L.smd_ps_destroy(ps)


def test_headless_driver_qatest(smd, orc, gpu, tmp_path):
    """smd_particles -qatest (particles.cpp benchmark/QA mode without GL):
    the dumped positions equal the CPU oracle stepping the reference
    constructor's CONFIG_GRID state (ParticleSystem(n, 64^3), test3) for the
    same number of steps, and a second run compares PASSED against the first
    dump."""
    import subprocess

    exe = os.path.join(os.path.dirname(smd.__file__), "smd_particles")
    assert os.path.exists(exe), "driver not built"
    n, it = 8192, 12
    out = tmp_path / "particles.bin"
    print("\n[gpu-test] headless driver", flush=True)
    r = subprocess.run([exe, f"-n={n}", f"-i={it}", "-qatest", f"-out={out}", f"-ref={tmp_path / 'none.bin'}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{n} particles, total time for {it} iterations" in r.stdout
    got = np.fromfile(out, np.float32).reshape(-1, 4)
    L = smd.lib()
    L.smd_ps_create_reference.restype = vp
    L.smd_ps_create_reference.argtypes = [ctypes.c_uint] * 4
    for f in ("smd_ps_update", "smd_ps_reset", "smd_ps_destroy"):
        getattr(L, f).argtypes = [vp] + ([ctypes.c_int] if f == "smd_ps_reset" else [])
    L.smd_ps_get_positions.argtypes = [vp, vp, ctypes.c_uint]
    L.smd_ps_get_types.argtypes = [vp, vp, ctypes.c_uint]
    ps = L.smd_ps_create_reference(n, 64, 64, 64)
    assert L.smd_ps_reset(ps, 1) == 0
    p0 = np.zeros((n, 4), np.float32)
    t0 = np.zeros(n, np.int32)
    L.smd_ps_get_positions(ps, p0.ctypes.data_as(vp), n)
    L.smd_ps_get_types(ps, t0.ctypes.data_as(vp), n)
    L.smd_ps_destroy(ps)
    sim = orc.sim_for(smd, _reference_spec(), p0, t0)
    sim.step(smd.gauss_table(), it)
    want, _ = sim.state()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    r2 = subprocess.run([exe, f"-n={n}", f"-i={it}", "-qatest", f"-out={tmp_path / 'again.bin'}", f"-ref={out}"],
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0 and "TEST PASSED!" in r2.stdout, r2.stdout + r2.stderr


def test_headless_driver_result_files(smd, gpu, tmp_path):
    """smd_particles -counts-every / -traj-every: the species-count CSV and the
    trajectory frames agree with each other and with the final dump."""
    import subprocess

    exe = os.path.join(os.path.dirname(smd.__file__), "smd_particles")
    n, it, k = 4096, 12, 4
    out, csv, traj = tmp_path / "p.bin", tmp_path / "counts.csv", tmp_path / "traj.bin"
    print("\n[gpu-test] headless driver result files", flush=True)
    r = subprocess.run([exe, f"-n={n}", f"-i={it}", "-qatest", f"-out={out}", f"-ref={tmp_path / 'none.bin'}",
                        f"-counts-every={k}", f"-counts-out={csv}", f"-traj-every={k}", f"-traj={traj}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = csv.read_text().strip().splitlines()
    assert lines[0].startswith("step,species_1")
    rows = [list(map(int, ln.split(","))) for ln in lines[1:]]
    assert [row[0] for row in rows] == list(range(0, it + 1, k))
    raw = traj.read_bytes()
    assert raw[:8] == b"SMDTRJ01"
    off, frames = 8, []
    while off < len(raw):
        step, m = np.frombuffer(raw, np.int32, 2, off)
        rec = np.frombuffer(raw, np.float32, 4 * m, off + 8).reshape(m, 4)
        frames.append((int(step), rec))
        off += 8 + 16 * int(m)
    assert [f[0] for f in frames] == [row[0] for row in rows]
    for (step, rec), row in zip(frames, rows):
        types = rec[:, 3].view(np.int32)
        assert len(types) == sum(row[1:])
        for t, c in enumerate(row[1:], start=1):
            assert int(np.count_nonzero(types == t)) == c
    final = np.fromfile(out, np.float32).reshape(-1, 4)
    assert np.array_equal(frames[-1][1][:, :3].view(np.uint32), final[:, :3].view(np.uint32))
