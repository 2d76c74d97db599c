"""gpusmoldyn_amd: MI355X (gfx950) Brownian-dynamics reaction-diffusion timestep.

Python host mirror of the C-ABI in include/gpusmoldyn.h (ctypes).  The HIP
library libgpusmoldyn.so must have been built in-tree
(``python -m gpusmoldyn_amd.build``); importing the GPU entry points without
it raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpusmoldyn.so")

MAX_TYPES = 16
MAX_RXN = 32
MAX_SPHERES = 4
MT_PARAM_SETS = 32768
GAUSS_DIM = 4096

BC_CLAMP, BC_REFLECT, BC_ABSORB, BC_PERIODIC = 0, 1, 2, 3
KNOB_PRODUCER_BLOCK, KNOB_PAIR_BLOCK, KNOB_SEGMENTED_SORT = 0, 1, 2
KNOB_VALUES = {KNOB_PRODUCER_BLOCK: (128, 256, 512), KNOB_PAIR_BLOCK: (64, 128, 256)}


class SmdError(RuntimeError):
    pass


class Reaction(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("reactants", ctypes.c_int * 2),
        ("products", ctypes.c_int * 2),
        ("nprod", ctypes.c_int),
        ("rate", ctypes.c_float),
        ("pgem", ctypes.c_float),
        ("prob", ctypes.c_float),
    ]


class SimSpec(ctypes.Structure):
    _fields_ = [
        ("num_types", ctypes.c_int),
        ("D", ctypes.c_float * MAX_TYPES),
        ("dt", ctypes.c_float),
        ("volume", ctypes.c_float),
        ("num_reactions", ctypes.c_int),
        ("rxn", Reaction * MAX_RXN),
        ("box_lo", ctypes.c_float * 3),
        ("box_hi", ctypes.c_float * 3),
        ("face", ctypes.c_int * 6),
        ("particle_radius", ctypes.c_float),
        ("num_spheres", ctypes.c_int),
        ("sphere", (ctypes.c_float * 4) * MAX_SPHERES),
        ("unbind_offsets", ctypes.c_int),
        ("grid_bits", ctypes.c_uint * 3),
        ("cell_size", ctypes.c_float * 3),
        ("origin", ctypes.c_float * 3),
        ("num_streams", ctypes.c_uint),
        ("seed", ctypes.c_uint),
    ]


class RxnDerived(ctypes.Structure):
    _fields_ = [
        ("prob", ctypes.c_float),
        ("bind_radius2", ctypes.c_float),
        ("bind_radius", ctypes.c_float),
        ("unbind_radius", ctypes.c_float),
        ("product_pos", ctypes.c_float * 6),
        ("setup_code", ctypes.c_int),
    ]


_lib = None


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SmdError(f"{LIB_PATH} not found: build it with `python -m gpusmoldyn_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32p, f32p, i32p = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
        L.smd_last_error.restype = ctypes.c_char_p
        L.smd_engine_create.restype = ctypes.c_void_p
        L.smd_engine_create.argtypes = [ctypes.c_int, ctypes.c_int]
        L.smd_engine_destroy.argtypes = [vp]
        L.smd_engine_add_replica.argtypes = [vp, ctypes.POINTER(SimSpec), f32p, i32p, ctypes.c_uint]
        L.smd_engine_finalize.argtypes = [vp, ctypes.c_uint]
        L.smd_engine_finalize_slots.argtypes = [vp] + [ctypes.c_uint] * 5
        L.smd_engine_admit.argtypes = [vp, ctypes.c_int, ctypes.POINTER(SimSpec), f32p, i32p, ctypes.c_uint]
        L.smd_engine_retire.argtypes = [vp, ctypes.c_int]
        L.smd_engine_step.argtypes = [vp, ctypes.c_int]
        L.smd_engine_sync.argtypes = [vp]
        L.smd_engine_counts.argtypes = [vp, u32p]
        L.smd_engine_particle_steps.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong)]
        L.smd_engine_set_graphs.argtypes = [vp, ctypes.c_int]
        L.smd_engine_set_sort.argtypes = [vp, ctypes.c_int]
        L.smd_engine_set_knob.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.smd_engine_get_knob.argtypes = [vp, ctypes.c_int]
        L.smd_engine_num_replicas.argtypes = [vp]
        L.smd_engine_radix_bits.argtypes = [vp]
        L.smd_engine_num_replicas.restype = ctypes.c_uint
        L.smd_engine_get.argtypes = [vp, ctypes.c_int, f32p, i32p, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        L.smd_engine_species_counts.argtypes = [vp, ctypes.c_int, u32p]
        L.smd_engine_get_streams.argtypes = [vp, ctypes.c_int, u32p, ctypes.c_uint]
        L.smd_engine_set_streams.argtypes = [vp, ctypes.c_int, u32p, ctypes.c_uint]
        L.smd_engine_rxn.argtypes = [vp, ctypes.c_int, ctypes.POINTER(RxnDerived)]
        L.smd_engine_debug_first.argtypes = [vp, u32p, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        L.smd_engine_debug_stage.argtypes = [vp, ctypes.c_int, f32p, i32p, u32p, u32p, i32p,
                                             ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]
        L.smd_engine_layout.argtypes = [vp] + [ctypes.POINTER(ctypes.c_uint)] * 4
        L.smd_engine_sort_plan.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.smd_engine_profile_step.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                              ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.smd_mt_generate.argtypes = [ctypes.c_int, u32p, ctypes.c_uint, ctypes.c_int, ctypes.c_int, u32p]
        L.smd_sort_pairs.argtypes = [ctypes.c_int, u32p, u32p, ctypes.c_uint, ctypes.c_uint, ctypes.c_int]
        L.smd_gauss_table.argtypes = [f32p, ctypes.c_int]
        L.smd_copy_probe.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        for f, n in (("numrxnrate", 3), ("actrxnrate", 2), ("unbindingradius", 4)):
            fn = getattr(L, "smd_" + f)
            fn.restype = ctypes.c_double
            fn.argtypes = [ctypes.c_double] * n
        L.smd_bindingradius.restype = ctypes.c_double
        L.smd_bindingradius.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int]
        L.smd_spec_derive.argtypes = [ctypes.POINTER(SimSpec), ctypes.POINTER(RxnDerived)]
        L.smd_spec_cells.argtypes = [ctypes.POINTER(SimSpec), f32p, f32p]
        L.smd_stream_seeds.argtypes = [ctypes.c_uint, ctypes.c_uint, u32p]
        L.smd_mt_params.argtypes = [ctypes.c_uint, u32p]
        L.smd_set_mt_dat_path.argtypes = [ctypes.c_char_p]
        L.smd_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc, what=""):
    if rc != 0:
        msg = lib().smd_last_error().decode(errors="replace")
        raise SmdError(f"{what} failed with status {rc}: {msg}")


def last_error():
    return lib().smd_last_error().decode(errors="replace")


# ---------------------------------------------------------------- host setup
def gauss_table(n=GAUSS_DIM):
    out = np.zeros(n, np.float32)
    _check(lib().smd_gauss_table(_ptr(out), n), "smd_gauss_table")
    return out


def numrxnrate(step, a, b):
    return lib().smd_numrxnrate(step, a, b)


def actrxnrate(step, a):
    return lib().smd_actrxnrate(step, a)


def bindingradius(rate, dt, difc, b=-1.0, rel=0):
    return lib().smd_bindingradius(rate, dt, difc, b, rel)


def unbindingradius(pgem, dt, difc, a):
    return lib().smd_unbindingradius(pgem, dt, difc, a)


def stream_seeds(seed, count):
    out = np.zeros(count, np.uint32)
    _check(lib().smd_stream_seeds(seed, count, _ptr(out)), "smd_stream_seeds")
    return out


def mt_params(index):
    out = np.zeros(3, np.uint32)
    _check(lib().smd_mt_params(index, _ptr(out)), "smd_mt_params")
    return out


def derive(spec):
    out = (RxnDerived * MAX_RXN)()
    rc = lib().smd_spec_derive(ctypes.byref(spec), out)
    _check(rc, "smd_spec_derive")
    return [out[k] for k in range(spec.num_reactions)]


def spec_cells(spec):
    cell = np.zeros(3, np.float32)
    origin = np.zeros(3, np.float32)
    _check(lib().smd_spec_cells(ctypes.byref(spec), _ptr(cell), _ptr(origin)), "smd_spec_cells")
    return cell, origin


def device_count():
    n = ctypes.c_int(0)
    lib().smd_device_count(ctypes.byref(n))
    return n.value


def mt_generate(seeds, count, device=0, batch3=False):
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros((len(seeds), count), np.uint32)
    _check(lib().smd_mt_generate(device, _ptr(seeds), len(seeds), count, int(batch3), _ptr(out)), "smd_mt_generate")
    return out


def sort_pairs(keys, values, key_bits, radix_bits=8, device=0):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(values, np.uint32).copy()
    _check(lib().smd_sort_pairs(device, _ptr(k), _ptr(v), len(k), key_bits, radix_bits), "smd_sort_pairs")
    return k, v


def copy_probe(nbytes, reps=10, device=0):
    """Average duration (us) of a plain float4 copy kernel over nbytes."""
    us = ctypes.c_double(0)
    _check(lib().smd_copy_probe(device, nbytes, reps, ctypes.byref(us)), "smd_copy_probe")
    return us.value


# ---------------------------------------------------------------- engine
class Engine:
    """An ensemble of independent replicas stepped together on one GPU (a
    plain simulation is an ensemble of one)."""

    def __init__(self, device=0, radix_bits=0):
        L = lib()
        h = L.smd_engine_create(device, radix_bits)
        if not h:
            raise SmdError("smd_engine_create: " + last_error())
        self._h = ctypes.c_void_p(h)
        self.specs = []
        self.finalized = False

    def close(self):
        if getattr(self, "_h", None):
            lib().smd_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_replica(self, spec, pos4, types):
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        types = np.ascontiguousarray(types, np.int32)
        assert len(pos4) == len(types)
        r = lib().smd_engine_add_replica(self._h, ctypes.byref(spec), _ptr(pos4), _ptr(types), len(types))
        if r < 0:
            _check(-r, "smd_engine_add_replica")
        self.specs.append(spec)
        return r

    def finalize(self, capacity=0):
        _check(lib().smd_engine_finalize(self._h, capacity), "smd_engine_finalize")
        self.finalized = True

    def finalize_slots(self, n_slots, slot_bucket_bits, slot_streams, capacity, inject_cap):
        """Serving layout: n_slots empty replica slots (see admit/retire)."""
        _check(lib().smd_engine_finalize_slots(self._h, n_slots, slot_bucket_bits, slot_streams, capacity,
                                               inject_cap), "smd_engine_finalize_slots")
        self.specs = [None] * n_slots
        self.finalized = True

    def admit(self, slot, spec, pos4, types):
        pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
        types = np.ascontiguousarray(types, np.int32)
        assert len(pos4) == len(types)
        _check(lib().smd_engine_admit(self._h, slot, ctypes.byref(spec), _ptr(pos4), _ptr(types), len(types)),
               "smd_engine_admit")
        self.specs[slot] = spec

    def retire(self, slot):
        _check(lib().smd_engine_retire(self._h, slot), "smd_engine_retire")

    def step(self, nsteps=1):
        _check(lib().smd_engine_step(self._h, nsteps), "smd_engine_step")

    def sync(self):
        _check(lib().smd_engine_sync(self._h), "smd_engine_sync")

    def counts(self):
        R = lib().smd_engine_num_replicas(self._h)
        out = np.zeros(R, np.uint32)
        _check(lib().smd_engine_counts(self._h, _ptr(out)), "smd_engine_counts")
        return out

    def set_sort(self, kind):
        """1: count / scan / scatter per digit (default), 0: single-sweep radix sort."""
        _check(lib().smd_engine_set_sort(self._h, int(kind)), "smd_engine_set_sort")

    def set_knob(self, knob, value):
        """Launch shape (KNOB_PRODUCER_BLOCK, KNOB_PAIR_BLOCK) or sort layout
        (KNOB_SEGMENTED_SORT: 1 sorts each replica on its local cells, 0 all
        keys at once); results are identical for every setting."""
        _check(lib().smd_engine_set_knob(self._h, int(knob), int(value)), "smd_engine_set_knob")

    def get_knob(self, knob):
        v = lib().smd_engine_get_knob(self._h, int(knob))
        if v < 0:
            _check(-v, "smd_engine_get_knob")
        return v

    def set_graphs(self, on):
        _check(lib().smd_engine_set_graphs(self._h, int(bool(on))), "smd_engine_set_graphs")

    def particle_steps(self):
        v = ctypes.c_ulonglong(0)
        _check(lib().smd_engine_particle_steps(self._h, ctypes.byref(v)), "smd_engine_particle_steps")
        return v.value

    def get(self, r=0):
        n = ctypes.c_uint(0)
        _check(lib().smd_engine_get(self._h, r, None, None, 0, ctypes.byref(n)), "smd_engine_get")
        pos = np.zeros((n.value, 4), np.float32)
        typ = np.zeros(n.value, np.int32)
        _check(lib().smd_engine_get(self._h, r, _ptr(pos), _ptr(typ), n.value, ctypes.byref(n)), "smd_engine_get")
        return pos, typ

    def species_counts(self, r=0):
        out = np.zeros(MAX_TYPES, np.uint32)
        _check(lib().smd_engine_species_counts(self._h, r, _ptr(out)), "smd_engine_species_counts")
        return out[: self.specs[r].num_types]

    def get_streams(self, r=0):
        S = self.specs[r].num_streams or MT_PARAM_SETS
        out = np.zeros((S, 20), np.uint32)
        _check(lib().smd_engine_get_streams(self._h, r, _ptr(out), S), "smd_engine_get_streams")
        return out

    def set_streams(self, r, words):
        words = np.ascontiguousarray(words, np.uint32)
        _check(lib().smd_engine_set_streams(self._h, r, _ptr(words), words.shape[0]), "smd_engine_set_streams")

    def rxn(self, r=0):
        out = (RxnDerived * MAX_RXN)()
        _check(lib().smd_engine_rxn(self._h, r, out), "smd_engine_rxn")
        return [out[k] for k in range(self.specs[r].num_reactions)]

    def radix_bits(self):
        return lib().smd_engine_radix_bits(self._h)

    def sort_plan(self):
        """(digit passes of the step's key sort, True when it sorts each
        replica on its local cells)."""
        p, g = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().smd_engine_sort_plan(self._h, ctypes.byref(p), ctypes.byref(g)), "smd_engine_sort_plan")
        return p.value, bool(g.value)

    def layout(self):
        v = [ctypes.c_uint(0) for _ in range(4)]
        _check(lib().smd_engine_layout(self._h, *[ctypes.byref(x) for x in v]), "smd_engine_layout")
        return dict(total_buckets=v[0].value, key_bits=v[1].value, total_streams=v[2].value, capacity=v[3].value)

    def debug_stage(self, stage, n_total):
        pos = np.zeros((n_total, 4), np.float32)
        typ = np.zeros(n_total, np.int32)
        keys = np.zeros(n_total, np.uint32)
        vals = np.zeros(n_total, np.uint32)
        cand = np.zeros(n_total, np.int32)
        nt, nl = ctypes.c_uint(0), ctypes.c_uint(0)
        _check(lib().smd_engine_debug_stage(self._h, stage, _ptr(pos), _ptr(typ), _ptr(keys), _ptr(vals), _ptr(cand),
                                            ctypes.byref(nt), ctypes.byref(nl)), "smd_engine_debug_stage")
        return dict(pos=pos, type=typ, keys=keys, vals=vals, cand=cand[: nl.value], n_total=nt.value,
                    n_live=nl.value)

    def debug_first(self):
        """Cell table after debug_stage(stage >= 2): first[h] for every key h
        up to the removed-particle key (= total buckets), then the particle
        count."""
        n = ctypes.c_uint(0)
        _check(lib().smd_engine_debug_first(self._h, None, 0, ctypes.byref(n)), "smd_engine_debug_first")
        out = np.zeros(n.value, np.uint32)
        _check(lib().smd_engine_debug_first(self._h, _ptr(out), n.value, ctypes.byref(n)), "smd_engine_debug_first")
        return out

    def profile(self, nsteps):
        us = (ctypes.c_double * 16)()
        names = ctypes.create_string_buffer(512)
        ns = ctypes.c_int(0)
        _check(lib().smd_engine_profile_step(self._h, nsteps, us, 16, names, 512, ctypes.byref(ns)),
               "smd_engine_profile_step")
        keys = names.value.decode().split("\n")
        return {k: us[i] for i, k in enumerate(keys[: ns.value])}
