"""Writes the golden vectors under tests/golden/ from the reference's own
compiled CPU code (oracle/_ref/libsmdref.so, built by `make -C oracle ref`
from rnd/genmtrand.cpp, stats.cpp, SFMT/SFMT.cpp and rxnparam.cpp).

  mt_streams.npz   first 4096 outputs of dcmt streams (records of
                   MersenneTwister.dat, fixed seeds) from genrand_mt
  gauss_table.npy  createGaussTable(4096)
  rxnparam.npz     numrxnrate / actrxnrate / bindingradius / unbindingradius
                   over a parameter grid that includes every workload value
  rdf.npz          rdfsteadystate (irreversible and reversible, b < 1 and
                   b > 1) on small radial grids: flux and both output rdfs
  sfmt_uniform.npz unirandOOF(-1, 1) after init_gen_rand(seed), seeds 0 and
                   4357: the randPos sequence (SFMT-19937)

Run here (the reference tree is not present on GPU machines):
    make -C oracle ref && python tests/golden/make_fixtures.py [only-name ...]
"""
import ctypes
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libsmdref.so")
DAT = os.path.join(ROOT, "gpusmoldyn_amd", "data", "MersenneTwister.dat")

MT_STREAMS = [0, 1, 2, 31, 32, 1023, 32767]
MT_COUNT = 4096


def mt_seed(k):
    return (4357 + 1000 * k) & 0xFFFFFFFF


def write_sfmt(L):
    L.smdref_sfmt_uniform.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_void_p]
    seeds = np.array([0, 4357], np.uint32)
    # 3 x 2000 draws crosses the 624-word refill boundary several times
    out = np.zeros((len(seeds), 6000), np.float32)
    for k, sd in enumerate(seeds):
        L.smdref_sfmt_uniform(int(sd), out.shape[1], -1.0, 1.0, out[k].ctypes.data_as(ctypes.c_void_p))
    np.savez_compressed(os.path.join(HERE, "sfmt_uniform.npz"), seeds=seeds, lo=np.float32(-1.0),
                        hi=np.float32(1.0), values=out)


def main():
    L = ctypes.CDLL(REF)
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            {"sfmt": write_sfmt, "rdf": write_rdf}[name](L)
        return 0
    dbl = ctypes.c_double
    L.smdref_mt_stream.argtypes = [ctypes.c_uint32] * 4 + [ctypes.c_int, ctypes.c_void_p]
    L.smdref_gauss_table.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.smdref_numrxnrate.restype = dbl
    L.smdref_numrxnrate.argtypes = [dbl] * 3
    L.smdref_actrxnrate.restype = dbl
    L.smdref_actrxnrate.argtypes = [dbl] * 2
    L.smdref_bindingradius.restype = dbl
    L.smdref_bindingradius.argtypes = [dbl] * 4 + [ctypes.c_int]
    L.smdref_unbindingradius.restype = dbl
    L.smdref_unbindingradius.argtypes = [dbl] * 4

    dat = np.fromfile(DAT, dtype="<u4").reshape(-1, 4)
    out = np.zeros((len(MT_STREAMS), MT_COUNT), np.uint32)
    seeds = np.array([mt_seed(k) for k in MT_STREAMS], np.uint32)
    for row, k in enumerate(MT_STREAMS):
        a, b, c = (int(x) for x in dat[k, :3])
        L.smdref_mt_stream(a, b, c, int(seeds[row]), MT_COUNT, out[row].ctypes.data_as(ctypes.c_void_p))
    np.savez_compressed(os.path.join(HERE, "mt_streams.npz"), streams=np.array(MT_STREAMS, np.int32),
                        seeds=seeds, outputs=out)

    g = np.zeros(4096, np.float32)
    L.smdref_gauss_table(g.ctypes.data_as(ctypes.c_void_p), 4096)
    np.save(os.path.join(HERE, "gauss_table.npy"), g)

    steps = np.geomspace(1e-3, 50.0, 23)
    radii = np.geomspace(1e-3, 10.0, 9)
    bred = np.array([-1.0, 0.0, 0.05, 0.3, 0.7, 0.99, 1.0, 1.5, 3.0, 10.0, 30.0])
    nr = []
    for s, a, b in itertools.product(steps, radii, bred):
        bb = b * a if b >= 0 else -1.0
        nr.append((s, a, bb, L.smdref_numrxnrate(s, a, bb)))
    ar = [(s, a, L.smdref_actrxnrate(s, a)) for s, a in itertools.product(steps, radii)]
    rates = list(np.geomspace(1e-4, 1e2, 13)) + [1e-3, 5e-3, 1e-2, 2e-3]
    dts = [1e-4, 1e-3, 1e-2, 0.1]
    difcs = [0.01, 0.5, 1.0, 1.5, 2.0]
    br, ub = [], []
    for rate, dt, d in itertools.product(rates, dts, difcs):
        a = L.smdref_bindingradius(rate, dt, d, -1.0, 0)
        br.append((rate, dt, d, a))
        for pg in (0.05, 0.2, 0.5, 0.9):
            ub.append((pg, dt, d, a, L.smdref_unbindingradius(pg, dt, d, a)))
    brr = [(rate, dt, d, b, L.smdref_bindingradius(rate, dt, d, b, 1))
           for rate, dt, d, b in itertools.product([1e-3, 1e-2, 1.0], [1e-3, 1e-2], [1.0, 2.0], [0.5, 2.0, 5.0])]
    np.savez_compressed(os.path.join(HERE, "rxnparam.npz"), numrxnrate=np.array(nr), actrxnrate=np.array(ar),
                        bindingradius=np.array(br), bindingradius_rel=np.array(brr),
                        unbindingradius=np.array(ub))
    write_rdf(L)
    write_sfmt(L)
    print("fixtures written to", HERE)


def rdf_grid(rmax, n):
    # the table maker's grid (rxnparam.cpp:500-505)
    r = np.zeros(n)
    dr = rmax / n
    i = 1
    while r[i - 1] < 1 and i - 1 < n:
        r[i] = r[i - 1] + dr
        i += 1
    r[i - 1] = 0.9999
    r[i] = 1.0001
    i += 1
    while i < n:
        r[i] = r[i - 1] + dr
        i += 1
    return r


def write_rdf(L):
    dbl = ctypes.c_double
    L.smdref_rdf_steady_state.restype = dbl
    L.smdref_rdf_steady_state.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, dbl, dbl, dbl]
    cases = [(10.0, 60, 0.3, -1.0), (10.0, 60, 2.0, -1.0), (5.0, 50, 0.7, 0.5), (5.0, 50, 1.5, 0.0),
             (5.0, 50, 1.0, 2.0)]
    rows = []
    for rmax, n, step, b in cases:
        r = rdf_grid(rmax, n)
        a0 = np.where(r < 1, 0.0, 1.0 - 1.0 / np.maximum(r, 1e-300))
        a = a0.copy()
        d = np.zeros(n)
        flux = L.smdref_rdf_steady_state(r.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
                                         d.ctypes.data_as(ctypes.c_void_p), n, step, b, 1e-6)
        rows.append((rmax, n, step, b, flux, r, a0, a, d))
    np.savez_compressed(os.path.join(HERE, "rdf.npz"),
                        params=np.array([(rm, n, s, b, f) for rm, n, s, b, f, *_ in rows]),
                        **{f"r{k}": row[5] for k, row in enumerate(rows)},
                        **{f"init{k}": row[6] for k, row in enumerate(rows)},
                        **{f"absorbed{k}": row[7] for k, row in enumerate(rows)},
                        **{f"diffused{k}": row[8] for k, row in enumerate(rows)})


if __name__ == "__main__":
    sys.exit(main())
