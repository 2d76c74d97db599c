"""Writes the golden vectors under tests/golden/ from the reference's own
compiled CPU code (oracle/_ref/libsmdref.so, built by `make -C oracle ref`
from rnd/genmtrand.cpp, stats.cpp, SFMT/SFMT.cpp and rxnparam.cpp).

  mt_streams.npz   first 4096 outputs of dcmt streams (records of
                   MersenneTwister.dat, fixed seeds) from genrand_mt
  gauss_table.npy  createGaussTable(4096)
  rxnparam.npz     numrxnrate / actrxnrate / bindingradius / unbindingradius
                   over a parameter grid that includes every workload value

Run here (the reference tree is not present on GPU machines):
    make -C oracle ref && python tests/golden/make_fixtures.py
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


def main():
    L = ctypes.CDLL(REF)
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
    print("fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
