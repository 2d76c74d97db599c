"""The rdf table maker batched on the GPU (one workgroup per sweep):
regenerates rows of the reference's rate tables (yi: 600 points, yr: 500
points, eps 1e-5, mean of the upward and downward sweeps) and agrees with
the host restatement on a small grid."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def _sweeps(smd, n, eps, modes, bs, downs):
    L = smd.lib()
    L.smd_rdf_sweeps_gpu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_char_p,
                                     vp, vp, vp]
    m = "".join(modes).encode()
    b = np.ascontiguousarray(bs, np.float64)
    d = np.ascontiguousarray(downs, np.int32)
    out = np.zeros((len(modes), 31))
    rc = L.smd_rdf_sweeps_gpu(0, n, eps, len(modes), m, b.ctypes.data_as(vp), d.ctypes.data_as(vp),
                              out.ctypes.data_as(vp))
    assert rc == 0, smd.last_error()
    return out


def test_rdf_gpu_matches_host_small(smd, gpu):
    L = smd.lib()
    L.smd_rdf_sweep.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, vp]
    print("\n[gpu-test] rdf gpu vs host n=80", flush=True)
    cases = [("i", -1.0, 0), ("i", -1.0, 1), ("r", math.exp(0.4), 0), ("b", 0.5, 1)]
    got = _sweeps(smd, 80, 1e-4, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases])
    for q, (mode, b, down) in enumerate(cases):
        want = np.zeros(31)
        assert L.smd_rdf_sweep(mode.encode(), 80, 1e-4, down, b, want.ctypes.data_as(vp)) == 0
        np.testing.assert_allclose(got[q], want, rtol=1e-9, err_msg=f"sweep {q}")


def test_rdf_gpu_regenerates_tables(smd, gpu):
    print("\n[gpu-test] rdf tables", flush=True)
    yi = _sweeps(smd, 600, 1e-5, ["i", "i"], [-1.0, -1.0], [0, 1])
    steps = [math.exp(-3.0 + 0.2 * k) for k in range(31)]
    mean = yi.mean(axis=0)
    want = np.array([smd.numrxnrate(s, 1.0, -1.0) for s in steps])
    np.testing.assert_allclose(mean, want, rtol=1e-4)
    rows = [0, 3, 8]  # yr rows b = exp(0.2 m)
    bs = [math.exp(0.2 * m) for m in rows for _ in (0, 1)]
    yr = _sweeps(smd, 500, 1e-5, ["r"] * len(bs), bs, [0, 1] * len(rows))
    for k, m in enumerate(rows):
        mean = yr[2 * k:2 * k + 2].mean(axis=0)
        want = np.array([smd.numrxnrate(s, 1.0, math.exp(0.2 * m)) for s in steps])
        np.testing.assert_allclose(mean, want, rtol=5e-4, err_msg=f"yr row {m}")
