"""The rdf table generator (rxnparam.cpp:325-564) restated in FP64: against
the reference's compiled rdfsteadystate (golden vectors) and against the
reaction-rate tables it produced (yi, yr), which numrxnrate returns at the
table nodes."""
import ctypes
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
vp = ctypes.c_void_p


def _lib(smd):
    L = smd.lib()
    L.smd_rdf_steady_state.argtypes = [vp] * 3 + [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.smd_rdf_steady_state.restype = ctypes.c_double
    L.smd_rdf_grid.argtypes = [ctypes.c_double, ctypes.c_int, vp]
    return L


def test_rdf_steady_state_matches_reference(smd):
    L = _lib(smd)
    g = np.load(os.path.join(GOLD, "rdf.npz"))
    for k, (rmax, n, step, b, flux) in enumerate(g["params"]):
        n = int(n)
        r = np.zeros(n)
        L.smd_rdf_grid(rmax, n, r.ctypes.data_as(vp))
        assert np.array_equal(r, g[f"r{k}"]), f"grid {k}"
        a = g[f"init{k}"].copy()
        d = np.zeros(n)
        f = L.smd_rdf_steady_state(r.ctypes.data_as(vp), a.ctypes.data_as(vp), d.ctypes.data_as(vp), n, step, b,
                                   1e-6)
        assert f == pytest.approx(flux, rel=1e-13, abs=0), f"flux {k}"
        np.testing.assert_allclose(a, g[f"absorbed{k}"], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(d, g[f"diffused{k}"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("mode,b", [("i", -1.0), ("r", math.exp(0.6))])
def test_rdf_regenerates_rate_table_large_steps(smd, mode, b):
    """Downward sweep of the table maker over the 10 largest reduced steps
    (600 points, eps 1e-5): within 2e-4 of the tabulated rates (the tables
    hold the mean of the upward and downward sweeps)."""
    L = _lib(smd)
    n = 600
    rmax = 10.0 if mode == "i" else b + 3.0
    r = np.zeros(n)
    L.smd_rdf_grid(rmax, n, r.ctypes.data_as(vp))
    a = np.where(r < 1, 0.0, 1.0)
    d = np.zeros(n)
    s = math.exp(3.0)
    inc = 1.0 / math.exp(0.2)
    for k in range(10):
        f = L.smd_rdf_steady_state(r.ctypes.data_as(vp), a.ctypes.data_as(vp), d.ctypes.data_as(vp), n, s, b, 1e-5)
        want = smd.numrxnrate(math.exp(-3.0 + 0.2 * (30 - k)), 1.0, b)
        assert abs(f / want - 1.0) < 2e-4, (k, s, f, want)
        s *= inc
