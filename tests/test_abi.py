"""The C-ABI library loads and exports every function include/*.h declares;
the host-side setup entry points work without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INC):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        if "class ParticleSystem" in text:
            continue  # C++ class header: its methods are C++ symbols
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.add(name)
    return names


def test_every_declared_symbol_is_exported(smd):
    lib_path = smd.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    decl = declared_functions()
    assert len(decl) > 60
    missing = sorted(n for n in decl if n not in exported)
    assert not missing, f"declared but not exported: {missing}"


def test_particle_system_class_symbols(smd):
    out = subprocess.run(["nm", "-DC", "--defined-only", smd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    for m in ("ParticleSystem::update()", "ParticleSystem::reset(", "ParticleSystem::getArray(",
              "ParticleSystem::setArray(", "ParticleSystem::ParticleSystem(unsigned int, smd_u3, bool)"):
        assert m in out, m


def test_struct_layouts(smd, tmp_path):
    # reference layouts: SimParams 104 B (particles_kernel.cuh:46-67),
    # MersenneTwisterState 84 B (MersenneTwister.h:43-47), Reaction 10 pointers
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "gpusmoldyn_legacy.h"
#include "gpusmoldyn.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(SimParams), sizeof(MersenneTwisterState), sizeof(Reaction),
         offsetof(SimParams, gridSize), offsetof(SimParams, cellSize), sizeof(smd_sim_spec));
  return 0;
}
"""
    tmp = str(tmp_path / "abi_layout")
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-std=c99", "-I", INC, "-o", tmp, tmp + ".c"], check=True)
    vals = [int(x) for x in subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[:5] == [104, 84, 80, 36, 64]


def test_spec_validation_errors(smd):
    from gpusmoldyn_amd import configs
    spec, _, _ = configs.c1()
    spec.grid_bits[0] = 1
    with pytest.raises(smd.SmdError):
        smd.derive(spec)
    spec, _, _ = configs.c1()
    spec.rxn[0].reactants[1] = 9
    with pytest.raises(smd.SmdError):
        smd.derive(spec)
    spec, _, _ = configs.c1(kf=1e4)  # binding radius larger than the cell
    with pytest.raises(smd.SmdError):
        smd.derive(spec)


def test_derived_c1_reaction(smd):
    from gpusmoldyn_amd import configs
    spec, _, _ = configs.c1()
    d = smd.derive(spec)[0]
    a = smd.bindingradius(1e-3, np.float64(np.float32(1e-3)), 2.0, -1, 0)
    assert d.bind_radius2 == np.float32(np.float32(a) * np.float32(a))
    assert d.prob == 1.0 and d.setup_code == 0
    assert abs(d.bind_radius - 6.2e-3) < 1e-4


def test_reversible_unreachable_pgem_falls_back(smd):
    from gpusmoldyn_amd import configs
    spec, _, _ = configs.c2(n=1000)
    d = smd.derive(spec)
    # C -> A + B: pgem 0.2 is beyond the largest reachable geminate
    # probability at s/a ~ 6, so products go to the reaction point
    assert d[1].setup_code == 8 and d[1].unbind_radius == 0.0 and all(v == 0.0 for v in d[1].product_pos)
    # decay probability of C: k/sum (1 - exp(-dt sum)) conditioned on the earlier channel
    # (smolreact.cpp:340 evaluates 1 - expf(-dt sum) in FP32: with dt sum ~ 1e-4
    # the cancellation leaves ~5e-4 relative error, reproduced here bit for bit)
    k1, k2, dt = 0.1, 0.01, 1e-3
    p1 = k1 / (k1 + k2) * (1 - np.exp(-dt * (k1 + k2)))
    assert abs(d[1].prob - p1) / p1 < 1e-3
    p2 = k2 / (k1 + k2) * (1 - np.exp(-dt * (k1 + k2))) / (1 - p1)
    assert abs(d[2].prob - p2) / p2 < 1e-3
    f32 = np.float32
    s = f32(k1) + f32(k2)
    q1 = f32(f32(k1) / s) * f32(f32(1.0) - f32(np.exp(f32(-f32(dt) * s), dtype=np.float32)))
    assert d[1].prob == q1 / f32(1.0)


def test_reachable_pgem_sets_offsets(smd):
    from gpusmoldyn_amd import configs
    spec = configs.make_spec(
        D=[0.0, 0.01, 0.01, 0.005], dt=1e-1,
        reactions=[dict(order=2, reactants=[1, 2], products=[3], rate=0.05),
                   dict(order=1, reactants=[3], products=[1, 2], rate=3.0, pgem=0.2)],
        box_lo=[0.0] * 3, box_hi=[4.0] * 3, faces=[1] * 6, grid_bits=(4, 4, 4))
    d = smd.derive(spec)
    b = d[1].unbind_radius
    a = d[0].bind_radius
    assert d[1].setup_code == 0 and b > a
    # offsets split the unbinding radius by the diffusion coefficients (smolreact.cpp:473-476)
    assert abs(d[1].product_pos[0] - b * 0.5) < 1e-7 and abs(d[1].product_pos[3] + b * 0.5) < 1e-7
    # the reverse reaction's geminate rebinding probability is the requested 0.2
    step = np.sqrt(2 * 0.02 * 0.1)
    ki = smd.numrxnrate(step, a, -1)
    kb = smd.numrxnrate(step, a, b)
    assert abs((1 - ki / kb) - 0.2) < 2e-3
