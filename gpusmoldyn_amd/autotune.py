"""Grid and sort-digit autotuner.

The hash grid (cells per axis, a power of two; the cell size follows from
the box) and the sort digit width do not change which particles react: a
pair reacts when each is the other's nearest reactive partner within the
binding radius, whatever the bucketing.  They do change the sorted order,
hence which stream serves a pair's draws (sorted index j uses stream j mod S),
so a tuned run is statistically, not bitwise, equal to an untuned one.

tune(spec, pos, types) times a few steps of each candidate on the GPU and
returns the fastest (grid_bits, radix_bits) with the timing table.
"""
import copy
import time

from . import Engine, SmdError, derive


def _candidates(spec, max_delta=2):
    base = [int(spec.grid_bits[d]) for d in range(3)]
    der = derive(spec) if spec.num_reactions else []
    max_bind = max([d.bind_radius for d in der] + [0.0])
    out = []
    for delta in range(-max_delta, max_delta + 1):
        bits = [b + delta for b in base]
        if min(bits) < 2 or max(bits) > 12 or sum(bits) > 28:
            continue
        cell_ok = all((spec.box_hi[d] - spec.box_lo[d]) / (1 << bits[d]) >= max_bind for d in range(3))
        if spec.cell_size[0] > 0:  # explicit cell size: keep it, vary only the wrap period
            cell_ok = True
        if cell_ok:
            out.append(tuple(bits))
    return out


def time_config(spec, pos, types, grid_bits, radix_bits, steps=20, warmup=3, device=0, cell=None):
    """Average wall time (us) of one step with the given grid, cell size and
    digits."""
    s = apply(spec, grid_bits, cell)
    e = Engine(device, radix_bits)
    try:
        e.add_replica(s, pos, types)
        e.finalize()
        e.step(warmup)
        e.sync()
        t0 = time.perf_counter()
        e.step(steps)
        e.sync()
        return (time.perf_counter() - t0) * 1e6 / steps
    finally:
        e.close()


def tune(spec, pos, types, steps=20, radix=(8, 4), max_delta=2, device=0, cells=None, grids=None):
    """Returns ((grid_bits, radix_bits, cell), table) with table[key] = us per
    step; configurations the engine rejects are left out.  Without `cells`
    the cell size follows the box (or stays the spec's explicit one) and the
    grids are the spec's scaled by -max_delta..max_delta bits per axis;
    `cells` (explicit cell sizes) and `grids` (grid_bits triples) give the
    candidates directly."""
    grid_list = [tuple(g) for g in grids] if grids else _candidates(spec, max_delta)
    table = {}
    for cell in (cells or [None]):
        for bits in grid_list:
            for rb in radix:
                try:
                    table[(bits, rb, cell)] = time_config(spec, pos, types, bits, rb, steps, device=device, cell=cell)
                except SmdError:
                    continue
    if not table:
        raise SmdError("no valid grid candidate")
    best = min(table, key=table.get)
    return best, table


def apply(spec, grid_bits, cell=None):
    """A copy of spec with the given grid (and explicit cell size)."""
    s = copy.deepcopy(spec)
    for d in range(3):
        s.grid_bits[d] = grid_bits[d]
        if cell is not None:
            s.cell_size[d] = cell
    return s
