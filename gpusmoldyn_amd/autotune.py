"""Autotuner: hash grid, sort digit width and kernel launch shapes.

The hash grid (cells per axis, a power of two; the cell size follows from
the box) does not change which particles react: a pair reacts when each is
the other's nearest reactive partner within the binding radius, whatever the
bucketing.  It does change the sorted order, hence which stream serves a
pair's draws (sorted index j uses stream j mod S), so a run on a tuned grid is
statistically, not bitwise, equal to an untuned one (and bitwise equal to the
oracle on the same grid).  The sort digit width and the launch shapes
(workgroup sizes of the producer and of the pair search) change no result.

tune(spec, pos, types)    grid and digit width of one simulation
tune_knobs(engine)        launch shapes of a finalized engine, in place
tune_ensemble(reps)       grid refinement per replica size class of an
                          ensemble; apply_ensemble(reps, deltas) applies it
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


# ---------------------------------------------------------------- launch shapes
def tune_knobs(engine, steps=10, values=None):
    """Time `steps` steps of `engine` (finalized, warmed up) under every
    combination of launch shapes and leave the fastest set.  The shapes do
    not change any result, so the timing runs continue one simulation.
    Returns (best {knob: value}, table {(producer, pair): us per step})."""
    from . import KNOB_PAIR_BLOCK, KNOB_PRODUCER_BLOCK, KNOB_VALUES

    values = values or KNOB_VALUES
    table = {}
    for pb in values[KNOB_PRODUCER_BLOCK]:
        for qb in values[KNOB_PAIR_BLOCK]:
            engine.set_knob(KNOB_PRODUCER_BLOCK, pb)
            engine.set_knob(KNOB_PAIR_BLOCK, qb)
            engine.step(2)  # capture the step graphs with this shape
            engine.sync()
            t0 = time.perf_counter()
            engine.step(steps)
            engine.sync()
            table[(pb, qb)] = (time.perf_counter() - t0) * 1e6 / steps
    best = min(table, key=table.get)
    engine.set_knob(KNOB_PRODUCER_BLOCK, best[0])
    engine.set_knob(KNOB_PAIR_BLOCK, best[1])
    return {KNOB_PRODUCER_BLOCK: best[0], KNOB_PAIR_BLOCK: best[1]}, table


# ---------------------------------------------------------------- ensembles
def size_class(n):
    """Replicas are tuned per power-of-two size class (n in [2^c, 2^(c+1)))."""
    return max(0, int(n).bit_length() - 1)


def _shift(spec, delta, per_stream=None, n=None):
    s = apply(spec, tuple(int(spec.grid_bits[d]) + delta for d in range(3)))
    if per_stream:
        s.num_streams = max(64, min(32768, -(-int(n) // int(per_stream))))
    return s


def _time_ensemble(reps, choice, steps, warmup, device):
    """Wall time per particle-step (ns) of the whole ensemble with the
    per-class choice applied; None when the engine rejects it."""
    e = Engine(device)
    try:
        for spec, pos, types in apply_ensemble(reps, choice):
            e.add_replica(spec, pos, types)
        e.finalize()
    except SmdError:
        e.close()
        return None
    try:
        e.step(warmup)
        e.sync()
        ps0 = e.particle_steps()
        t0 = time.perf_counter()
        e.step(steps)
        e.sync()
        return (time.perf_counter() - t0) * 1e9 / max(e.particle_steps() - ps0, 1)
    finally:
        e.close()


def tune_ensemble(reps, deltas=(-1, 0, 1), per_stream=(None,), steps=10, warmup=4, device=0):
    """Per size class, the grid refinement (delta bits per axis on each
    replica's own grid) and the particles per random stream (None: the
    replica's own stream count), timed on the whole ensemble: the key width
    of the sort and the size of the cell table depend on every replica's
    grid together, so classes are not timed apart.  First one choice for
    every class, then each class in turn (largest first) tries the other
    refinements with the rest fixed.  Timings a few percent apart are within
    the noise of short runs: the common choices within 3% of the fastest are
    timed again over three times the steps, in the order given (so ties keep
    the earlier grid and the more streams), a later one replacing an earlier
    one only for a gain above 1%, and a class changes only for a gain above
    1%.  reps: [(spec, pos, types)].  Returns
    ({class: (delta, per_stream)}, {label: ns per particle-step})."""
    classes = sorted({size_class(len(t)) for _, _, t in reps}, reverse=True)
    table = {}
    timed = []
    for d in deltas:
        for p in per_stream:
            ch = {c: (d, p) for c in classes}
            t = _time_ensemble(reps, ch, steps, warmup, device)
            if t is None:
                continue
            table[f"all {d:+d} /{p}"] = t
            timed.append((t, d, p, ch))
    if not timed:
        raise SmdError("no valid ensemble grid")
    fastest = min(x[0] for x in timed)
    close = [x for x in timed if x[0] <= 1.03 * fastest]
    best, best_t = None, None
    for _, d, p, ch in close:
        t = _time_ensemble(reps, ch, 3 * steps, warmup, device) if len(close) > 1 else table[f"all {d:+d} /{p}"]
        if t is None:
            continue
        if len(close) > 1:
            table[f"all {d:+d} /{p} again"] = t
        if best_t is None or t < 0.99 * best_t:
            best, best_t = ch, t
    if best is None:
        raise SmdError("no valid ensemble grid")
    for c in classes:
        for d in deltas:
            if d == best[c][0]:
                continue
            ch = dict(best)
            ch[c] = (d, best[c][1])
            t = _time_ensemble(reps, ch, steps, warmup, device)
            if t is None:
                continue
            table[f"2^{c} {d:+d}"] = t
            if t < 0.99 * best_t:
                best, best_t = ch, t
    return best, table


def apply_ensemble(reps, choice):
    """The replicas with each one's grid refined (and stream count set) by its
    size class's choice: {class: (delta, per_stream)} or {class: delta}."""
    out = []
    for spec, pos, types in reps:
        ch = choice.get(size_class(len(types)), 0)
        dlt, per = ch if isinstance(ch, tuple) else (ch, None)
        out.append((_shift(spec, dlt, per, len(types)), pos, types))
    return out
