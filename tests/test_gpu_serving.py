"""Serving: a job's result is the same alone and in the mix.

Each job is run alone through a plain one-replica engine for K steps, and
together with other jobs through a slot engine (continuous batching: jobs
admitted as slots free up, slots reused in the step that retires their
previous job).  Positions and types must be bit-identical: a job's buckets,
random streams and compaction segment are its own, so the mix cannot change
a single draw.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs():
    from gpusmoldyn_amd import serving

    spec = [(1800, 37, 3), (4000, 21, 4), (2500, 60, 2), (3300, 15, 3), (1200, 44, 4), (3900, 30, 3)]
    return [serving.make_job(k, 0.0, n, steps, sp, 9000 + k) for k, (n, steps, sp) in enumerate(spec)]


def _alone(smd, job):
    e = smd.Engine(0)
    e.add_replica(job.spec, job.pos, job.types)
    e.finalize()
    e.step(job.steps)
    pos, typ = e.get(0)
    e.close()
    return pos, typ


@pytest.mark.parametrize("graphs,slot_bits", [(True, 12), (False, 12), (True, 14)])
def test_job_alone_equals_in_mix(smd, gpu, graphs, slot_bits):
    """slot_bits 14: every job's grid (2^12 buckets at most) leaves most of its
    slot's range unused, so the cell table skips those spans."""
    from gpusmoldyn_amd import serving

    jobs = _jobs()
    srv = serving.Server(n_slots=4, slot_bucket_bits=slot_bits, slot_streams=1024, capacity=1 << 16,
                         inject_cap=1 << 14, max_chunk=16, graphs=graphs)
    results, stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == len(jobs)
    assert stats["particle_steps"] > 0
    slots_used = {r.slot for r in results}
    assert len(slots_used) <= 4
    by_id = {r.jid: r for r in results}
    for job in jobs:
        r = by_id[job.jid]
        pos, typ = _alone(smd, job)
        assert r.types.shape == typ.shape, (job.jid, r.types.shape, typ.shape)
        np.testing.assert_array_equal(r.types, typ, err_msg=f"job {job.jid}")
        np.testing.assert_array_equal(r.pos, pos, err_msg=f"job {job.jid}")


def test_slot_reuse_after_retire(smd, gpu):
    """A slot emptied by retire holds no particles after the next step and
    serves a second job exactly as a fresh slot would."""
    from gpusmoldyn_amd import serving

    jobs = _jobs()[:2]
    e = smd.Engine(0)
    e.finalize_slots(2, 12, 1024, 1 << 15, 1 << 13)
    e.admit(0, jobs[0].spec, jobs[0].pos, jobs[0].types)
    e.step(jobs[0].steps + 1)
    e.retire(0)
    e.admit(0, jobs[1].spec, jobs[1].pos, jobs[1].types)  # same step removes job 0, injects job 1
    e.step(jobs[1].steps + 1)
    assert int(e.counts()[1]) == 0
    pos, typ = e.get(0)
    e.retire(0)
    e.step(1)
    assert int(e.counts().sum()) == 0
    e.close()
    apos, atyp = _alone(smd, jobs[1])
    np.testing.assert_array_equal(typ, atyp)
    np.testing.assert_array_equal(pos, apos)


def test_admit_rejects_oversized(smd, gpu):
    from gpusmoldyn_amd import serving

    e = smd.Engine(0)
    e.finalize_slots(2, 10, 256, 1 << 14, 1 << 12)
    job = serving.make_job(0, 0.0, 4000, 10, 3, 1)  # grid of 2^12 buckets > 2^10
    with pytest.raises(smd.SmdError):
        e.admit(0, job.spec, job.pos, job.types)
    e.close()
