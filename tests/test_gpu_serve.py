"""GPU: the serving loop (mset_amd/serve.py) -- every job's results are the same alone and in a mix of
concurrent jobs under a memory budget that forces evictions and re-admissions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(specs, order, workers, budget):
    from mset_amd import serve
    srv = serve.Server(specs, workers=workers, budget_bytes=budget, batch=64)
    reqs = [srv.submit(j) for j in order]
    for r in reqs:
        assert r.done.wait(300), 'request did not finish'
    assert all(not r.failure for r in reqs), [r.failure for r in reqs if r.failure]
    states = {s.job_id: srv.state_of(s.job_id) for s in specs}
    errors = {s.job_id: list(srv.jobs[s.job_id].errors) for s in specs}
    rep = srv.report()
    srv.close()
    return errors, states, rep


def test_results_are_the_same_alone_and_in_the_mix():
    from mset_amd import serve, lib
    specs = [serve.JobSpec(job_id=0, vol=32, P=64, tilts=[-20.0, 10.0], scan=6, seed=11),
             serve.JobSpec(job_id=1, vol=32, P=64, tilts=[0.0, 30.0, -40.0], scan=8, seed=12),
             serve.JobSpec(job_id=2, vol=64, P=128, tilts=[15.0, -15.0], scan=5, seed=13)]
    big = max(serve.footprint_bytes(s, lib.batch_slots(s.P)) for s in specs)
    alone_err, alone_state = {}, {}
    for s in specs:
        print('case: serve alone job', s.job_id, flush=True)
        e, st, _ = _run([s], [s.job_id] * 3, 1, 4 * big)
        alone_err.update(e)
        alone_state.update(st)
    print('case: serve mix', flush=True)
    order = [0, 1, 2, 1, 0, 2, 2, 0, 1]
    e, st, rep = _run(specs, order, 3, int(1.2 * big))      # the largest job never fits beside another
    print(rep, flush=True)
    assert rep['evictions'] >= 1 and rep['admissions'] > len(specs)
    for s in specs:
        assert e[s.job_id] == alone_err[s.job_id], s.job_id
        for a, b in zip(st[s.job_id], alone_state[s.job_id]):
            if isinstance(a, list):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y)
            else:
                assert np.array_equal(a, b)
