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


def test_c5_job_draw_alone_and_in_the_mix():
    """The 16-job serving draw (volumes 32..128^3, probes 64..256^2, 8..30 tilts, 16^2..40^2 scans) with 24
    requests arriving as a Poisson stream at 20/s, under a 60 GB budget (about three of the largest jobs):
    every job's errors and final state are the same, bit for bit, when its requests run alone and when they
    run in the mix, where requests overlap on the device and jobs are evicted and re-admitted."""
    import time
    from mset_amd import serve
    specs = serve.draw_jobs(16, 5)
    offsets, jobs = serve.poisson_arrivals(24, 20.0, 16, 6)
    count = np.bincount(jobs, minlength=len(specs))
    alone_err, alone_state = {}, {}
    for s in specs:
        if count[s.job_id] == 0:
            continue
        print('case: C5 job %d alone (P=%d, %d tilts, scan %d, %d requests)' % (s.job_id, s.P, len(s.tilts), s.scan,
                                                                               count[s.job_id]), flush=True)
        srv = serve.Server([s], workers=1, batch=512)
        reqs = [srv.submit(s.job_id) for _ in range(count[s.job_id])]
        for r in reqs:
            assert r.done.wait(300), 'request did not finish'
        assert all(not r.failure for r in reqs), [r.failure for r in reqs]
        alone_err[s.job_id] = list(srv.jobs[s.job_id].errors)
        alone_state[s.job_id] = srv.state_of(s.job_id)
        srv.close()
    print('case: C5 mix, 24 requests, 4 workers, 60 GB budget', flush=True)
    srv = serve.Server(specs, workers=4, budget_bytes=int(60e9), batch=512)
    t0 = time.perf_counter()
    reqs = []
    for off, j in zip(offsets, jobs):
        now = time.perf_counter() - t0
        if off > now:
            time.sleep(off - now)
        reqs.append(srv.submit(int(j)))
    for r in reqs:
        assert r.done.wait(600), 'request did not finish'
    assert all(not r.failure for r in reqs), [r.failure for r in reqs]
    mix_err = {s.job_id: list(srv.jobs[s.job_id].errors) for s in specs}
    mix_state = {s.job_id: srv.state_of(s.job_id) for s in specs if count[s.job_id]}
    rep = srv.report()
    srv.close()
    print('C5 mix report', {k: rep[k] for k in ('requests', 'p50_ms', 'p99_ms', 'throughput', 'evictions',
                                               'admissions', 'overlap_s', 'max_concurrent')}, flush=True)
    assert rep['evictions'] >= 1 and rep['max_concurrent'] >= 2 and rep['overlap_s'] > 0
    for j in alone_err:
        assert mix_err[j] == alone_err[j], j
        a, b = mix_state[j], alone_state[j]
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), j
        for x, y in zip(a[2] + a[3], b[2] + b[3]):
            assert np.array_equal(x, y), j
