"""Control flow of the reconstruction driver (mset_amd/recon.py, Run_MSET_recon.m:102-292) against a
scripted stand-in for the device tilt series: which passes run, when the line search restores and shrinks
the steps, when it gives up, and the TV error normalisation."""
import numpy as np

from mset_amd.recon import Driver


class FakeRecon:
    """Errors of successive error passes are scripted; every call is logged."""

    def __init__(self, errors, tv_error=0.0):
        self.errors = list(errors)
        self.log = []
        self.tv_error = tv_error

    def snapshot(self):
        self.log.append('snapshot')

    def restore(self):
        self.log.append('restore')

    def set_steps(self, s):
        self.log.append(('steps', tuple(np.round(s, 12))))

    def iterate(self, path, batch, use_graph, positivity, want_error=True):
        self.log.append(('iterate', positivity))

    def error_pass(self):
        self.log.append('errors')
        return np.full(3, self.errors.pop(0))

    def tv(self, lam, n_iter):
        self.log.append(('tv', lam, n_iter))
        return self.tv_error


def test_accepts_decreasing_errors_without_line_search():
    r = FakeRecon([5.0, 4.0, 3.0])
    h = Driver(r, (100.0, 1.0, 2.0), 3, (8, 16, 16)).run(3)
    assert h.mean_error == [5.0, 4.0, 3.0] and h.bls == [] and not h.stopped_early
    assert r.log.count('snapshot') == 3 and 'restore' not in r.log


def test_worse_error_restores_and_scales_every_step():
    r = FakeRecon([5.0, 6.0, 4.5])                  # iteration 2: first attempt worse, second better
    h = Driver(r, (100.0, 1.0, 2.0), 3, (8, 16, 16), bls_parameter=0.1).run(2)
    assert h.mean_error == [5.0, 4.5]
    assert h.bls == [(1, 6.0, 100.0)]
    assert np.allclose(h.steps[1], [10.0, 0.1, 0.2])
    assert r.log.count('restore') == 1
    # the steps stay reduced for the next iteration (STEM_data.step_size is not reset)
    r2 = FakeRecon([5.0, 6.0, 4.5, 4.0])
    h2 = Driver(r2, (100.0, 1.0, 2.0), 3, (8, 16, 16)).run(3)
    assert np.allclose(h2.steps[2], [10.0, 0.1, 0.2])


def test_line_search_gives_up_below_bls_to_the_fifth():
    r = FakeRecon([5.0] + [9.0] * 10)
    h = Driver(r, (100.0, 0.0, 0.0), 3, (8, 16, 16), bls_parameter=0.5).run(4)
    # rejected at 100, 50, 25, 12.5, 6.25 and 3.125 (not below 0.5^5 * 100 = 3.125); 1.5625 is below: stop
    assert h.mean_error == [5.0] and h.stopped_early
    assert len(h.bls) == 6 and r.log.count('restore') == 6


def test_tv_term_normalisation():
    r = FakeRecon([2.0], tv_error=48.0)
    h = Driver(r, (1.0, 0.0, 0.0), 3, (8, 16, 4), use_tv=True, tv_lambda=0.02).run(1)
    # total = mean + TV_error / N_angle / size(rec, 1) / size(rec, 2), size(rec, 1) = nx = 4, size(rec, 2) = 16
    assert h.total_mean_error == [2.0 + 48.0 / 3 / 4 / 16]
    assert ('tv', 0.02, 300) in r.log
