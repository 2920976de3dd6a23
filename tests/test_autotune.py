"""The autotuner's control flow against a scripted tilt series (no GPU): every candidate is timed from the
snapshot, the state is restored after each, the fastest grid is left set and the choice is cached."""
import json
import time

from mset_amd import autotune


class FakeRecon:
    shape = (8, 64, 64)

    def __init__(self, cost, shape=(8, 64, 64)):
        self.cost = cost                # (slots, batch) or (engine, slots, batch) -> seconds per pass
        self.shape = shape
        self.npos = [24, 24]
        self.slots = 0
        self.engine = 0
        self.calls = []
        self.dirty = False

    def set_slot_limit(self, s):
        self.slots = s
        self.calls.append(('slots', s))

    def set_engine(self, e):
        self.engine = e
        self.calls.append(('engine', e))

    def snapshot(self):
        self.calls.append(('snapshot',))

    def restore(self):
        self.dirty = False
        self.calls.append(('restore',))

    def iterate(self, path, batch, use_graph=False, want_error=True):
        assert not use_graph
        time.sleep(self.cost.get((self.engine, self.slots, batch), self.cost.get((self.slots, batch), 0.0)))
        self.dirty = True
        self.calls.append(('iterate', self.slots, batch))

    def profile(self, enable=True):
        pass

    def profile_read(self, name):
        self.calls.append(('read', name))
        return 1, 0.0


def test_tuner_times_every_candidate_restores_and_caches(tmp_path):
    cost = {(24, 8): 0.03, (24, 24): 0.02, (12, 8): 0.04, (12, 24): 0.005}
    r = FakeRecon(cost)
    path = str(tmp_path / 'tune.json')
    best = autotune.tune_recon(r, batches=(8, 24), slots=[24, 12], engines=(0,), cache_path=path)
    assert (best['slots'], best['batch']) == (12, 24)
    assert r.slots == 12 and not r.dirty
    timed = [c for c in r.calls if c[0] == 'iterate']
    # one untimed sizing pass at the largest batch per grid, then one pass per batch
    assert timed == [('iterate', 24, 24), ('iterate', 24, 8), ('iterate', 24, 24),
                     ('iterate', 12, 24), ('iterate', 12, 8), ('iterate', 12, 24)]
    # every pass is followed by a restore (timed passes after reading every timer)
    calls = [c for c in r.calls if c[0] != 'read']
    for i, c in enumerate(calls):
        if c[0] == 'iterate':
            assert calls[i + 1] == ('restore',)
    assert r.calls.count(('read', 'rotate')) == r.calls.count(('read', 'position_batch')) == 4
    # in-memory cache: no new passes; the disk cache holds the same choice
    n = len(r.calls)
    assert autotune.tune_recon(r, batches=(8, 24), slots=[24, 12], engines=(0,)) is best
    assert not any(c[0] == 'iterate' for c in r.calls[n:])
    with open(path) as fh:
        disk = json.load(fh)
    assert [(v['slots'], v['batch']) for v in disk.values()] == [(12, 24)]


def test_tuner_sweeps_the_plane_routine_at_256():
    """At 256 x 256 both plane routines (separable and general) and both workgroup shapes (32- and 16-line
    tiles) are candidates; the fastest is left set."""
    from mset_amd import lib
    cost = {(lib.ENGINE_AUTO, 24, 8): 0.03, (lib.ENGINE_ALT_SHAPE, 24, 8): 0.02, (lib.ENGINE_GENERAL, 24, 8): 0.01}
    r = FakeRecon(cost, shape=(16, 256, 256))
    best = autotune.tune_recon(r, batches=(8,), slots=[24])
    assert len(best['table']) == 3 and best['engine'] == lib.ENGINE_GENERAL and r.engine == lib.ENGINE_GENERAL
    assert autotune.engine_name(lib.ENGINE_GENERAL | lib.ENGINE_ALT_SHAPE) == 'general routine, alternative shape'


def test_tuner_sweeps_the_workgroup_shape_at_every_size():
    """Below 256 the candidates are the two workgroup shapes of the plane size; the grid candidates start with
    0, every resident workgroup of the engine's own shape."""
    from mset_amd import lib
    for n in (16, 24, 32, 48, 64, 96, 128):
        assert autotune.engine_candidates(n) == (lib.ENGINE_AUTO, lib.ENGINE_ALT_SHAPE)
    cost = {(lib.ENGINE_AUTO, 0, 8): 0.02, (lib.ENGINE_ALT_SHAPE, 0, 8): 0.005}
    r = FakeRecon(cost, shape=(16, 128, 128))
    best = autotune.tune_recon(r, batches=(8,), slots=[0])
    assert best['engine'] == lib.ENGINE_ALT_SHAPE and r.engine == lib.ENGINE_ALT_SHAPE and r.slots == 0
