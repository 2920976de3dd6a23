"""Host driver of the reconstruction loop on the device-resident tilt series.

Restates src/Run_MSET_recon.m:102-292 (device == 2 flow) over `lib.Recon`: per iteration, a snapshot of
(rec, probes, positions) (:113-115), the update pass over every tilt (:119-187, one call), the error pass
(:189-239), the backtracking line search on the object step (:241-257: on a worse mean error restore the
snapshot and scale every step by bls_parameter; stop once step(1) < bls_parameter^5 * step(1) at start),
and the optional TV regulariser on the accepted volume (:266-275).  Every array stays on the device; the
host only sees the per-tilt errors.

`recon` may be any object with the lib.Recon methods used here (set_steps, iterate, error_pass, snapshot,
restore, tv), which is how the control flow is tested without a GPU.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class History:
    mean_error: list = field(default_factory=list)         # mat_save.mean_error / STEM_data.errorlist
    total_mean_error: list = field(default_factory=list)   # with the TV term (use_TV only)
    bls: list = field(default_factory=list)                # (iteration, rejected mean error, object step before)
    steps: list = field(default_factory=list)              # step sizes of each accepted iteration
    stopped_early: bool = False                            # the line search ran out (break at :259-262)


class Driver:
    def __init__(self, recon, step_size, n_tilts, shape, bls_parameter=0.1, use_positivity=True, use_tv=False,
                 tv_lambda=0.0, tv_iter=300, path=1, batch=256, use_graph=False):
        """shape: (nz, ny, nx) of rec, for the TV error normalisation (size(rec, 1) * size(rec, 2))."""
        self.r = recon
        self.original = np.asarray(step_size, np.float64).copy()
        self.steps = self.original.copy()
        self.n_tilts = int(n_tilts)
        self.shape = tuple(shape)
        self.bls = float(bls_parameter)
        self.positivity = bool(use_positivity)
        self.use_tv = bool(use_tv)
        self.tv_lambda = float(tv_lambda)
        self.tv_iter = int(tv_iter)
        self.path, self.batch, self.use_graph = path, int(batch), bool(use_graph)
        self.old_mean_error = 1e31

    def _exhausted(self):
        return self.steps[0] < self.bls ** 5 * self.original[0]

    def run(self, n_iter):
        h = History()
        for i in range(int(n_iter)):
            self.r.snapshot()                                        # :113-115
            while True:
                self.r.set_steps(self.steps)
                self.r.iterate(self.path, self.batch, self.use_graph, self.positivity, want_error=False)
                errs = np.asarray(self.r.error_pass(), np.float64)   # :189-239
                if self.old_mean_error > errs.mean():                # :241-243
                    break
                h.bls.append((i, float(errs.mean()), float(self.steps[0])))
                self.r.restore()                                     # :245-247
                self.steps = self.bls * self.steps                   # :249
                if self._exhausted():                                # :252-255
                    break
            if self._exhausted():                                    # :259-262
                h.stopped_early = True
                break
            mean_error = float(errs.mean())
            if self.use_tv:                                          # :266-272
                tv_err = self.r.tv(self.tv_lambda, self.tv_iter)
                h.total_mean_error.append(mean_error + tv_err / self.n_tilts / self.shape[2] / self.shape[1])
            h.mean_error.append(mean_error)
            h.steps.append(self.steps.copy())
            self.old_mean_error = mean_error                         # :292
        return h
