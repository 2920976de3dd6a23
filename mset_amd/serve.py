"""Serving loop: many reconstruction jobs sharing one device.

Each job is a tilt series with its own device-resident state (a lib.Recon: volume, probes, positions and
patterns stay on the device between iterations).  A request asks for one forward + gradient iteration of
one job (an update pass over all its tilts, followed by its loss).  Requests are served continuously:

  * admission by device memory: a job becomes resident when its estimated footprint fits the budget;
    when it does not, idle resident jobs are evicted (least recently used first) -- their state is
    copied to the host and restored exactly on re-admission.  The earliest request that waits for
    memory holds back the resident jobs it needs evicted (no new requests start on them), so a large
    job is not starved by a stream of requests for small resident ones;
  * up to `workers` requests run at once, each job on its own HIP stream.  A request starting while k
    others run gets 1/2^ceil(log2(k+1)) of the device's resident workgroups as its persistent grid (all
    of them when it runs alone), so the persistent kernels of concurrent jobs co-reside on the CUs and
    their work overlaps instead of queueing behind one another.  Gradient slabs are assigned by position,
    not by workgroup, so the grid share never changes a result.  One job never runs two requests at a
    time, so its iterations stay ordered;
  * every request's latency (arrival to completion) is recorded; p50 / p99 and the aggregate
    probe-position-slice throughput are reported.

Kernels are deterministic and every job owns its workspace, so a job's results do not depend on what
else runs: the same alone and in the mix.
"""
import argparse
import json
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

PROBE_SIDES = (64, 128, 256)
VOLUME_SIDES = (32, 64, 96, 128)


@dataclass
class JobSpec:
    """One reconstruction: P x P probe, a vol^3 volume padded to P x P x vol, tilts with square scans."""
    job_id: int
    vol: int
    P: int
    tilts: list
    scan: int
    seed: int
    dx: float = 0.4
    binning: int = 4
    steps: tuple = (100.0, 0.0, 0.0)

    @property
    def L(self):
        return -(-self.vol // min(self.binning, self.vol))

    def positions(self):
        return len(self.tilts) * self.scan * self.scan

    def pos_slices(self):
        return self.positions() * self.L


def draw_jobs(n_jobs, seed, vol_sides=VOLUME_SIDES, probe_sides=PROBE_SIDES, tilts=(8, 30), scan=(16, 40)):
    """Job sizes of the serving configuration: volume and probe sides drawn from the given sets (probe side
    at least the volume side: the probe window covers the whole lateral field), tilt count and scan side
    uniform in the given ranges."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        vol = int(rng.choice(vol_sides))
        P = int(rng.choice([p for p in probe_sides if p >= vol]))
        nt = int(rng.integers(tilts[0], tilts[1] + 1))
        angles = list(np.linspace(-60.0, 60.0, nt)) if nt > 1 else [0.0]
        side = int(rng.integers(scan[0], scan[1] + 1))
        jobs.append(JobSpec(job_id=j, vol=vol, P=P, tilts=angles, scan=side, seed=int(rng.integers(1 << 30))))
    return jobs


def footprint_bytes(spec, slots):
    """Device bytes a resident job holds (the allocations of capi.hip for its shapes)."""
    n2 = spec.P * spec.P
    nv = n2 * spec.vol
    L, nt, npos = spec.L, len(spec.tilts), spec.scan * spec.scan
    per_tilt = nv * 4 + npos * n2 * 4 * 2 + n2 * 8 * 2 + npos * 24     # V, patterns (+ staging), probe, prop
    shared = (slots * (L + 2) * n2 * 8 + slots * L * n2 * 4 + slots * n2 * 8 + L * n2 * 8 * 2 + n2 * 24
              + nv * 8 * 2)                                            # save, gslab, pslab, trans, G, rotation
    if spec.P >= 256:
        shared += slots * 2 * n2 * 8                                   # plane buffers
    return int(1.1 * (nt * per_tilt + shared + 2 * nv * 4))            # + rec and its snapshot, 10 % margin


def build_inputs(spec):
    """Synthetic, seeded inputs of one job: volume, probe, propagator, per-tilt scans and patterns."""
    from . import physics, synth
    rng = np.random.default_rng(spec.seed)
    lam = physics.electron_wavelength(300e3)
    sigma = physics.interaction_constant(300e3)
    V, _ = synth.fcc_volume(spec.vol, spec.dx, 0.08, rng)
    rec = synth.pad_lateral(V, spec.P)
    prop = physics.propagator(spec.P, spec.dx, lam, spec.binning)
    probe = physics.normalise_probe(physics.probe(spec.P, spec.dx, lam, 21.0, 0.0), 1e4).astype(np.complex64)
    step = max(0.5, min(2.0, 0.5 * spec.P / spec.scan))
    cen = np.floor((spec.P + 1) / 2 + 0.5)
    x, y = physics.scan_grid(spec.scan, spec.scan, step, cen - (spec.scan - 1) * step / 2)
    # a bank of positive patterns, repeated over the positions (the serving loop measures time, not fit)
    bank = rng.poisson(1e4 / (spec.P * spec.P), (16, spec.P, spec.P)).astype(np.float32)
    npos = spec.scan * spec.scan
    M = bank[np.arange(npos) % len(bank)]
    tilts = [dict(angles=(0.0, float(a), 0.0), xl=x, yl=y, M=M) for a in spec.tilts]
    return dict(rec=rec, prop=prop, probe=probe, sigma=sigma, tilts=tilts)


def make_recon(spec, state=None):
    """A lib.Recon for the job, from its synthetic inputs or from a saved (rec, probes, xls, yls) state."""
    from . import lib
    inp = build_inputs(spec)
    rec = inp['rec'] if state is None else state[0]
    r = lib.Recon(rec, inp['prop'], len(spec.tilts), inp['sigma'], spec.dx, slice_binning=spec.binning,
                  step_size=spec.steps)
    for j, t in enumerate(inp['tilts']):
        probe = inp['probe'] if state is None else state[1][j]
        xl = t['xl'] if state is None else state[2][j]
        yl = t['yl'] if state is None else state[3][j]
        r.set_tilt(j, t['angles'], probe, xl, yl, t['M'], n_scan=(spec.scan, spec.scan))
    return r


@dataclass
class Request:
    job_id: int
    arrival: float
    seq: int
    start: float = 0.0
    end: float = 0.0
    slots: int = 0                # the persistent grid this request ran with
    concurrent: int = 0           # requests running (this one included) when it started
    error: float = float('nan')
    admitted: bool = False        # this request made its job resident
    failure: str = ''             # set when the request failed (the job's state is unchanged)
    done: threading.Event = field(default_factory=threading.Event)


class _Job:
    def __init__(self, spec, nbytes):
        self.spec = spec
        self.bytes = nbytes
        self.recon = None
        self.saved = None          # host copy of an evicted job's state
        self.busy = False
        self.last_used = 0.0
        self.errors = []


class Server:
    def __init__(self, specs, workers=4, budget_bytes=None, device=0, batch=512, use_graph=True,
                 recon_factory=make_recon, slots_of=None, clock=time.perf_counter):
        """specs: the jobs; budget_bytes: device memory the jobs may hold (default 90 % of what is free)."""
        if slots_of is None:
            from . import lib
            slots_of = lambda n: lib.batch_slots(n, device)
        if budget_bytes is None:
            from . import lib
            budget_bytes = int(0.9 * lib.device_memory(device)[0])
        self.budget = int(budget_bytes)
        self.slots_of = {s.P: slots_of(s.P) for s in specs}
        self.jobs = {s.job_id: _Job(s, footprint_bytes(s, self.slots_of[s.P])) for s in specs}
        for j in self.jobs.values():
            if j.bytes > self.budget:
                raise ValueError('job %d needs ~%.1f GB, more than the budget %.1f GB'
                                 % (j.spec.job_id, j.bytes / 1e9, self.budget / 1e9))
        self.factory = recon_factory
        self.batch, self.use_graph = batch, use_graph
        self.clock = clock
        self.cv = threading.Condition()
        self.pending = []
        self.completed = []
        self.resident_bytes = 0
        self.running = 0
        self.workers = int(workers)
        self.busy_log = []            # (time, requests running) at every start and end
        self.evictions = 0
        self.closing = False
        self.seq = 0
        self.threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(int(workers))]
        for t in self.threads:
            t.start()

    # ------------------------------------------------------------------ client side
    def submit(self, job_id, arrival=None):
        with self.cv:
            req = Request(job_id=job_id, arrival=self.clock() if arrival is None else arrival, seq=self.seq)
            self.seq += 1
            self.pending.append(req)
            self.cv.notify_all()
        return req

    def close(self):
        with self.cv:
            self.closing = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        for j in self.jobs.values():
            if j.recon is not None:
                j.recon.close()
                j.recon = None

    def state_of(self, job_id):
        """(rec, probes, xlists, ylists) of a job (resident or saved)."""
        j = self.jobs[job_id]
        with self.cv:
            while j.busy:
                self.cv.wait()
            if j.recon is not None:
                return j.recon.get()
            return j.saved

    # ------------------------------------------------------------------ scheduling (cv held)
    def _pick(self):
        """The earliest pending request whose job is idle and resident, or admissible after evicting idle
        jobs; returns (request, jobs to evict) or (None, None).  The first request that cannot be
        admitted yet reserves the resident jobs (least recently used first) whose eviction would make
        room for it: later requests of those jobs wait."""
        held = set()
        for i, req in enumerate(self.pending):
            job = self.jobs[req.job_id]
            if job.busy or req.job_id in held:
                continue
            if job.recon is not None:
                return self.pending.pop(i), []
            need = job.bytes - (self.budget - self.resident_bytes)
            evict = []
            if need > 0:
                idle = sorted((j for j in self.jobs.values() if j.recon is not None and not j.busy),
                              key=lambda j: j.last_used)
                for j in idle:
                    if need <= 0:
                        break
                    evict.append(j)
                    need -= j.bytes
            if need <= 0:
                return self.pending.pop(i), evict
            if not held:
                need = job.bytes - (self.budget - self.resident_bytes)
                for j in sorted((j for j in self.jobs.values() if j.recon is not None), key=lambda j: j.last_used):
                    if need <= 0:
                        break
                    held.add(j.spec.job_id)
                    need -= j.bytes
        return None, None

    def _share(self, P, running):
        """Persistent grid of a request that starts with `running` requests on the device (itself included):
        the resident workgroups divided by the next power of two, so requests that start together fit on the
        device side by side, and every share divides the slab count.  A request keeps the share it started
        with for its whole pass."""
        div = 1
        while div < running:
            div *= 2
        return max(1, self.slots_of[P] // div)

    def _worker(self):
        while True:
            with self.cv:
                while True:
                    if self.closing and not self.pending:
                        return
                    req, evict = self._pick()
                    if req is not None:
                        break
                    self.cv.wait()
                job = self.jobs[req.job_id]
                job.busy = True
                for j in evict:
                    j.busy = True
                admit = job.recon is None
                if admit:
                    self.resident_bytes += job.bytes
                for j in evict:
                    self.resident_bytes -= j.bytes
                self.running += 1
                req.concurrent = self.running
                req.slots = self._share(job.spec.P, self.running)
                self.busy_log.append((self.clock(), self.running))
            req.start = self.clock()
            try:
                for j in evict:                                    # outside the lock: device work
                    j.saved = j.recon.get()
                    j.recon.close()
                    j.recon = None
                if admit:
                    job.recon = self.factory(job.spec, job.saved)
                    job.saved = None
                    req.admitted = True
                job.recon.set_slot_limit(req.slots)
                req.error = job.recon.iterate(1, self.batch, self.use_graph, True, want_error=True)
            except Exception as e:                                  # report it on the request, keep serving
                req.failure = '%s: %s' % (type(e).__name__, e)
            req.end = self.clock()
            with self.cv:
                if admit and job.recon is None:                    # admission failed: release the reservation
                    self.resident_bytes -= job.bytes
                job.busy = False
                job.last_used = req.end
                if not req.failure:
                    job.errors.append(req.error)
                for j in evict:
                    j.busy = False
                    if j.recon is None:
                        self.evictions += 1
                    else:                                          # its eviction failed: it is still resident
                        self.resident_bytes += j.bytes
                self.running -= 1
                self.busy_log.append((req.end, self.running))
                self.completed.append(req)
                self.cv.notify_all()
            req.done.set()

    # ------------------------------------------------------------------ report
    def report(self):
        lat = np.array([r.end - r.arrival for r in self.completed]) * 1e3
        t0 = min(r.arrival for r in self.completed)
        t1 = max(r.end for r in self.completed)
        units = sum(self.jobs[r.job_id].spec.pos_slices() for r in self.completed)
        # time with two or more requests on the device, from the start / end log
        log = sorted(self.busy_log)
        overlap = sum(b[0] - a[0] for a, b in zip(log, log[1:]) if a[1] >= 2)
        return dict(requests=len(self.completed), p50_ms=float(np.percentile(lat, 50)),
                    p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                    throughput=units / (t1 - t0), unit='probe-position·slices/s', makespan_s=t1 - t0,
                    admissions=sum(r.admitted for r in self.completed), evictions=self.evictions,
                    budget_gb=self.budget / 1e9, overlap_s=overlap,
                    max_concurrent=max((r.concurrent for r in self.completed), default=0),
                    failures=sum(bool(r.failure) for r in self.completed))


def poisson_arrivals(n_requests, rate, n_jobs, seed):
    """Arrival offsets (s) of a Poisson stream of the given rate and the job of each request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    return t - t[0], rng.integers(0, n_jobs, n_requests)


def main():
    ap = argparse.ArgumentParser(description='serving loop over synthetic reconstruction jobs')
    ap.add_argument('--jobs', type=int, default=16)
    ap.add_argument('--requests', type=int, default=64)
    ap.add_argument('--rate', type=float, default=20.0, help='mean arrivals per second (Poisson)')
    ap.add_argument('--workers', type=int, default=4)
    ap.add_argument('--seed', type=int, default=5)
    ap.add_argument('--budget-gb', type=float, default=None)
    ap.add_argument('--max-probe', type=int, default=256)
    ap.add_argument('--max-tilts', type=int, default=30)
    ap.add_argument('--max-scan', type=int, default=40)
    args = ap.parse_args()
    specs = draw_jobs(args.jobs, args.seed, probe_sides=[p for p in PROBE_SIDES if p <= args.max_probe],
                      vol_sides=[v for v in VOLUME_SIDES if v <= args.max_probe],
                      tilts=(min(8, args.max_tilts), args.max_tilts), scan=(min(16, args.max_scan), args.max_scan))
    offsets, jobs_of = poisson_arrivals(args.requests, args.rate, args.jobs, args.seed + 1)
    srv = Server(specs, workers=args.workers,
                 budget_bytes=None if args.budget_gb is None else int(args.budget_gb * 1e9))
    t0 = time.perf_counter()
    reqs = []
    for off, jid in zip(offsets, jobs_of):
        now = time.perf_counter() - t0
        if off > now:
            time.sleep(off - now)
        reqs.append(srv.submit(int(jid)))
        print('request %d job %d' % (len(reqs), jid), file=sys.stderr, flush=True)
    for r in reqs:
        r.done.wait()
    srv.close()
    out = srv.report()
    out['jobs'] = [dict(job=s.job_id, vol=s.vol, P=s.P, tilts=len(s.tilts), scan=s.scan) for s in specs]
    print(json.dumps(out))


if __name__ == '__main__':
    main()
