"""Synthetic tilt-series workloads for the benchmark configurations.

The reference ships one small example and no simulator (its data came from an external PRISM-type
code), so measured data here are produced by this package's own forward model: a potential of Gaussian
atoms on FCC sites is rotated to each tilt and propagated on the device (mset_simulate), and the
intensities are Poisson-sampled at the stated dose.  Everything is seeded.
"""
import numpy as np

from . import physics

# sigma = 0.3 A Gaussian atoms, peak 1.0 (Au) / 0.8 (Cu), 50/50, on an FCC lattice with a = 4.0 A
ATOM_SIGMA_A = 0.3
FCC_A = 4.0

CONFIGS = {
    # plain method, CPU-runnable: 24^3 at 0.5 A in a 64 x 64 field, 8 x 8 scan at 0.8 A, 3 tilts, 2 A slices
    'C1': dict(vol=24, P=64, dx=0.5, kV=300.0, alpha=25.0, defocus=0.0, scan=(8, 8), step_A=0.8, jitter=0.0,
               tilts=[-30.0, 0.0, 30.0], slice_A=2.0, dose=1e5, occupancy=0.25),
    # plain method on the GPU: 64^3 at 0.4 A in a 128 x 128 field, 32 x 32 scan (2 px + U[-0.5, 0.5] px),
    # 11 tilts -50..50 deg, 1.6 A slices, 5e4 e/A^2
    'C2': dict(vol=64, P=128, dx=0.4, kV=300.0, alpha=21.0, defocus=0.0, scan=(32, 32), step_A=0.8, jitter=0.5,
               tilts=list(np.linspace(-50.0, 50.0, 11)), slice_A=1.6, dose=5e4, occupancy=0.08),
    # batched: 128^3 at 0.4 A in a 256 x 256 field, 21 tilts -60..60 deg, each with its own square scan of
    # side U{24..48} (2 px steps), 1.6 A slices (32 planes)
    'C3': dict(vol=128, P=256, dx=0.4, kV=300.0, alpha=21.0, defocus=0.0, scan=(24, 48), step_A=0.8, jitter=0.5,
               tilts=list(np.linspace(-60.0, 60.0, 21)), slice_A=1.6, dose=5e4, occupancy=0.08, uneven=True),
    # adjoint verification: 32^3 at 0.4 A in 64 x 64, 4 x 4 scan, 2 tilts, no binning
    'C4': dict(vol=32, P=64, dx=0.4, kV=300.0, alpha=21.0, defocus=0.0, scan=(4, 4), step_A=1.6, jitter=0.0,
               tilts=[-20.0, 20.0], slice_A=0.4, dose=1e5, occupancy=0.08),
}


def fcc_volume(n, dx, occupancy, rng, sigma_A=ATOM_SIGMA_A):
    """n^3 voxels at dx A: Gaussian atoms on randomly occupied FCC sites, [z, y, x] float32."""
    L = n * dx
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * FCC_A
    m = int(np.ceil(L / FCC_A)) + 1
    cells = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing='ij'), -1).reshape(-1, 3) * FCC_A
    sites = (cells[:, None, :] + basis[None]).reshape(-1, 3)
    margin = 3 * sigma_A
    sites = sites[np.all((sites >= margin) & (sites <= L - margin), axis=1)]
    occ = rng.random(len(sites)) < occupancy
    sites = sites[occ]
    peaks = np.where(rng.random(len(sites)) < 0.5, 1.0, 0.8)
    V = np.zeros((n, n, n), np.float64)
    r = int(np.ceil(4 * sigma_A / dx))
    ax = np.arange(-r, r + 1)
    for (x, y, z), pk in zip(sites, peaks):
        c = np.floor(np.array([x, y, z]) / dx).astype(int)
        ix, iy, iz = [np.clip(c[i] + ax, 0, n - 1) for i in range(3)]
        gx = np.exp(-((ix + 0.5) * dx - x) ** 2 / (2 * sigma_A ** 2))
        gy = np.exp(-((iy + 0.5) * dx - y) ** 2 / (2 * sigma_A ** 2))
        gz = np.exp(-((iz + 0.5) * dx - z) ** 2 / (2 * sigma_A ** 2))
        V[np.ix_(iz, iy, ix)] += pk * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return V.astype(np.float32), len(sites)


def pad_lateral(V, P):
    """Centre an n^3 volume in a P x P x n field (the probe side must equal the lateral volume side)."""
    n = V.shape[0]
    out = np.zeros((V.shape[0], P, P), np.float32)
    o = (P - V.shape[1]) // 2
    out[:, o:o + V.shape[1], o:o + V.shape[2]] = V
    return out


class Workload:
    """Inputs of one benchmark configuration: truth, probe, propagator, per-tilt scan lists and data."""

    def __init__(self, name, seed=None, scan=None, tilts=None):
        cfg = dict(CONFIGS[name])
        if scan is not None:
            cfg['scan'] = scan
        if tilts is not None:
            cfg['tilts'] = tilts
        self.name, self.cfg = name, cfg
        self.seed = int(name[1:]) if seed is None else seed
        rng = np.random.default_rng(self.seed)
        E0 = cfg['kV'] * 1e3
        self.lam = physics.electron_wavelength(E0)
        self.sigma = physics.interaction_constant(E0)
        self.dx = cfg['dx']
        self.P = cfg['P']
        self.binning = max(1, int(round(cfg['slice_A'] / cfg['dx'])))
        V, self.n_atoms = fcc_volume(cfg['vol'], cfg['dx'], cfg['occupancy'], rng)
        self.truth = pad_lateral(V, self.P)
        self.nz = self.truth.shape[0]
        self.L = -(-self.nz // self.binning)
        self.prop = physics.propagator(self.P, self.dx, self.lam, self.binning)
        step_px = cfg['step_A'] / cfg['dx']
        cen = np.floor((self.P + 1) / 2 + 0.5)
        self.angles = [(0.0, float(t), 0.0) for t in cfg['tilts']]
        self.xlists, self.ylists, self.n_scans = [], [], []
        for _ in self.angles:
            if cfg.get('uneven'):
                side = int(rng.integers(cfg['scan'][0], cfg['scan'][1] + 1))   # square grid, side U{lo..hi}
                nsx = nsy = side
            else:
                nsx, nsy = cfg['scan']
            origin = cen - (nsx - 1) * step_px / 2
            x, y = physics.scan_grid(nsx, nsy, step_px, origin)
            j = cfg['jitter']
            if j:
                x = (x + rng.uniform(-j, j, x.shape)).astype(np.float32)
                y = (y + rng.uniform(-j, j, y.shape)).astype(np.float32)
            self.xlists.append(x)
            self.ylists.append(y)
            self.n_scans.append((nsx, nsy))
        self.n_scan = self.n_scans[0]
        self.counts = cfg['dose'] * cfg['step_A'] ** 2
        self.probe = physics.normalise_probe(physics.probe(self.P, self.dx, self.lam, cfg['alpha'], cfg['defocus']),
                                             self.counts).astype(np.complex64)
        self.rng = rng
        self.measured = None

    @property
    def npos(self):
        """positions of the first tilt (every tilt of an even configuration)"""
        return self.n_scan[0] * self.n_scan[1]

    def npos_of(self, j):
        return self.n_scans[j][0] * self.n_scans[j][1]

    def positions(self):
        return sum(self.npos_of(j) for j in range(len(self.angles)))

    def pos_slices(self):
        return self.positions() * self.L

    def simulate(self, ctx):
        """Poisson data for every tilt from the device forward model (one mset_ctx, reused)."""
        self.measured = []
        for j, ang in enumerate(self.angles):
            zeros = np.zeros((self.npos_of(j), self.P, self.P), np.float32)
            ctx.upload(self.truth, self.probe, self.prop, self.xlists[j], self.ylists[j], zeros, self.sigma, self.dx,
                       slice_binning=self.binning, n_scan=self.n_scans[j])
            ctx.rotate_resident(ang, inverse=True)
            I = ctx.simulate()
            self.measured.append(self.rng.poisson(np.maximum(I, 0).astype(np.float64)).astype(np.float32))
        return self.measured
