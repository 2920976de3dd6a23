"""Seeded synthetic states shared by the CPU and GPU tests (inputs built with the oracle's physics)."""
import numpy as np

from oracle import oracle as O


def small_case(n=16, nz=6, binning=2, npos=3, strength=0.3, seed=0, alpha=60.0, defocus=20.0, dx=0.5,
               spread=2.0):
    """A strong-phase state: phases ~strength rad per plane, random positive patterns.

    Returns a dict with the volume V [nz, n, n] (float64), probe and propagators in MATLAB layout,
    scan positions (MATLAB 1-based pixels), measured patterns [npos, n, n] and the physics scalars."""
    rng = np.random.default_rng(seed)
    lam, sig = O.wavelength(300e3), O.interaction_sigma(300e3)
    V = rng.uniform(0, 1, (nz, n, n)) * strength / (sig * dx * binning)
    p = O.probe(n, dx, lam, alpha, C10=defocus)
    p = O.probe_normalise(p, 1e3)
    c = n // 2 + 1
    xl = (c + rng.uniform(-spread, spread, npos)).astype(np.float32)
    yl = (c + rng.uniform(-spread, spread, npos)).astype(np.float32)
    M = (rng.uniform(0.5, 1.5, (npos, n, n)) * 1e3 / n ** 2).astype(np.float32)
    prop = O.propagator(n, dx, lam, binning)
    bprop = O.propagator(n, dx, lam, binning, back=True)
    return dict(V=V, probe=p, xl=xl, yl=yl, M=M, prop=prop, bprop=bprop, bin=binning, sig=sig, dx=dx, n=n,
                lam=lam, nz=nz, npos=npos)
