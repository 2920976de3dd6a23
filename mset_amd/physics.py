"""Host-side inputs of the multislice path: constants, propagator, probe, probe normalisation.

These are computed once per reconstruction on the host, as the reference's MATLAB initialisation does
(STEP00_INIT.m:99-100, 195, 211, 214-218); they are inputs to the device path, not part of it.
Layouts follow the MEX: planes are numpy [ny, nx] (MATLAB (x, y) reversed), DC-centred in Fourier space.
"""
import numpy as np

_M = 9.109383e-31      # electron mass (kg), electron_wavelength.m:3
_E = 1.602177e-19      # elementary charge (C)
_C = 299792458.0       # speed of light (m/s)
_H = 6.62607e-34       # Planck constant (J s)


def electron_wavelength(E0_eV):
    """Relativistic electron wavelength in Angstrom (electron_wavelength.m:1-9)."""
    return _H / np.sqrt(2 * _M * _E * E0_eV) / np.sqrt(1 + _E * E0_eV / 2 / _M / _C ** 2) * 1e10


def interaction_constant(E0_eV):
    """Interaction constant sigma in rad/(V A) (interaction_constant.m:1-10)."""
    lam = electron_wavelength(E0_eV)
    return (2 * np.pi / lam / E0_eV) * (_M * _C ** 2 + _E * E0_eV) / (2 * _M * _C ** 2 + _E * E0_eV)


def _mround(v):
    return int(np.floor(v + 0.5))


def _freq_grid(n, dx):
    """(qx, qy) on the centred grid of Func_generate_propagator.m:17-26: (k - round((n+1)/2)) / (n dx)."""
    k = np.arange(1, n + 1) - _mround((n + 1) / 2)
    q = k / (n * dx)
    qy, qx = np.meshgrid(q, q, indexing='ij')     # [y, x]
    return qx, qy


def propagator(n, dx, lam, slice_binning, back=False):
    """Fresnel propagator exp(-+ i pi lambda dz q^2), dz = dx * slice_binning, rounded to single
    (Func_generate_propagator.m:36-37); MATLAB (centred) layout, complex64 [n, n]."""
    qx, qy = _freq_grid(n, dx)
    sgn = 1.0 if back else -1.0
    return np.exp(1j * sgn * np.pi * lam * (dx * slice_binning) * (qx ** 2 + qy ** 2)).astype(np.complex64)


def probe(n, dx, lam, alpha_mrad, defocus=0.0, C30=0.0, C50=0.0, C12=0.0, Phi12=0.0, C21=0.0, Phi21=0.0,
          C23=0.0, Phi23=0.0):
    """Aperture-limited probe with aberrations (Func_generate_probe_wave.m:14-56; C10 = -probeDefocus,
    Func_check_probe_wave_parameters.m:136-142).  complex128 [n, n]."""
    qx, qy = _freq_grid(n, dx)
    q2 = qx ** 2 + qy ** 2
    qa = lam * np.sqrt(q2)
    phi = np.arctan2(qy, qx)
    mask = q2 < (alpha_mrad / 1000.0 / lam) ** 2
    C10 = -defocus
    chi = (0.5 * qa ** 2 * C10 + 0.25 * qa ** 4 * C30 + 1.0 / 6 * qa ** 6 * C50
           + 0.5 * qa ** 2 * C12 * np.cos(2 * (phi - Phi12)) + 1.0 / 3 * qa ** 3 * C21 * np.cos(phi - Phi21)
           + 1.0 / 3 * qa ** 3 * C23 * np.cos(3 * (phi - Phi23)))
    chi = 2 * np.pi / lam * chi
    a = mask * np.exp(-1j * chi)
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(a)))


def normalise_probe(p, mean_intensity):
    """probe * sqrt(mean_intensity / sum(|fft2(probe)|^2)) (STEP00_INIT.m:214-218)."""
    return p * np.sqrt(mean_intensity / np.sum(np.abs(np.fft.fft2(p)) ** 2))


def mean_pattern_intensity(measured, n_scan_x, n_scan_y):
    """Sum over the scan of each pattern's total counts / (N_scan_x N_scan_y) (STEP00_INIT.m:63-79)."""
    return float(np.sum(measured, dtype=np.float64) / n_scan_x / n_scan_y)


def scan_grid(n_scan_x, n_scan_y, step_px, origin_px):
    """Raster positions (x fastest, MATLAB linear cell order), 1-based pixels (STEP00_INIT.m:168-179)."""
    iy, ix = np.meshgrid(np.arange(n_scan_y), np.arange(n_scan_x), indexing='ij')
    return ((ix.ravel() * step_px + origin_px).astype(np.float32), (iy.ravel() * step_px + origin_px).astype(np.float32))
