"""Extract golden vectors from the reference's bundled CuAu example (data files only).

Run in a container that has the reference mounted and h5py available:

    /opt/conda/bin/python3.9 tests/golden/extract_cuau.py /root/reference/examples/CuAu_24size_dataset

Nothing from the reference is executed or imported: the MAT v7.3 files are HDF5 containers read with
h5py's dataset reader, and the two MAT v5 files with scipy.io.loadmat (a plain data parser).

Output (tests/golden/cuau/):
  snapshot.npz     the saved STEM_data state of the last position of the last update pass
                   (examples/CuAu_24size_dataset/MSET_recon.mat:/STEM_data/*)
  recon.npz        rec_save (5 iterations), mean_error, per-tilt error, ground-truth potential
  measured.npz     the 8 tilts' measured patterns, (tilt, position k, qy, qx) float32
  manifest.json    layout notes and scalar parameters

Layout convention of every array written here: C order with the MATLAB dimensions reversed, i.e. a
MATLAB array A(x, y, z) becomes a[z, y, x]; position k is MATLAB's column-major linear cell index.
"""
import json
import os
import sys

import h5py
import numpy as np
import scipy.io as sio


def cplx(ds):
    a = np.array(ds)
    if a.dtype.names:
        return (a['real'] + 1j * a['imag']).astype(np.complex64 if a['real'].dtype == np.float32 else np.complex128)
    return a


def scalar(ds):
    return float(np.array(ds).ravel()[0])


def cell_patterns(f, ds):
    refs = np.array(ds)              # h5py shape (N_scan_y, N_scan_x) = MATLAB (N_scan_x, N_scan_y) reversed
    flat = refs.ravel(order='C')     # C order over (y1, x1) == MATLAB column-major linear index k
    return np.stack([np.array(f[r], dtype=np.float32) for r in flat])


def main(src):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'cuau')
    os.makedirs(out, exist_ok=True)
    f = h5py.File(os.path.join(src, 'MSET_recon.mat'), 'r')
    S = f['STEM_data']

    snap = dict(
        prop=cplx(S['prop']), back_prop=cplx(S['back_prop']), prop2D=cplx(S['prop2D']),
        trans=cplx(S['trans']), potz=np.array(S['potz']), RVol=np.array(S['RVol']), rec=np.array(S['rec']),
        tmp_probe_wfn=cplx(S['tmp_probe_wfn']), probe_wfn=cplx(S['probe_wfn']),
        init_wave2D=cplx(S['init_wave2D']), save_wave_f=cplx(S['save_wave_f']), resi_vec=cplx(S['resi_vec']),
        grad=cplx(S['grad']),
        scan_pos=np.array(S['scan_pos']),            # (tilt, 2, k)
        tilt_angles=np.array(S['tilt_angles']).T,     # (tilt, 3)
        error=np.array(S['error']).ravel(), errorlist=np.array(S['errorlist']).ravel(),
    )
    np.savez_compressed(os.path.join(out, 'snapshot.npz'), **snap)

    info = h5py.File(os.path.join(src, 'CuAu_s24_p72_Np_8_STEM_data_info.mat'), 'r')['STEM_data']
    recon = dict(rec_save=np.array(f['rec_save']), mean_error=np.array(f['mean_error']).ravel(),
                 pot3D=np.array(info['pot3D']))
    np.savez_compressed(os.path.join(out, 'recon.npz'), **recon)

    meas = []
    for t in range(1, 9):
        g = h5py.File(os.path.join(src, 'CuAu_s24_p72_Np_8_4DSTEM_data_%d.mat' % t), 'r')
        meas.append(cell_patterns(g, g['Full_4D_STEM_data']))
    meas = np.stack(meas)
    snapshot_meas = cell_patterns(f, S['measured_4D_data'])
    assert np.array_equal(snapshot_meas, meas[int(scalar(S['Nth_angle'])) - 1]), 'snapshot data != its tilt file'
    np.savez_compressed(os.path.join(out, 'measured.npz'), measured=meas)

    sp = sio.loadmat(os.path.join(src, 'scan_pos.mat'))['scan_pos'].astype(np.float32)   # MATLAB (576, 2)
    ta = sio.loadmat(os.path.join(src, 'tilt_angles.mat'))['Tilt_Angles']                # MATLAB (8, 3)

    manifest = dict(
        source='jhlee0667/MSET examples/CuAu_24size_dataset (MSET_recon.mat, *_4DSTEM_data_{1..8}.mat, '
               'CuAu_s24_p72_Np_8_STEM_data_info.mat, scan_pos.mat, tilt_angles.mat)',
        layout='C order, MATLAB dims reversed: A(x,y,z) -> a[z,y,x]; patterns (tilt, k, qy, qx); '
               'k = MATLAB column-major cell index (x fastest)',
        E0_kV=scalar(S['E0']), lambda_A=scalar(S['lambda']), sigma=scalar(S['sigma']),
        potential_pixelsize=scalar(S['potential_pixelsize']), slice_binning=int(scalar(S['slice_binning'])),
        step_size=np.array(S['step_size']).ravel().tolist(), alpha_mrad=scalar(S['alpha']),
        probeDefocus_A=float(np.array(S['probeDefocus']).ravel()[0]),
        row=scalar(S['row']), col=scalar(S['col']), k=int(scalar(S['k'])), Nth_angle=int(scalar(S['Nth_angle'])),
        bls_parameter=scalar(S['bls_parameter']), use_positivity=bool(scalar(S['use_positivity'])),
        N_scan_x=int(scalar(S['N_scan_x'])), N_scan_y=int(scalar(S['N_scan_y'])),
        numPlanes=int(scalar(S['numPlanes'])),
        scan_pos_file=sp.tolist(), tilt_angles_file=ta.tolist(),
    )
    with open(os.path.join(out, 'manifest.json'), 'w') as fh:
        json.dump(manifest, fh, indent=1)
    for name in sorted(os.listdir(out)):
        print(name, os.path.getsize(os.path.join(out, name)))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/examples/CuAu_24size_dataset')
