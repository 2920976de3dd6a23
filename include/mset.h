/* mset.h -- C ABI of the MI355X multislice forward/adjoint engine (libmset_amd.so).
 *
 * Drop-in for the CUDA MEX `mset_cuda_kernel` of jhlee0667/MSET:
 *   mexFunction dispatch          src/CUDA/mex_main.cu:19-39
 *   'upload'                      src/CUDA/mex_main.cu:42-95   -> mset_upload
 *   'run_mset'                    src/CUDA/mex_main.cu:98-155  -> mset_run
 *   'error_mset'                  src/CUDA/mex_main.cu:157-184 -> mset_error
 *   `clear mset_cuda_kernel`      src/Run_MSET_recon.m:174,235 -> mset_destroy
 * called from src/Run_MSET_recon.m:162-174 (update pass) and :226-235 (error pass).
 *
 * Array conventions are the MEX's: every array is MATLAB column-major, i.e. a C array with the
 * MATLAB dimensions reversed.  RVol(x, y, z) is float[nz][ny][nx]; complex arrays are interleaved
 * (re, im) float pairs (MATLAB R2018a interleaved complex); the measured 4D data is the cell array
 * measured_4D_data{ix, iy} laid out as float[num_scan_pos][ny][nx] in MATLAB's linear cell order
 * (ix fastest), each pattern with its DC at the centre as stored by MATLAB.  Scan positions are
 * MATLAB 1-based pixel coordinates (scan_xlist along dim 1, scan_ylist along dim 2).
 *
 * All entry points return an MSET_* status; on failure mset_last_error() describes it.  No call
 * throws across the ABI.  A context may be used from one host thread at a time; independent
 * contexts (one per reconstruction job) may live side by side on one device.
 */
#ifndef MSET_H
#define MSET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSET_OK 0
#define MSET_EINVAL 1          /* bad argument, shape or call order */
#define MSET_ENOMEM 2          /* device allocation failed */
#define MSET_EDEVICE 3         /* HIP runtime / kernel failure */
#define MSET_EUNSUPPORTED 4    /* valid for the reference but not provided here (e.g. a plane size not built) */

#define MSET_METHOD_MSET 1     /* 'run_mset' / 'error_mset': multislice, STEP02_FORWARD_4D.m / STEP03_BACK.m */
#define MSET_METHOD_SSET 2     /* 'run_sset' / 'error_sset': single slice, STEP02_FORWARD_4D_SSET.m /
                                  STEP03_BACK_SSET.m (src/CUDA/mset.cu:345-546, 700-810): the whole depth
                                  projected into one transmission, no propagation, object term paired with
                                  the incident wave, the same probe and position updates */

#define MSET_GRAD_REF 0        /* STEP03_BACK.m:14-15 pairing: conj of the wave after plane a */
#define MSET_GRAD_EXACT 1      /* exact adjoint: wave entering plane a; spectral position derivative */

typedef struct mset_ctx mset_ctx;

/* The STEM_data fields the MEX 'upload' reads (src/CUDA/mex_main.cu:53-75, mex_tools.hpp:15-165). */
typedef struct mset_params {
    int32_t nx, ny, nz;           /* size(RVol); nx == ny == size(init_wave2D) (mex_tools.hpp:120-122) */
    int32_t n_scan_x, n_scan_y;   /* N_scan_x, N_scan_y: the error normalisation */
    int32_t num_scan_pos;         /* positions processed, in cell order */
    int32_t slice_binning;        /* slices per propagation plane */
    float alpha, sigma, E0;       /* sigma in rad/(V A); alpha, E0 are carried like the MEX, unused */
    float potential_pixelsize;    /* dx, A */
    float step_size[3];           /* object, probe, position step sizes */
    int32_t grad_mode;            /* MSET_GRAD_REF (reference semantics) or MSET_GRAD_EXACT */
    int32_t diffraction_transpose; /* STEM_data.diffraction_transpose (STEP00_INIT.m:58-73): nonzero transposes
                                      every measured pattern at ingest, as MATLAB's transpose() of each cell */
} mset_params;

/* ---- lifetime ------------------------------------------------------------------------------- */
int mset_create(mset_ctx** ctx, int device);
int mset_destroy(mset_ctx* ctx);
const char* mset_last_error(void);
/* Stream the context's work is issued on (a hipStream_t); 0 before mset_create succeeds. */
void* mset_stream(mset_ctx* ctx);

/* ---- MEX-compatible entry points -------------------------------------------------------------- */
/* 'upload' (mex_main.cu:42-95).  Copies every input (the MEX borrows the measured cell array; here
 * the patterns are ingested once as sqrt(ifftshift(M)) and stay on the device).  prop2D is the
 * MATLAB-layout (centred) propagator of Func_generate_propagator.m; its ifftshift and conjugate are
 * taken here (mex_main.cu:80-83). */
int mset_upload(mset_ctx* ctx, const mset_params* p, const float* RVol, const float* init_wave2D,
                const float* prop2D, const float* scan_xlist, const float* scan_ylist, const float* measured);

/* 'run_mset' / 'run_sset' (mex_main.cu:98-155): one sequential pass over the positions with the per-position
 * updates of Run_MSET_recon.m:134-158; returns the updated RVol, probe, scan lists and, if error_out
 * is not NULL, the error (sum over positions of mean((sqrt M - sqrt I)^2) / (N_scan_x N_scan_y)). */
int mset_run(mset_ctx* ctx, int method, float* RVol_out, float* probe_out, float* scan_xlist_out,
             float* scan_ylist_out, float* error_out);

/* 'error_mset' / 'error_sset' (mex_main.cu:157-184): forward + loss over every position, no updates. */
int mset_error(mset_ctx* ctx, int method, float* error_out);

/* ---- device-resident entry points (no host copies of the state) ------------------------------- */
/* Same as mset_run / mset_error on the resident state; error accumulated in double. */
int mset_run_resident(mset_ctx* ctx, int method, double* error_out);
int mset_error_resident(mset_ctx* ctx, int method, double* error_out);
/* Copy the resident state to the host (any pointer may be NULL). */
int mset_download(mset_ctx* ctx, float* RVol, float* probe, float* scan_xlist, float* scan_ylist);
/* Replace the resident volume / probe / positions (any pointer may be NULL). */
int mset_set_state(mset_ctx* ctx, const float* RVol, const float* probe, const float* scan_xlist,
                   const float* scan_ylist);

/* ---- batched path ----------------------------------------------------------------------------- */
/* Forward + gradient at the fixed resident (V, probe) for positions [first, first + count): the sums
 * over those positions of the per-position terms STEP03_BACK.m forms.  Outputs (host, any may be NULL):
 *   dV     float[L][ny][nx], L = ceil(nz / slice_binning): sum_k Re(grad_k) of each plane (every voxel
 *          of plane a's bin receives this value in the update RVol += step(1) * real(grad));
 *   dprobe interleaved complex [ny][nx]: sum_k of the probe step direction (the probe update is
 *          probe - step(2) * dprobe);
 *   dpos   float[count][2]: per-position (row, col) update terms (position += step(3) * dpos);
 *   error  sum of the positions' error terms, normalised as in mset_run. */
int mset_grad(mset_ctx* ctx, int first, int count, float* dV, float* dprobe, float* dpos, double* error_out);
/* (mset_grad, mset_waves, mset_simulate and the mset_recon_* tilt series use the multislice operator.) */

/* Allocation-free mset_grad.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It runs the same kernels as mset_grad on the resident state of an uploaded context, but its scratch
 * (transmission planes, fftn(probe), the stored waves and gradient slabs of the resident positions) is a
 * static device array of MSET_NOALLOC_SCRATCH_BYTES per GPU, held from library load to process end; it
 * calls no hipMalloc, hipFree or allocating routine.  As many positions are resident at once as that
 * array holds (at most the mset_set_slot_limit cap); a problem whose one resident position does not fit
 * is rejected with MSET_EINVAL.  Outputs are DEVICE pointers with mset_grad's layouts: dV float[L][ny][nx],
 * dprobe interleaved complex [ny][nx], dpos float[count][2], err_pos double[count] (per-position sums
 * over pixels of (sqrt M - sqrt I)^2, unnormalised).  The work is only issued on the context's stream
 * (mset_stream): synchronise it before reading the outputs.  All calls on one GPU share its array and
 * nothing in it needs clearing (every region is written before it is read); calls that overlap in time
 * on one GPU, from any contexts or threads, race on it and must be serialised by the caller. */
#define MSET_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int mset_grad_noalloc(mset_ctx* ctx, int first, int count, float* dV, float* dprobe, float* dpos,
                      double* err_pos);

/* The incident wave and the wave after each plane for position k at the resident state, i.e.
 * tmp_probe_wave and save_wave_f of STEP02_FORWARD_4D.m:27-35: interleaved complex [L + 1][ny][nx]. */
int mset_waves(mset_ctx* ctx, int k, float* waves_out);

/* Exit waves of the batched path.  STORE keeps every plane's wave for the adjoint (L + 2 planes per
 * resident position, the reference's save_wave3D); RECOMPUTE keeps 5 and rebuilds psi_a in the adjoint by
 * inverting the forward sweep, psi_a = conj(t_a) .* ifftn(fftn(psi_{a+1}) .* conj(prop)), for 50 % more
 * transforms.  It needs a unit-modulus propagator (the Fresnel propagator of Func_generate_propagator.m
 * is one; EINVAL otherwise).  CHECKPOINT keeps every K-th wave (K = ceil(sqrt(L + 1))) and the last
 * segment of K, L / K + K + 3 planes, and rebuilds each earlier segment from its checkpoint with the
 * forward sweep's own steps: any propagator, the stored variant's waves, about one forward sweep more.
 * AUTO (the default) stores unless the store scratch (waves and gradient slabs) would exceed a quarter of
 * the device memory; then it checkpoints if that scratch fits, else recomputes (unit-modulus propagator),
 * else checkpoints.  RECOMPUTE with fewer than 4 planes, and every variant under the single-slice operator, store.
 * No counterpart in the reference, which stores.  mset_exit_waves reports the variant and the per-position
 * scratch bytes of the last batched call. */
#define MSET_EXIT_WAVES_AUTO 0
#define MSET_EXIT_WAVES_STORE 1
#define MSET_EXIT_WAVES_RECOMPUTE 2
#define MSET_EXIT_WAVES_CHECKPOINT 3
int mset_set_exit_waves(mset_ctx* ctx, int mode);
int mset_exit_waves(mset_ctx* ctx, int* resolved, size_t* bytes_per_slot);
/* Cap on the persistent grid of the batched kernels (resident positions, each with its own scratch);
 * 0 (the default) = every workgroup the device keeps resident (mset_batch_slots).  A tuning knob with
 * no counterpart in the reference: fewer slots hold less scratch and may keep more of it in cache. */
int mset_set_slot_limit(mset_ctx* ctx, int slots);
/* Engine flags of the position kernels, a tuning knob with no counterpart in the reference; every choice
 * agrees with the FP64 oracle to the same tolerances.  AUTO (0, the default): at 256 x 256, when the
 * propagator factors as p(kx, ky) = px(kx) py(ky) (the Fresnel propagator of Func_generate_propagator.m
 * does, checked at upload), the forward kernels and the reference-pairing batched gradient with stored
 * waves run the separable routine, which applies the two 1D half-steps of neighbouring planes in one pass
 * over the plane (half the passes).  GENERAL: always the 2D routine.  ALT_SHAPE: the batched kernel with
 * stored waves runs the alternative workgroup shape of its plane size: twice the threads per position of
 * the LDS engine up to 64, 1024 threads at 96, at 128 the register engine with a half-size transpose
 * buffer and a 128-register budget (two workgroups per CU instead of one), at 256 16-line tiles with 256
 * threads (three workgroups per CU instead of two).  The flags combine. */
#define MSET_ENGINE_AUTO 0
#define MSET_ENGINE_GENERAL 1
#define MSET_ENGINE_ALT_SHAPE 2
int mset_set_engine(mset_ctx* ctx, int engine);

/* Mini-batch variant of 'run_mset': positions are taken in order in batches of `batch`; each batch's
 * gradient is formed at the state left by the previous batch and then applied.  batch == 1 equals
 * the sequential semantics up to summation order; step sizes 0 make it a forward + gradient pass. */
int mset_run_batched(mset_ctx* ctx, int method, int batch, double* error_out);

/* ---- 3D Fourier-shear rotation (Func_rotate_3Dvol_ROTmat_FourierShear.m) ---------------------- */
/* Rotates a device or host volume float[nz][ny][nx] by the 3x3 row-major matrix R (the MATLAB MAT
 * argument), with rotStepInit the largest angle per step (11 in the reference).  in/out are host
 * arrays; in may equal out. */
int mset_rotate(mset_ctx* ctx, const float* in, float* out, int nx, int ny, int nz, const double* R,
                double rot_step_init);
/* Func_rot_3Dvol_FourierShear (inverse = 0) / Func_inv_rot_3Dvol_FourierShear (inverse = 1) on the
 * resident volume: RVol <- rot(RVol) with vec1..vec3 the rotation axes and (phi, theta, psi) in degrees;
 * identity when all |angles| <= 0.5. */
int mset_rotate_resident(mset_ctx* ctx, const double* vec1, const double* vec2, const double* vec3,
                         double phi, double theta, double psi, int inverse);

/* ---- total-variation regulariser (TV_method.m:5-84) -------------------------------------------- */
/* out = TV_method(in, lambda, n_iter) for a host volume float[nz][ny][nx] (in may equal out): Chambolle's
 * projection for min_u 1/2 |u - in|^2 + lambda TV(u), periodic forward differences, tau = 1/12.
 * tv_error (nullable) gets lambda * sum |grad u|, the error TV_method returns. */
int mset_tv(mset_ctx* ctx, const float* in, float* out, int nx, int ny, int nz, float lambda, int n_iter,
            double* tv_error);

/* ---- forward model as a data generator -------------------------------------------------------- */
/* |fftn(exit wave)|^2 of every position at the resident state, DC-centred (the layout of the measured
 * data): float[num_scan_pos][ny][nx].  The measured patterns uploaded with the state are not used. */
int mset_simulate(mset_ctx* ctx, float* intensity_out);

/* ---- device-resident tilt series (the update / error passes of Run_MSET_recon.m:111-262) ------ */
/* A reconstruction keeps rec (the lab-frame volume) and every tilt's probe, positions and patterns on
 * the device.  mset_recon_iterate runs one update pass over all tilts: for tilt j, RVol = inverse
 * Fourier-shear rotation of rec (STEP01_ROTATION.m), the positions (STEP02-04), rec = forward rotation
 * of RVol and optional positivity (Run_MSET_recon.m:178-185).  The error pass, BLS rollback (snapshot /
 * restore) and step changes are separate calls so a host driver can restate Run_MSET_recon.m:189-262. */
#define MSET_PATH_PLAIN 0      /* one position per launch, sequential updates (reference semantics) */
#define MSET_PATH_BATCHED 1    /* mini-batches of positions per launch, update per batch */

typedef struct mset_recon mset_recon;

/* common: nx = ny, nz, slice_binning, sigma, potential_pixelsize, step_size, grad_mode (the scan fields
 * are per tilt); vecs: vec1, vec2, vec3 (9 doubles, NULL = [0 0 1], [0 1 0], [1 0 0]); rec float[nz][ny][nx];
 * prop2D: the MATLAB-layout propagator. */
int mset_recon_create(mset_recon** r, int device, const mset_params* common, int n_tilts, const double* vecs,
                      const float* rec, const float* prop2D);
int mset_recon_destroy(mset_recon* r);
/* tilt j: angles (phi, theta, psi) in degrees, its probe, scan grid and measured patterns */
int mset_recon_set_tilt(mset_recon* r, int j, const double* angles, const float* probe, int n_scan_x, int n_scan_y,
                        int num_scan_pos, const float* xlist, const float* ylist, const float* measured);
int mset_recon_set_steps(mset_recon* r, const float* steps);
/* One update pass.  batch: positions per mini-batch (MSET_PATH_BATCHED).  use_graph: capture the pass in
 * a hipGraph on first use and replay it afterwards.  error_out (nullable): mean over tilts of the
 * update-pass errors. */
int mset_recon_iterate(mset_recon* r, int path, int batch, int use_graph, int positivity, double* error_out);
/* Error pass (Run_MSET_recon.m:189-239): errs[j] for every tilt at the current rec. */
int mset_recon_error_pass(mset_recon* r, double* errs);
/* rec float[nz][ny][nx]; probes interleaved complex [n_tilts][ny][nx]; scan lists concatenated by tilt */
int mset_recon_get(mset_recon* r, float* rec, float* probes, float* xlists, float* ylists);
/* TV step of the driver (Run_MSET_recon.m:266-269) on the resident rec: rec = sf_max * TV_method(rec /
 * sf_max, lambda, n_iter) with sf_max = max(rec); tv_error (nullable) as for mset_tv.  EINVAL when the
 * maximum is zero (MATLAB would fill the volume with NaN). */
int mset_recon_tv(mset_recon* r, float lambda, int n_iter, double* tv_error);
/* exit-wave variant of the tilt series' batched passes (MSET_EXIT_WAVES_*, as mset_set_exit_waves) */
int mset_recon_set_exit_waves(mset_recon* r, int mode);
/* slot cap of the tilt series' batched passes (as mset_set_slot_limit) */
int mset_recon_set_slot_limit(mset_recon* r, int slots);
/* mset_set_engine for the series' passes (drops a captured graph when it changes). */
int mset_recon_set_engine(mset_recon* r, int engine);
/* BLS rollback (Run_MSET_recon.m:113-115, 245-247) */
int mset_recon_snapshot(mset_recon* r);
int mset_recon_restore(mset_recon* r);
int mset_recon_profile_enable(mset_recon* r, int enable);
int mset_recon_profile_read(mset_recon* r, const char* name, int* launches, double* total_ms);
void* mset_recon_stream(mset_recon* r);

/* ---- page-locked host buffers ------------------------------------------------------------------- */
/* Host memory for the arrays a caller uploads (patterns, volumes): copies from it run at the full link
 * rate instead of through the driver's staging of pageable memory.  No counterpart in the reference,
 * whose MATLAB arrays are pageable; any host pointer stays valid for every entry point. */
int mset_host_alloc(size_t bytes, void** ptr);
int mset_host_free(void* ptr);

/* ---- device capacity (admission control for several reconstructions on one device) ------------- */
/* Free and total device memory in bytes. */
int mset_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* Workgroups the batched position kernel keeps resident for plane size n: the number of per-slot
 * workspaces (saved waves, gradient slabs) a context allocates. */
int mset_batch_slots(int device, int n, int* slots);

/* ---- instrumentation -------------------------------------------------------------------------- */
/* When enabled, every launch of the named kernels is bracketed by HIP events on the context stream. */
int mset_profile_enable(mset_ctx* ctx, int enable);
/* name: "position_plain", "position_batch", "position_fwd", "rotate"; returns launches and total ms. */
int mset_profile_read(mset_ctx* ctx, const char* name, int* launches, double* total_ms);
int mset_synchronize(mset_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MSET_H */
