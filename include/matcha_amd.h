/* matcha_amd: C ABI of the MI355X (gfx950) orientation-search library.
 *
 * The reference (swing-research/Matcha) has no FFI: its hot path is a set of
 * duck-typed Python objects (Matcha, CrossCorrelationMatcher, ShiftMatcher)
 * over torch, numba.cuda kernels, cufinufft and torch-harmonics.  Each entry
 * point below replaces one reference function or kernel group, cited as
 * path:line in the reference repository.  The Python classes in
 * matcha_amd/core bind these through ctypes.
 *
 * Conventions
 *  - every array argument is device memory owned by the caller, except the
 *    plan, which owns its tables and workspaces (allocated at create time:
 *    hot calls allocate nothing);
 *  - complex arrays are interleaved float2 (torch complex64);
 *  - launches go on the caller's stream; no call synchronises the host;
 *  - return 0 on success, an MCH_ERR_* or HIP code otherwise; the message is
 *    in mch_last_error() (per host thread);
 *  - a plan is not re-entrant: calls that share a plan must be serialised.
 */
#ifndef MATCHA_AMD_H
#define MATCHA_AMD_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCH_OK 0
#define MCH_ERR_ARG 1001
#define MCH_ERR_SIZE 1002
#define MCH_ERR_HIP 1003
#define MCH_ERR_UNSUPPORTED 1004

/* Last error message of this host thread. */
const char* mch_last_error(void);

/* ---------------------------------------------------------------- SO(3) --- */

/* Plan for the SO(3) search: SOFFT tables (d^l(beta_j), l < L0, FP64->FP32),
 * DFT twiddles, and workspaces for max_batch particles x max_candidates.
 * Replaces SOFFT._precompute (src/matcha/core/SOFFT.py:30-64) and the
 * WignerDMatrices/WignerDGradients constructors (WignerDMatrices.py:1076-1176,
 * WignerDGradients.py:81-152). */
int mch_so3_plan_create(int L_max, int L0, int K, int max_batch, int max_candidates,
                        void** out_plan);
int mch_so3_plan_destroy(void* plan);

/* Full score grid f[b, beta, alpha, gamma] (float32, (B, n, n, n), n = 2*L0*K).
 * Replaces SOFFT.eval (src/matcha/core/SOFFT.py:67-98). sigma: (B, >= compute_size(L0))
 * complex64 rows of stride ld_sigma (elements). */
int mch_sofft_eval(void* plan, const void* sigma, long long ld_sigma, int B, float* grid,
                   hipStream_t stream);

/* Top-C grid points per particle as Euler angles and scores, each (B, C).
 * Replaces SOFFT.eval + sparse_topk_per_batch_jit + SOFFT.ids_to_angles
 * (src/matcha/core/Matcha.py:27-58,207-236; SOFFT.py:100-124). Ties: lower
 * flat grid index first. */
int mch_sofft_topk(void* plan, const void* sigma, long long ld_sigma, int B, int C, float* alpha,
                   float* beta, float* gamma, float* score, hipStream_t stream);

/* Score, gradient and Hessian of Re sum_{l<=L} D(alpha,beta,gamma) sigma at
 * (B, C) given angles; out is (B*C, 10): s, g_a, g_b, g_g, H_aa, H_ab, H_ag,
 * H_bb, H_bg, H_gg.  Replaces WignerDMatrices.D + the score einsum +
 * WignerDGradients.set_so3_coeffs/get_derivatives
 * (WignerDMatrices.py:1291-1375; Matcha.py:184,288; WignerDGradients.py:155-302). */
int mch_so3_evaluate(void* plan, const void* sigma, long long ld_sigma, int B, int C, int L,
                     const float* alpha, const float* beta, const float* gamma, float* out,
                     hipStream_t stream);

/* Newton (newton=1) or gradient (newton=0) march of (B, C) candidates over the
 * level schedule; angles in/out, score out.  Replaces Matcha.optimize and
 * Matcha._take_step with the singular-Hessian skip and the lr rule
 * (src/matcha/core/Matcha.py:11-15,150-205,239-367). */
int mch_newton_march(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                     const int* levels, int n_levels, int num_steps, int newton, float base_lr,
                     float* alpha, float* beta, float* gamma, float* score, hipStream_t stream);

/* Batch-global early exit of the march (the reference's stop_early config
 * field): when set, a level's remaining Newton steps are skipped once one of
 * its steps accepted no candidate of the batch.  mch_newton_march and
 * mch_search_orientations then run one launch per (level, step), reading the
 * accept counts of the previous launches on the device (no host sync);
 * num_steps is then limited to 63.  Replaces the stop_early branch of
 * Matcha._take_step and the break in Matcha.optimize
 * (src/matcha/core/Matcha.py:198-199,358-361). */
int mch_so3_set_stop_early(void* plan, int stop_early);

/* Best candidate per particle (first maximum).  Replaces reduce_over_candidates
 * (src/matcha/core/Matcha.py:19-24). */
int mch_reduce_candidates(const float* alpha, const float* beta, const float* gamma,
                          const float* score, int B, int C, float* out_alpha, float* out_beta,
                          float* out_gamma, float* out_score, hipStream_t stream);

/* Matcha.search_orientations: grid top-C, march, reduction; outputs (B,).
 * Replaces src/matcha/core/Matcha.py:369-403. */
int mch_search_orientations(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                            const int* levels, int n_levels, int num_steps, int newton,
                            float base_lr, float* out_alpha, float* out_beta, float* out_gamma,
                            float* out_score, hipStream_t stream);

/* mch_search_orientations with a march end per particle: march_ends is a
 * device array of B ints; particle b marches the levels of `levels` (strictly
 * increasing) below march_ends[b], then march_ends[b] (<= the plan's L_max) --
 * the schedule of a Matcha built with L_max = march_ends[b]
 * (src/matcha/core/Matcha.py:76,381-386), so one launch of the grid, the
 * march and the reduction serves a batch of mixed bandlimit schedules that
 * the reference runs as one Matcha per schedule
 * (src/matcha/align_subtomograms.py:61-134).  A march end outside
 * [1, L_max] gives that particle NaN outputs.  Not combined with stop_early. */
int mch_search_orientations_ends(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                                 const int* levels, int n_levels, const int* march_ends,
                                 int num_steps, int newton, float base_lr, float* out_alpha,
                                 float* out_beta, float* out_gamma, float* out_score,
                                 hipStream_t stream);

/* ------------------------------------------------------- projection --- */

/* Plan of the symmetry-adapted ball-harmonic projection: orbit table of the
 * fundamental voxels (V0 x 16 flat indices, image p + 4q + 8t of the voxel,
 * -1 where that image is outside the array), FP32 basis (Kpad x Ntot: the
 * cos-weighted row of voxel v at k = v, the sin-weighted one at
 * k = Kpad/2 + v, Kpad a multiple of 128 with Kpad/2 >= V0; the classes
 * m = 3 mod 4 hold their sin rows negated: they share the operand of the
 * classes m = 1 mod 4; NULL: built later by mch_fle_build_basis), per-column
 * output maps.  Replaces FLEBasis3D.precomp (src/matcha/core/FLEBasis3D.py:
 * 111-540: NUFFT plans, SHT, interpolation matrices). */
int mch_fle_plan_create(int N, int V0, int Kpad, int Ntot, const int* class_colpad,
                        const int* orbit, const float* basis, const int* chan_plus,
                        const int* chan_minus, const int* chan_m, int ne, int max_batch,
                        void** out_plan);
int mch_fle_plan_destroy(void* plan);

/* The plan's projection table built on the device in FP64 (rounded to FP32)
 * instead of uploaded: pass basis = NULL to mch_fle_plan_create, then this.
 * vox (V0 x 4 doubles, host): (r, cos theta, phi, h w) per fundamental voxel;
 * chan_lm (Ntot x 2 ints, host): (l, m) per column, l < 0 for padding;
 * chan_val (Ntot x 3 doubles, host): (lambda_lk, c_lk, sign of the sin half).
 * Entry (v, col) = h w j_l(lambda r / radius) P_l^m(cos theta) c cos(m phi),
 * and the sign times the same with sin(m phi) at row Kpad/2 + v.  Replaces the
 * basis evaluation of FLEBasis3D.precomp (src/matcha/core/FLEBasis3D.py:
 * 111-540, there a radial table sampled and interpolated).  Synchronous. */
int mch_fle_build_basis(void* plan, const double* vox, const int* chan_lm, const double* chan_val,
                        double radius);
/* Copy of the plan's table W (Kpad x Ntot floats, device) into out (device);
 * no reference counterpart (tests). */
int mch_fle_copy_basis(void* plan, float* out, hipStream_t stream);

/* a = B^H f: volumes (B, N^3) float32 (row stride ld_vol) -> coefficients
 * (B, ne) complex64 (row stride lda), reference order.  Replaces
 * FLEBasis3D.evaluate_t (src/matcha/core/FLEBasis3D.py:945-976,1026-1267). */
int mch_fle_project(void* plan, const float* vols, long long ld_vol, int B, void* coeffs,
                    long long lda, hipStream_t stream);
/* The same on the patches vol[a + flip(shift_zyx[b])] (shift_zyx (B, 3)
 * float32 on the device, NULL for none), read straight from the volumes:
 * evaluate_t of extract_shifted_patch_batch's output
 * (src/matcha/utils/io_utils.py:113-178) without the patch copy. */
int mch_fle_project_shifted(void* plan, const float* vols, long long ld_vol, int B, const float* shift_zyx,
                            void* coeffs, long long lda, hipStream_t stream);
/* The coefficients of the four volumes of mch_shift_tangent (the patch at a
 * continuous shift t (Bp, 3) and its three exact t-derivatives; volumes of
 * Nv^3, N <= Nv <= 1024) into rows b,
 * Bp + b, 2 Bp + b, 3 Bp + b of coeffs (4 Bp <= the plan's batch): what
 * mch_fle_project of mch_shift_tangent's output gives, with the volumes formed
 * per orbit image inside the operand build instead of written to and read
 * from HBM.  FP32 and BF16x3 precision (FP16x3: MCH_ERR_UNSUPPORTED).  No
 * reference counterpart (the reference differentiates the score in Euler
 * angles only, src/matcha/core/WignerDGradients.py:247-302); it serves the
 * differentiable score's shift derivative. */
int mch_fle_project_tangents(void* plan, const float* vols, long long ld_vol, int Nv, int Bp, const float* t,
                             void* coeffs, long long lda, hipStream_t stream);

/* The projection GEMM alone on the operand of the last mch_fle_project
 * (timing aid; no reference counterpart). */
int mch_fle_gemm_only(void* plan, int B, hipStream_t stream);

/* Compute precision of the projection GEMM: 0 = FP32 MFMA (the plain path,
 * the default), 1 = BF16 hi/lo split operands, 2 = FP16 hi/lo split with
 * power-of-2 row/column scales; the 16-bit modes take three MFMA products
 * (hi*hi + hi*lo + lo*hi) with FP32 accumulation.  The reference runs this
 * projection in FP32/complex64 (FLEBasis3D.py:945-976); the reference's own
 * precision_mode (:39-109) selects the expansion accuracy, not this. */
int mch_fle_set_precision(void* plan, int precision);

/* Tile configuration the 16-bit GEMM was autotuned to when the plan first
 * switched to a 16-bit mode: 10000 * (128-column tiles per workgroup - 1) +
 * 100 * (K splits) + 10 * (waves per workgroup, 2, 4 or 8) + pipeline depth
 * (2 or 3 LDS stages); 0 before that.  No reference counterpart (reporting
 * aid). */
int mch_fle_tuned_depth(void* plan);

/* Pin the 16-bit GEMM to one of the autotuner's tile configurations
 * (stages, waves per workgroup, 128-column tiles per workgroup, K splits);
 * errors for a configuration the plan cannot run.  No reference counterpart
 * (tuning and testing aid). */
int mch_fle_set_gemm_tiles(void* plan, int stages, int waves, int column_tiles, int k_splits);

/* Index table (lmax+1, 2*lmax+1, kdim) of coefficient positions for
 * (l, m, k) (-1 = absent).  Replaces CrossCorrelationMatcher._precompute_indices
 * (src/matcha/core/CrossCorrelationMatcher.py:133-181). */
int mch_corr_plan_create(int lmax, int kdim, const int* idx_host, int max_batch, void** out_plan);
int mch_corr_plan_destroy(void* plan);

/* sigma[b, Dindex(l,u,v)] = (-1)^{u+v} sum_k aT[l,k,-v] conj(aV[b,l,k,-u]) and
 * ||aV[b]|| (norms may be NULL).  Replaces CrossCorrelationMatcher.compute_sigma
 * and the norm of get_sigma (src/matcha/core/CrossCorrelationMatcher.py:183-249). */
int mch_compute_sigma(void* plan, const void* aT, const void* aV, long long ldv, int ne, int B,
                      void* sigma, long long lds, float* norms, hipStream_t stream);

/* mch_compute_sigma for the shells l <= L only (entries above L are not
 * written): the search reads sigma up to its L_max, below the expansion's
 * lmax.  Same reference counterpart. */
int mch_compute_sigma_degree(void* plan, const void* aT, const void* aV, long long ldv, int ne, int B, int L,
                             void* sigma, long long lds, float* norms, hipStream_t stream);

/* ------------------------------------------------------------ shift --- */

/* Plan of the FFT shift match: the centred Fourier reference (complex64
 * (N,N,N), host memory, as built by ShiftMatcher.set_reference), the cosine
 * background mask radius (box_size / (2 voxel_size)), FFT tables and
 * workspaces.  N must be a multiple of 4 of the form 2^a 3^b, <= 128.
 * Replaces ShiftMatcher.set_reference + setup_utils mask tables
 * (src/matcha/core/ShiftMatcher.py:805-822,442-472; src/matcha/utils/setup_utils.py:239-301). */
int mch_shift_plan_create(int N, const void* ref_host, float box_radius, int max_batch,
                          void** out_plan);
int mch_shift_plan_destroy(void* plan);

/* Sub-voxel refinement: upsample factor `up` (> 1) over a window of
 * 2 * radius + 1 voxels, evaluated as the reference's local matrix DFT around
 * zero (the coarse lag is discarded, as compute_shift does in mode "local");
 * up == 1 selects integer lags over the whole box.  Replaces the
 * upsample_factor > 1 branch of compute_shift / _upsampled_dft
 * (src/matcha/core/ShiftMatcher.py:46-92,241-282) called from
 * estimate_shift_like_relion (:658-670). */
int mch_shift_plan_set_upsample(void* plan, int up, int radius);

/* Integer-lag shift search (shift_upsample_factor 1): rotate the reference by
 * rot (B, 3x3 row-major acting on (x, y, z) = array axes (2, 1, 0)), extract
 * the patch at the prior shift_zyx (B, 3) (NULL = 0), cosine-masked
 * background, roll, rFFT, cross power with the masked half spectrum (x ctf
 * (B, N, N, N/2+1) if not NULL), inverse FFT with the half axis zero-padded,
 * first argmax.  out_shift (B, 3) in the reference's flipped order, out_score
 * (B,) normalised by the spectra's norms.  Replaces estimate_shift_like_relion
 * + compute_shift_local (src/matcha/core/ShiftMatcher.py:94-282,491-674). */
int mch_shift_search(void* plan, const float* subtomos, long long ld_vol, int B,
                     const float* shift_zyx, const float* rot, const float* ctf, float* out_shift,
                     float* out_score, hipStream_t stream);

/* Storage precision of the shift searches (no reference counterpart: the
 * reference keeps its rotated reference and the cross-correlation spectra in
 * complex64, src/matcha/core/ShiftMatcher.py:171-282,491-537).  reduced = 0
 * (default): complex FP32 throughout, the reference's values.  reduced = 1:
 * the rotated reference sampled from FP16 quads scaled by a power of two (two
 * 16-byte gathers per trilinear sample instead of four, weights rounded to
 * FP16, products accumulated in FP32) and, on the integer-lag path, the work
 * spectrum between passes stored as complex BF16 (half the bytes); all
 * arithmetic stays FP32.  A throughput mode: lags agree with the FP32 path
 * where the correlation has a clear peak, scores to about 2e-3, and lags of
 * near-tied correlation values may differ by one voxel. */
int mch_shift_set_precision(void* plan, int reduced);

/* Per-kernel timing of mch_shift_search (a measurement aid, no reference
 * counterpart): when enabled, every search records events around its four
 * launches on its stream; mch_shift_kernel_ms waits for the last search and
 * returns the four durations in ms (forward x, plane, inverse x / kx, final). */
int mch_shift_plan_set_timing(void* plan, int enable);
int mch_shift_kernel_ms(void* plan, float* ms4);

/* Shifted patches patch[a] = vol[a + flip(shift_zyx)] (trilinear for fractional
 * shifts, zeros outside, border slice zeroed), volumes (B, N^3) float32; shift_zyx
 * (B, 3) or NULL.  Replaces extract_shifted_patch_batch without normalisation
 * (src/matcha/core/ShiftMatcher.py:558-616). */
int mch_extract_patches(const float* vols, long long ld_vol, int B, int N, const float* shift_zyx,
                        float* out, long long ld_out, hipStream_t stream);

/* Shift-search rotations (B, 9) from search_orientations' Euler angles:
 * rb = P R^T P with R = Rz(alpha) Ry(beta) Rz(gamma) on array axes (0,1,2) and P
 * the axis reversal; the device-side equivalent of update_rotation_estimate's
 * x<->z swap followed by search_shifts' conj + Euler round trip
 * (src/matcha/utils/rotation_ops.py:23-34; src/matcha/core/ShiftMatcher.py:913). */
int mch_euler_to_shift_rot(const float* alpha, const float* beta, const float* gamma, int B,
                           float* rot, hipStream_t stream);

/* One alternation's particles: the N^3 patch of each (Nv^3) particle centred
 * on Nv/2 and displaced by shift (x, y, z) (trilinear, zeros outside),
 * normalised by its mean and population standard deviation over the binary
 * sphere |i - N/2| < radius, times that mask, then rotated: out[i] = that
 * volume at c + M (i - c), c = N/2, (x, y, z) index coordinates with x the
 * last axis, trilinear with zeros outside.  vols (B, Nv^3) float32 with
 * stride ld_vol; shift (B, 3) float32 or NULL; rot (B, 9) row-major M = R(tracker);
 * work B * N^3 floats (needed when shift is given; used for fractional
 * shifts); partial B * N * 4 floats; out (B, N^3) with stride ld_out.
 * Replaces BatchProcessor._extract_rotated_batch
 * (src/matcha/align_subtomograms.py:240-252): extract_subtomogram_patch_batch
 * with normalize=True (src/matcha/utils/io_utils.py:94-178) and rotate_volumes
 * (src/matcha/utils/volume_rotation.py:9-125). */
int mch_extract_rotated(const float* vols, long long ld_vol, int Nv, int B, int N, const float* shift,
                        const float* rot, double radius, float* work, float* partial, float* out,
                        long long ld_out, hipStream_t stream);

/* Rotation trackers (B, 4) float32 (w, x, y, z), updated in place when
 * alpha/beta/gamma are given: tracker <- normalised(tracker * conj(q~)), q~ the
 * z-y-z quaternion of (alpha, beta, gamma) with its x and z components
 * exchanged.  Then rot_particle = R(tracker) (the next alternation's particle
 * rotation) and rot_reference = R(tracker)^T (the shift search's reference
 * rotation), each (B, 9) row-major or NULL.  Replaces update_rotation_estimate
 * (src/matcha/utils/rotation_ops.py:23-34) and the Euler round trips of
 * rotate_volumes (src/matcha/utils/volume_rotation.py:30,95) and search_shifts
 * (src/matcha/core/ShiftMatcher.py:913). */
int mch_tracker_update(const float* alpha, const float* beta, const float* gamma, int B, float* tracker,
                       float* rot_particle, float* rot_reference, hipStream_t stream);

/* ------------------------------------------------- differentiable score --- */
/* The score s = Re sum_{l<=L} D^l(R(omega)) sigma^l(aT, aV(t)) of the reference's
 * search (src/matcha/core/Matcha.py:184,288) with R(omega) = exp([omega]x) and aV(t)
 * the coefficients of the particle patch at shift t, differentiated with respect to
 * omega, t and aT.  The reference has Euler-angle derivatives only
 * (src/matcha/core/WignerDGradients.py:247-302); these entry points have no direct
 * counterpart. */

/* omega (B, 3) float32 -> z-y-z Euler angles (B, 3) FP64 of exp([omega]x) and the
 * right Jacobian J_r (B, 9) FP64 of exp. */
int mch_score_pose(const float* omega, int B, double* angles, double* jr, hipStream_t stream);

/* Patch at shift t (B, 3) (reference shift order: t0 along the last axis) of each
 * (Nv^3) particle, trilinear with zeros outside, and its derivatives in t0, t1, t2:
 * rows b, B+b, 2B+b, 3B+b of out (4B, N^3) with row stride ld_out.  The patch is
 * extract_subtomogram_patch_batch's (src/matcha/utils/io_utils.py:113-178). */
int mch_shift_tangent(const float* vols, long long ld_vol, int Nv, int B, int N, const float* t, float* out,
                      long long ld_out, hipStream_t stream);

/* Score and gradients: sigma (4B rows, stride ld_sigma >= Dsize(lmax)) of the
 * patch and of its three t-derivatives (as from mch_compute_sigma of their
 * coefficients), idx the coefficient index table (lmax+1, 2lmax+1, kdim) of
 * mch_corr_plan_create (device copy), angles/jr from mch_score_pose, partial
 * B*(L+1)*8 floats of scratch.  out (B, 7) float32: s, ds/domega (3), ds/dt (3).
 * With G (B rows, stride ldG >= ne, complex64) and aV (the patch coefficients,
 * B rows, stride ld_aV): the per-particle template adjoint
 * G_{l k,-v} = sum_u conj(D^l_uv) (-1)^{u+v} aV_{l k,-u} (zero above L). */
int mch_score_gradients(const int* idx, int lmax, int kdim, const void* sigma, long long ld_sigma, int B,
                        const void* aV, long long ld_aV, int L, const double* angles, const double* jr,
                        float* partial, float* out, void* G, long long ldG, hipStream_t stream);

/* mch_score_gradients without the partial argument and without allocating:
 * the per-shell partials live in a static device array of the library
 * (B * (L + 1) <= 131072, 4 MiB per GPU for the life of the process).  No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (HIP graph capture, serving loops).  Calls on one GPU share that
 * array, so calls that overlap on one GPU (different streams or host
 * threads) race; calls ordered on one stream do not. */
int mch_score_gradients_noalloc(const int* idx, int lmax, int kdim, const void* sigma, long long ld_sigma, int B,
                                const void* aV, long long ld_aV, int L, const double* angles, const double* jr,
                                float* out, void* G, long long ldG, hipStream_t stream);

/* Backward of the score under upstream weights (B,): d_omega (B, 3) and d_t (B, 3)
 * (either may be NULL) from out of mch_score_gradients, and with G the template
 * gradient d_aT (ne,) complex64 = sum_b weights[b] G_b (d/dRe + i d/dIm). */
int mch_score_backward(const float* weights, const float* grads, int B, const void* G, long long ldG, int ne,
                       float* d_omega, float* d_t, void* d_aT, hipStream_t stream);

/* Host FP64 Wigner d^l_{m'm}(beta), 0 <= l <= L, D layout (host memory).
 * Replaces spherical.Wigner(L).d (used by src/matcha/core/SOFFT.py:31-39). */
int mch_wigner_d_host(double beta, int L, double* out);

#ifdef __cplusplus
}
#endif

#endif /* MATCHA_AMD_H */
