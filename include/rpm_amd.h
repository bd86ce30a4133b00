/*
 * rpm_amd.h -- C ABI of the MI355X radical-pair solver library (librpm_amd.so).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved FP64
 * (re, im) pairs ("c128").  Pointers named d_* are device pointers that the
 * caller owns (the library never frees or retains them); streams are
 * hipStream_t passed as void*.  Every function returns 0 on success or a
 * negative RPM_E_* code; rpm_last_error() returns the message of the most
 * recent failure on the calling thread.
 *
 * Reference interfaces each entry point replaces (R/ = "Realistic system
 * size/code/" of sam-not-smart/RPM):
 *   rpm_operator_create_pair   CompiledFullPairHamiltonian.__init__   R/rpm_statevector_engine.py:142-219
 *   rpm_operator_create_local  MatrixFreeHamiltonian.__init__         R/rpm_statevector_engine.py:49-93
 *   rpm_operator_apply         .matvec / .matmat                       R/rpm_statevector_engine.py:95-114,221-253
 *   rpm_lanczos_basis          AdaptiveLanczosPropagator._basis        R/rpm_statevector_engine.py:651-688
 *   rpm_krylov_candidates      AdaptiveLanczosPropagator._candidate    R/rpm_statevector_engine.py:690-715
 *                              and _projected_coefficients             R/rpm_statevector_engine.py:569-582
 *   rpm_singlet_gram           SingletPopulationObservable.project_basis R/rpm_statevector_engine.py:521-529
 *   rpm_krylov_outputs         the per-output-time loop of propagate   R/rpm_statevector_engine.py:861-902
 *                              with evaluate_projected                 R/rpm_statevector_engine.py:531-538
 *   rpm_krylov_combine         endpoint update psi <- |psi| V c        R/rpm_statevector_engine.py:904
 *   rpm_norm                   np.linalg.norm(state)                   R/rpm_statevector_engine.py:656,774,909
 *   rpm_singlet_population     singlet_population                      R/rpm_statevector_engine.py:492-497
 *
 * Batched entry points (many trajectories per launch; each replaces the same
 * reference function as its single-trajectory form above, called once per
 * trajectory by the drivers R/rpm_full14_quadrupole_experiment.py:154-194 and
 * R/rpm_quantum_needle_jd_sweeps.py:466-546):
 *   rpm_operator_apply_pairs   .matvec with a per-vector electron 4x4 (field orientation/strength)
 *   rpm_lanczos_basis_batch    AdaptiveLanczosPropagator._basis        R/rpm_statevector_engine.py:651-688
 *   rpm_lanczos_basis_batch_prefix  the same, built in pieces, optionally two steps per read of the basis
 *                              R/rpm_statevector_engine.py:651-688
 *   rpm_krylov_tasks           _candidate / evaluate_projected         R/rpm_statevector_engine.py:690-715,531-538
 *   rpm_singlet_gram_batch     SingletPopulationObservable.project_basis R/rpm_statevector_engine.py:521-529
 *   rpm_krylov_combine_batch   endpoint update psi <- |psi| V c        R/rpm_statevector_engine.py:904
 *   rpm_yield_accumulate       integrate_yield (Simpson) + coarse yield + propagation estimate
 *                              R/rpm_quantum_needle_jd_sweeps.py:418-420, R/rpm_full14_quadrupole_experiment.py:179-183
 *
 * Segmented propagation and adjoint gradients:
 *   rpm_axis_apply             local_unitary_apply                     R/poster_reproduction/core.py:285-292
 *   rpm_operator_apply_affine  Chebyshev/Clenshaw step of the adjoint sweep (no reference counterpart: the
 *                              reference has no gradients; dH/dp is its Zeeman form R/poster_reproduction/core.py:186-194)
 *   rpm_cross_density          electron 4x4 of <lambda| dH/dp |psi> for electron-only dH/dp (no reference counterpart)
 *
 * Factorized spectral oracle (J = D = 0):
 *   rpm_eigh_batched           diagonalize_radical's eigh, batched       R/rpm_quantum_needle_reproduction.py:127-142
 *   rpm_bin_lines              transition_lines + binned_spectrum        R/rpm_quantum_needle_reproduction.py:145-172
 *   rpm_spectral_yields        convolution + yields_from_spectrum sums   R/rpm_quantum_needle_reproduction.py:175-214
 */
#ifndef RPM_AMD_H
#define RPM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPM_OK 0
#define RPM_E_INVALID (-1)   /* bad argument: maps to ValueError */
#define RPM_E_HIP (-2)       /* HIP runtime error */
#define RPM_E_NOMEM (-3)     /* device allocation failed */
#define RPM_E_UNSUPPORTED (-4)

/* Message of the last failure on this thread (NUL-terminated, truncated to len). */
int rpm_last_error(char* buffer, size_t len);

/* Library version string. */
const char* rpm_version(void);

typedef struct rpm_operator rpm_operator;

/*
 * Two-radical compiled Hamiltonian.  State layout (C order):
 *   index = ((eA * 2 + eB) * NL + l) * NR + r
 * with l the C-order multi-index over the left (radical A) nuclear axes and r
 * over the right (radical B) nuclear axes; NL = prod(left_dims), NR = prod(right_dims).
 * nuclear_blocks: for each nucleus (left ones first, then right ones), the
 *   (2d x 2d) complex matrix of every term acting on (its radical's electron,
 *   that nucleus) or on the nucleus alone, row-major in the basis
 *   (electron m_e, nuclear m): row = e * d + m.  Packed back to back.
 * electron_pair: 4x4 complex on (eA, eB), row = eA*2 + eB; it holds the
 *   electron-pair term and the single-electron terms (Zeeman).
 */
int rpm_operator_create_pair(int n_left, const int* left_dims, int n_right, const int* right_dims,
                             const double* nuclear_blocks, const double* electron_pair,
                             rpm_operator** out);

/*
 * Generic sum of local terms on a tensor-product state (any axis order).
 * term_naxes[t] axes per term, term_axes packed, term_ops packed complex
 * (D_t x D_t row-major with D_t = prod of the term's axis dimensions,
 * C order over the term's axes as listed).
 */
int rpm_operator_create_local(int naxes, const int* dims, int nterms, const int* term_naxes,
                              const int* term_axes, const double* term_ops, rpm_operator** out);

int rpm_operator_destroy(rpm_operator* op);

/* Dimension N of the operator's state space (or -1 for a NULL handle). */
long long rpm_operator_dimension(const rpm_operator* op);

/*
 * d_out[b] = H d_in[b] for b < batch; vectors are ld complex elements apart.
 * d_in and d_out must not overlap.
 */
int rpm_operator_apply(const rpm_operator* op, const void* d_in, void* d_out, int batch,
                       long long ld, void* stream);

/*
 * d_out[b] = H_b d_in[b] where H_b is the pair operator with its electron 4x4
 * (electron_pair of rpm_operator_create_pair) replaced by d_pairs[b]
 * (16 complex per vector, row-major, row = eA*2 + eB).  Pair operators only.
 */
int rpm_operator_apply_pairs(const rpm_operator* op, const void* d_in, void* d_out, int batch,
                             long long ld, const void* d_pairs, void* stream);

/*
 * H·psi kernel selection (no counterpart in the reference; for tests and
 * benchmarks): 0 automatic (the LDS-tiled pair kernel when the shape allows
 * it), 1 the per-amplitude gather kernel, 2 the register-tiled kernel, 3 the
 * LDS-tiled kernel, 4 the LDS-tiled kernel that closes the first left axis
 * (dimension 3) in its tile instead of the last two (2-4 give
 * RPM_E_UNSUPPORTED when the operator shape has no such tiled form; the
 * per-vector Lanczos partials of kernel 4 come one per 3-row group, so a
 * workspace sized under another kernel may be too small, which the basis
 * entry points report as RPM_E_INVALID).  rpm_operator_kernel reports the
 * kernel the next apply will use: 1-4 as above, 9 the generic local-term
 * kernel.
 */
int rpm_operator_set_kernel(rpm_operator* op, int kernel);
int rpm_operator_kernel(const rpm_operator* op);

/* Bytes of device workspace rpm_lanczos_basis needs for max dimension m. */
size_t rpm_lanczos_workspace_bytes(const rpm_operator* op, int m);

/*
 * Lanczos basis of d_state (not modified), up to `target` vectors, fully
 * reorthogonalised (three-term step plus one full projection per vector, the
 * projection applied one step late so that each step reads the basis once).
 * d_basis: target * N complex (vector j at offset j*N).
 * d_alpha[target], d_beta[target] (FP64, device), d_info[4] (int, device):
 *   info[0] = basis size actually built (< target on breakdown),
 *   info[2] = explicit corrections of the lagged projection applied (diagnostic),
 * d_norm (device double): receives ||state||.
 * Matvec count on the host side equals info[0] (one per basis vector).
 */
int rpm_lanczos_basis(const rpm_operator* op, const void* d_state, int target, void* d_basis,
                      double* d_alpha, double* d_beta, int* d_info, double* d_norm,
                      double breakdown_tolerance, void* d_workspace, size_t workspace_bytes,
                      void* stream);

/*
 * rpm_lanczos_basis without a workspace argument and without any device allocation: no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as rpm_lanczos_basis (bitwise: the same
 * kernels); the workspace is a 16 MiB array in device memory that the library holds per device for the
 * life of the process, enough for N = 221,184 at target 64.  A basis whose workspace would exceed it
 * gives RPM_E_INVALID.  Every call on a device shares that device's array, so calls that overlap on one
 * device (other host threads or other streams) race and must be serialised by the caller.
 */
int rpm_lanczos_basis_noalloc(const rpm_operator* op, const void* d_state, int target, void* d_basis,
                              double* d_alpha, double* d_beta, int* d_info, double* d_norm,
                              double breakdown_tolerance, void* stream);

/*
 * Batched basis (as rpm_lanczos_basis): trajectory b < batch starts from
 * d_states + b*ld (ld == N) and
 * builds min(d_targets[b], max_target) vectors into d_basis + b*m*N (vector j
 * at + j*N).  d_targets[b] <= 0 leaves trajectory b untouched (its basis,
 * tridiagonal, norm and info keep their values).  d_pairs (nullable): per
 * trajectory electron 4x4 as in rpm_operator_apply_pairs.  d_alpha, d_beta:
 * batch x m (row b at b*m); d_info: batch x 4 (info[4b] = size built,
 * info[4b+1] = 1 if no breakdown, info[4b+2] = explicit corrections); d_norm[batch].
 * max_target <= m <= 64.
 */
size_t rpm_lanczos_batch_workspace_bytes(const rpm_operator* op, int m, int batch);
int rpm_lanczos_basis_batch(const rpm_operator* op, int batch, const void* d_states, long long ld,
                            const int* d_targets, int max_target, const void* d_pairs, void* d_basis,
                            int m, double* d_alpha, double* d_beta, int* d_info, double* d_norm,
                            double breakdown_tolerance, void* d_workspace, size_t workspace_bytes,
                            void* stream);

/*
 * rpm_lanczos_basis_batch in pieces (same reference function, R/rpm_statevector_engine.py:651-688): the
 * first vectors of a basis now, the rest in a later call.  Trajectory b with d_targets[b] > 0 starts a
 * new basis (eventual size min(d_targets[b], max_target)) from d_states + b*ld and builds its first
 * d_stops[b] vectors (all of them when d_stops[b] <= 0); with d_targets[b] == 0 and d_stops[b] > 0 it
 * continues the basis an earlier call of this function left in the same workspace and buffers, up to
 * d_stops[b] vectors; otherwise it is left untouched.  Lanczos steps j_begin <= j < max_stop are
 * launched (j_begin no larger than the smallest size a continued basis has, 0 when a basis starts;
 * max_stop the largest d_stops).  info[4b] = vectors built so far (less than d_stops[b] on breakdown),
 * info[4b+1] = 1 if no breakdown, info[4b+2] = explicit corrections, info[4b+3] = 1 if the call stopped
 * the basis early (see flags); the tridiagonal of the built vectors is in d_alpha/d_beta as for the
 * whole basis.  The workspace is rpm_lanczos_prefix_workspace_bytes(op, m, batch) bytes.
 * flags = 0: each step reads the basis once (bitwise equal to one uninterrupted build).
 * flags = RPM_LANCZOS_PAIRED: from step 12 on, steps j and j+1 (j even) share one read of the basis (the
 * second step's projections formed by linearity from H applied to the unprojected three-term vector);
 * every call still ends with its vectors materialised, so calls of either kind continue each other's
 * bases.  A near-breakdown inside a pair stops that trajectory's basis early with info[4b+3] = 1 and
 * info[4b+1] = 1: the vectors so far are valid, and a further call (d_targets[b] = 0) continues it.
 * The result is deterministic per trajectory (independent of the other trajectories of the batch).
 * flags |= RPM_LANCZOS_FORCE_SHORT (with RPM_LANCZOS_PAIRED; a test hook): every pair takes that
 * early stop at its first step, so the continuation path can be exercised; the basis built over the
 * continuing calls is the same to round-off.
 */
#define RPM_LANCZOS_PAIRED 1
#define RPM_LANCZOS_FORCE_SHORT 2
size_t rpm_lanczos_prefix_workspace_bytes(const rpm_operator* op, int m, int batch);
int rpm_lanczos_basis_batch_prefix(const rpm_operator* op, int batch, const void* d_states, long long ld,
                                   const int* d_targets, int max_target, const int* d_stops, int j_begin,
                                   int max_stop, const void* d_pairs, void* d_basis, int m, double* d_alpha,
                                   double* d_beta, int* d_info, double* d_norm, double breakdown_tolerance,
                                   int flags, void* d_workspace, size_t workspace_bytes, void* stream);

/*
 * Batched small-matrix evaluations.  d_tasks: ntasks records of 32 bytes
 *   { int32 b, dim, low, out; float64 t; int32 kind, pad; }
 * evaluating c = exp(-i T_dim t) e_1 of trajectory b's tridiagonal
 * (d_alpha/d_beta + b*ab_stride) with dim <= that trajectory's basis size.
 *   kind 0 (candidate): d_v0[out] = safety*norm[b]*||c - pad(c_low)|| (0 if low <= 0 or low >= dim)
 *   kind 1 (output):    d_v0[out] = norm[b]^2 Re(c^H P_b c) with P_b = d_P + b*p_stride
 *                       (dim x dim row-major), d_v1[out] = the point error as above
 * d_coeffs (nullable): row out (m_stride complex) receives c padded with zeros.
 */
int rpm_krylov_tasks(const void* d_tasks, int ntasks, const double* d_alpha, const double* d_beta,
                     int ab_stride, const double* d_norm, const void* d_P, long long p_stride,
                     double safety, double* d_v0, double* d_v1, void* d_coeffs, int m_stride,
                     void* stream);

/*
 * Singlet yield accumulated on the device as the outputs are produced (R/rpm_quantum_needle_jd_sweeps.py:418-420,
 * Phi = simpson(exp(-t/tau) P_S / tau, t), and the coarse-grid yield and propagation estimate of
 * R/rpm_full14_quadrupole_experiment.py:179-183).  For each trajectory b < batch, d_spans[3b..3b+2] =
 * {row0, count, w0} selects its outputs P_i = d_values[row0 + i] (point errors d_point_err[row0 + i],
 * nullable = 0) and weight rows W_i = d_weights[2 (w0 + i) .. +1] = (fine, coarse), i < count, and
 *   d_phi[3b]   += sum_i W_i.fine P_i
 *   d_phi[3b+1] += sum_i W_i.coarse P_i
 *   d_phi[3b+2] += sum_i W_i.fine min(1, 2 e_i + e_i^2),  e_i = d_acc[b] (nullable = 0) + point error.
 * The weights are the Simpson weights of the output grid times exp(-t/tau)/tau (coarse: those of the
 * coarse sub-grid, 0 elsewhere).  The sum order depends only on the trajectory's own outputs.  Rows or
 * weights beyond nvalues / nweights contribute nothing.
 */
int rpm_yield_accumulate(const double* d_values, const double* d_point_err, long long nvalues,
                         const long long* d_spans, int batch, const double* d_weights, long long nweights,
                         const double* d_acc, double* d_phi, void* stream);

/* Batched rpm_singlet_gram: P_b (k_b x k_b at d_P + b*m*m) of the first k_b = min(d_k[b], m)
 * vectors of basis b (d_basis + b*m*N), m <= 64; k_b <= 0 skips b. */
size_t rpm_singlet_gram_batch_workspace_bytes(long long nn, int batch);
int rpm_singlet_gram_batch(const void* d_basis, int m, const int* d_k, int batch, long long nn, void* d_P,
                           void* d_workspace, size_t workspace_bytes, void* stream);

/* Batched rpm_krylov_combine: d_out + b*out_ld = scale[b] * sum_{j<k_b} basis_b[j] c_b[j],
 * basis_b = d_basis + b*m*n, c_b = d_c + b*c_stride, k_b = min(d_k[b], m, c_stride) (<= 0 skips b). */
int rpm_krylov_combine_batch(const void* d_basis, int m, const int* d_k, int batch, long long n,
                             const void* d_c, int c_stride, const double* d_scale, void* d_out,
                             long long out_ld, void* stream);

/*
 * d_out[b] = s * H_b d_in[b] + a_in * d_in[b] + (c0_re + i c0_im) * d_x0[b] + c1 * d_x1[b]
 * (pair operators; H_b as in rpm_operator_apply_pairs when d_pairs is given).  d_x0 / d_x1
 * are nullable; d_x1 may be d_out (in place), d_in may alias none of the others.
 */
int rpm_operator_apply_affine(const rpm_operator* op, const void* d_in, void* d_out, int batch, long long ld,
                              const void* d_pairs, double s, double a_in, const void* d_x0, double c0_re,
                              double c0_im, const void* d_x1, double c1, void* stream);

/* d_acc[b*16 + e*4 + f] += weight * sum_n conj(a_b[e*nn + n]) b_b[f*nn + n]  (e, f = eA*2 + eB). */
size_t rpm_cross_density_workspace_bytes(long long nn, int batch);
int rpm_cross_density(const void* d_a, const void* d_b, int batch, long long ld, long long nn, double weight,
                      void* d_acc, void* d_workspace, size_t workspace_bytes, void* stream);

/* d_out[b] = (1 (x) U (x) 1) d_in[b]: U (d x d complex, row-major) on tensor axis `axis` of the
 * C-order state with dimensions dims[naxes]; d = dims[axis] <= 16. */
int rpm_axis_apply(int naxes, const int* dims, int axis, const void* d_U, const void* d_in, void* d_out,
                   int batch, long long ld, void* stream);

/*
 * Adjoint sweep of the Simpson singlet yield Phi = sum_k W_k <psi_k|P|psi_k> on an output grid
 * (no reference counterpart; the reference computes Phi only, R/rpm_quantum_needle_jd_sweeps.py:418-420).
 * The grid is cut into nmacro steps [t_a, t_b]; step k has class cls[k] (series length
 * nterms[c] = K <= kmax <= 255, coefficients coef[c][0..K] (host, interleaved complex) of
 * exp(-i H (t_b - t_a)) = sum_n c_n T_n(X), X = (H - shift)/half_width, and the (kmax+1)^2 complex
 * matrix d_G[c] (device) with G[m][j] = sum_{kicks t in (t_a, t_b]} 2 w(t) conj(c_m(t - t_a)) c_j(t - t_a)),
 * a weight multiplier mult[k] (W = mult[k] w), electron-matrix segment seg[k] < nseg (d_pairs: nseg x batch
 * x 16 complex, nullable; seg and nseg are read only with d_pairs) and accumulator slot slot[k] < nslot
 * (d_acc: nslot x batch x 16 complex).  Out-of-range classes, slots or segments give RPM_E_INVALID.  Running backward from d_psi = psi(t_end)
 * (overwritten with psi(0)), each step carries psi back by Clenshaw's recurrence, forms
 * u_j = T_j(X) psi(t_a), and runs Clenshaw over the costate with the vector seeds
 * z_m = conj(c_m) g + mult P sum_j G[m][j] u_j, accumulating
 *   d_acc[slot][b] += sum_n w_n / half_width * C(nu_{n+1}, u_n)   (w_0 = 1, w_n = 2,
 * C as in rpm_cross_density), so that dPhi/dp = Re sum_{e,f} (dH_b/dp)[e][f] acc[e][f] for every
 * electron-only dH/dp.
 */
size_t rpm_adjoint_workspace_bytes(const rpm_operator* op, int batch, int kmax);
int rpm_adjoint_sweep(const rpm_operator* op, int batch, void* d_psi, int nmacro, double shift,
                      double half_width, int ncls, const int* nterms, int kmax, const double* coef,
                      const void* d_G, const int* cls, const double* mult, const void* d_pairs,
                      const int* seg, int nseg, void* d_acc, const int* slot, int nslot, void* d_workspace,
                      size_t workspace_bytes, void* stream);

/*
 * Factorized spectral yields (R/rpm_quantum_needle_reproduction.py:145-172, transition_lines and
 * binned_spectrum): the transition lines of one radical (or of one symmetry block of it), frequency
 * lambda_i - lambda_j and weight Tp[i][j] conj(Tq[i][j]) / weight_divisor, deposited by linear
 * interpolation at grid index (lambda_i - lambda_j) / delta + half, for the six pairs
 * (p, q) = (0,0) (1,1) (2,2) (0,1) (0,2) (1,2). weight_divisor is the radical's nuclear dimension,
 * divided by the block's multiplicity when a block stands for several identical ones.
 * d_eigenvalues: norient x d, ascending per orientation. d_ops: norient x 3 x d x d complex (T_x, T_y,
 * T_z in the eigenbasis, row-major, Hermitian). d_grid: norient x 6 x grid_stride complex: with
 * accumulate = 0 written in full (bins 0 .. 2 half hold the spectrum, the rest zeros); with
 * accumulate = 1 the lines are added to what it holds. If a line falls outside bins 0 .. 2 half,
 * *d_overflow (device int) is incremented.
 */
int rpm_bin_lines(const double* d_eigenvalues, const void* d_ops, int d, int norient, double weight_divisor,
                  double delta, long long half, void* d_grid, long long grid_stride, int accumulate, int* d_overflow,
                  void* stream);

/*
 * Yields of the convolved spectra by Parseval (R/rpm_quantum_needle_reproduction.py:175-214, the FFT
 * product, inverse FFT and yields_from_spectrum sums): for orientation b and lifetime t,
 *   partial[blk][b][t] summed over blk = sum_f Re(cd_b(f) LD_t(f) + co_b(f) LO_t(f)),
 * cd_b = sum over pairs 0..2 of FA_b,k FB_b,k and co_b over pairs 3..5 (the rpm_bin_lines order).
 * d_fa, d_fb: norient x 6 x n complex (FFTs of the two radicals' binned spectra); d_ld, d_lo:
 * nlifetimes x n complex (inverse FFTs of K_t and of 2 Re K_t on the output grid, zero beyond it).
 * d_partial: nblocks x norient x nlifetimes doubles, written. nlifetimes <= 8.
 */
int rpm_spectral_yields(const void* d_fa, const void* d_fb, const void* d_ld, const void* d_lo, int norient,
                        long long n, int nlifetimes, int nblocks, double* d_partial, void* stream);

/*
 * Batched Hermitian eigen-decomposition of row-major n x n complex matrices d_a (batch of them,
 * contiguous), in place: on return d_a[b] holds V^H (row k = conjugated eigenvector k) and
 * d_w[b*n + k] the eigenvalues in ascending order. method 0: divide and conquer (heevd), 1: Jacobi
 * (heevdj). The per-matrix status (0 = converged) is left in the first batch ints of d_ws.
 * Replaces scipy.linalg.eigh(..., driver="evd") of R/rpm_quantum_needle_reproduction.py:135-140,
 * one matrix per orientation, with one call for the batch.
 */
size_t rpm_eigh_workspace_bytes(int n, int batch);
int rpm_eigh_batched(void* d_a, int n, int batch, double* d_w, void* d_ws, size_t ws_bytes, int method, void* stream);

/*
 * Embedded error estimates for a ladder of candidate dimensions.
 * For each candidate i: dim = min(dims[i], size); low = max(1, dim - offset);
 *   err[i] = safety * state_norm * || c_dim(t) - pad(c_low(t)) ||  (0 if low == dim)
 * with c_k(t) = exp(-i T_k t) e_1 from the device tridiagonal eigensolver.
 * d_coeffs: ncand * m_stride complex, c_dim(t) of each candidate (padded with 0).
 * Host arrays: dims[ncand]; d_alpha/d_beta/d_size/d_norm are device pointers.
 * d_err: device double[ncand].
 */
int rpm_krylov_candidates(const double* d_alpha, const double* d_beta, const int* d_info,
                          const double* d_norm, int ncand, const int* dims, int offset,
                          double safety, double t, int m_stride, double* d_err, void* d_coeffs,
                          void* d_scratch, void* stream);

/* Bytes of scratch rpm_krylov_candidates / rpm_krylov_outputs need for m_stride. */
size_t rpm_krylov_scratch_bytes(int ncand, int m_stride);

/* P = A^H A with A = (V[e=01] - V[e=10]) / sqrt2 over the nuclear index, Hermitized.
 * d_basis: k vectors of length N (= 4 * nn); d_P: k*k complex row-major. */
int rpm_singlet_gram(const void* d_basis, int k, long long nn, void* d_P, void* d_workspace,
                     size_t workspace_bytes, void* stream);
size_t rpm_singlet_gram_workspace_bytes(int k, long long nn);

/*
 * Values at output times inside one accepted step.
 * sel: accepted dimension; low = max(1, sel - offset), no low model if exact or low == sel.
 * For each offset t_i: c = c_sel(t_i); P_S = norm^2 Re(c^H P c);
 *   point_err = safety * norm * || c - pad(c_low(t_i)) ||.
 * low = 0 means no embedded model (point_err = 0).
 * d_values / d_point_err: device double[nt].  d_coeffs_out (nullable): nt x sel
 * complex, the coefficient vectors c_sel(t_i) (for point states).
 */
int rpm_krylov_outputs(const double* d_alpha, const double* d_beta, const double* d_norm,
                       const void* d_P, int sel, int low, int nt, const double* d_offsets,
                       double safety, double* d_values, double* d_point_err, void* d_coeffs_out,
                       void* d_scratch, size_t scratch_bytes, void* stream);

/* d_out = scale * sum_{j<k} d_basis[j] * c[j]; c: device complex[k]; scale: device double. */
int rpm_krylov_combine(const void* d_basis, int k, long long n, const void* d_c,
                       const double* d_scale, void* d_out, void* stream);

/* d_out[b] = ||x_b|| for b < batch (FP64). */
int rpm_norm(const void* d_x, long long n, int batch, long long ld, double* d_out,
             void* d_workspace, size_t workspace_bytes, void* stream);

/* d_out[b] = singlet population of x_b (N = 4 * nn). */
int rpm_singlet_population(const void* d_x, long long nn, int batch, long long ld, double* d_out,
                           void* d_workspace, size_t workspace_bytes, void* stream);

/* Workspace for rpm_norm / rpm_singlet_population. */
size_t rpm_reduce_workspace_bytes(long long n, int batch);

#ifdef __cplusplus
}
#endif
#endif /* RPM_AMD_H */
