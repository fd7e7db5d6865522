/*
 * taurus_gpu.h -- C ABI of the MI355X (gfx950) PNVAP projected-gradient library.
 *
 * Drop-in boundary for the hot path of TAURUS_vap: the gauge-angle loop of
 * module_projection, the two-body contraction of module_fields and the gradient
 * step of module_gradient.  Every entry point is a plain extern "C" function on
 * host pointers and sizes, so the reference's Fortran host binds it with
 * iso_c_binding (bind(C) interfaces, `value` for scalars).
 *
 * Conventions (those of the reference's module arrays):
 *   - matrices are column-major (Fortran order), leading dimension ndim unless
 *     an lda argument is given;
 *   - complex(r64) arrays are interleaved (re, im) doubles, i.e. `double _Complex`
 *     / complex(c_double_complex);
 *   - single-particle indices in abcd/sp_tr are 1-based, as stored by the
 *     reference;
 *   - return value: 0 on success, > 0 a LAPACK-style info, < 0 an error
 *     (TG_E_*); tg_last_error() gives a message.
 * The library owns its device buffers; no host pointer is retained after a
 * call returns.  A context is not re-entrant: call it from one host thread.
 */
#ifndef TAURUS_GPU_H
#define TAURUS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK                0
#define TG_E_ARG            -1   /* invalid argument */
#define TG_E_HIP            -2   /* HIP runtime error */
#define TG_E_NOMEM          -3   /* device allocation failed */
#define TG_E_NODEVICE       -4   /* no HIP device */
#define TG_E_IMAG_GRADIENT  -5   /* calculate_projected_H20: Im(gradient) > 1e-8 */
#define TG_E_SINGULAR       -6   /* singular overlap matrix (zgetrf info /= 0) */
#define TG_E_UNSUPPORTED    -7   /* a case the device routine leaves to the host */

/* Context flags */
#define TG_FLAG_PLAIN        1   /* one gauge angle per launch sequence, no batching */
#define TG_FLAG_TBME_FP32    2   /* keep TBMEs in FP32 on the device (FP64 accumulation) */
#define TG_FLAG_GRAPH        4   /* replay the projection as a captured hipGraph */
#define TG_FLAG_CONJ_PAIRS   8   /* [0, 2 pi) grids: compute one angle of each (phi, -phi)
                                    pair (per-angle quantities of real U, V at -phi are the
                                    complex conjugates) and add it as 2 Re(.) */

typedef struct tg_ctx tg_ctx;
typedef struct tg_group tg_group;

/* Message of the last failing call on this thread. */
const char* tg_last_error(void);

/*
 * tg_init -- basis + Hamiltonian upload, once per run.
 * Replaces the storage set up by set_hamiltonian / read_hamiltonian_reduced
 * (src/module_hamiltonian.f90:79-106, 533-600: hamil_H0, hamil_H1, hamil_H2dim,
 * hamil_abcd, hamil_H2, hamil_trperm, :53-58) and the basis arrays
 * HOsp_2mj, HOsp_2mt, HOsp_l, HOsp_tr (src/module_basis.f90:33-41, 209-238).
 *   H1      complex ndim x ndim
 *   abcd    int16[4*ntb] (1-based a,b,c,d per stored element)
 *   H2      ntb values, float (h2_bytes = 4, rH2 = real32) or double (8)
 *   trperm  int8[ntb] time-reversal permutation codes
 * flags: TG_FLAG_* .
 */
int tg_init(tg_ctx** ctx, int ndim, const int32_t* sp_2mj, const int32_t* sp_2mt,
            const int32_t* sp_l, const int32_t* sp_tr, double H0_re, double H0_im,
            const double* H1, int64_t ntb, const int16_t* abcd, const void* H2,
            int h2_bytes, const int8_t* trperm, int flags);

/* Frees the context and its device memory. */
int tg_finalize(tg_ctx* ctx);

/* Switches between batched (flags = 0) and plain per-angle (TG_FLAG_PLAIN) execution. */
int tg_set_mode(tg_ctx* ctx, int plain);

/*
 * tg_set_batch_limit -- at most max_items (state, angle) items per launch
 * sequence of the batched projection (0: no limit, the default).  Independently
 * of it, when the per-item work matrices of all items do not fit into the free
 * device memory, the projection runs the items chunk by chunk and keeps only
 * A_phi and H20_phi of every item for sum_gauge (the memory-short mode).  With
 * several states a chunk holds whole states.  tg_project_observables needs all
 * angles at once and returns TG_E_NOMEM otherwise.
 */
int tg_set_batch_limit(tg_ctx* ctx, int max_items);

/*
 * tg_calculate_fields -- calculate_fields(rhoLR,kappaLR,kappaRL,gammaLR,hspLR,
 * deltaLR,deltaRL,ndim), src/module_fields.f90:99-189.  All complex ndim x ndim.
 */
int tg_calculate_fields(tg_ctx* ctx, const double* rhoLR, const double* kappaLR,
                        const double* kappaRL, double* gammaLR, double* hspLR,
                        double* deltaLR, double* deltaRL);

/*
 * tg_calculate_fields_diag -- calculate_fields_diag(rhoLR,kappaLR,gammaLR,hspLR,
 * deltaLR,deltaRL,ndim), src/module_fields.f90:204-317 (deltaRL may be NULL,
 * the Fortran optional argument).
 */
int tg_calculate_fields_diag(tg_ctx* ctx, const double* rhoLR, const double* kappaLR,
                             double* gammaLR, double* hspLR, double* deltaLR,
                             double* deltaRL);

/*
 * tg_inverse -- the ZGETRF + ZGETRI pair of calculate_thouless
 * (src/module_projection.f90:702-705) and calculate_overlap (:860-883) for
 * `count` complex matrices of order n at A + k*lda*n, replaced in place by their
 * inverses (Gauss-Jordan with the partial pivoting of ZGETRF: first maximum of
 * |Re|+|Im|).  det: complex[count], the determinants (the product of the pivots
 * with the interchange signs, as calculate_overlap forms detR :866-877);
 * info: int[count], 0 or the first zero pivot column (1-based, ZGETRF's info).
 */
int tg_inverse(int n, int count, double* A, int lda, double* det, int* info);

/*
 * tg_zskpfa -- ZSKPFA('U','P',n,A,lda,pfaff,...) of src/subroutines_pfaffian.f:12-256
 * (Parlett-Reid, upper triangle read, INFO > 0 from a zero pivot column gives
 * pfaff = 0 and info = 0).  `count` matrices laid out at A + k*lda*n (complex).
 * pfaff: complex[count]; info: int[count].  A is not modified.
 */
int tg_zskpfa(int n, int count, const double* A, int lda, double* pfaff, int* info);

/*
 * tg_zskpfa_noalloc -- tg_zskpfa without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Every device buffer of the call (the copy of A, the work matrices, the
 * per-matrix results) comes from one static device array of
 * TG_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a call that
 * needs more (count * 16 * (lda*n + n*n) bytes plus a few per matrix) returns
 * TG_E_ARG.  Calls on one device share that array: calls that overlap on one
 * device race.  Results are those of tg_zskpfa.
 */
#define TG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tg_zskpfa_noalloc(int n, int count, const double* A, int lda, double* pfaff, int* info);

/*
 * tg_calculate_overlap -- calculate_overlap(nocc0,nemp0,ovac0,voveru0,D0,nocc1,
 * nemp1,ovac1,voveru1,D1,ROT,overlap,ndim), src/module_projection.f90:782-937.
 * D0, D1, ROT complex ndim x ndim; voveru0 real[ndim-nocc0], voveru1 real[ndim-nocc1].
 * Returns TG_E_SINGULAR where the reference stops on zgetrf/zgetri info /= 0.
 */
int tg_calculate_overlap(int ndim, int nocc0, int nemp0, double ovac0, const double* voveru0,
                         const double* D0, int nocc1, int nemp1, double ovac1,
                         const double* voveru1, const double* D1, const double* ROT,
                         double* overlap);

/*
 * tg_project_wavefunction -- project_wavefunction(0,iprint,Mphip,Mphin,ndim)
 * followed by calculate_projected_H20(ndim), src/module_projection.f90:201-373,
 * 729-759, for the real state (U0, V0) (bogo_U0/bogo_V0).
 * The canonical basis of the state is an input, as built once per iteration by
 * construct_canonical_basis (src/module_wavefunctions.f90:777-945, called at
 * module_projection.f90:219-234): D0 real ndim x ndim (bogo_zD0), nocc0, nemp0,
 * ovac0 and voveru0 real[ndim-nocc0].
 * iprint > 0 also computes the parity overlaps rot_pari (:279-284) and pnp_pari.
 * facpi = proj_facpi; valence_Z/N; factor_delta (module_fields.f90:39).
 * Outputs:
 *   pnp_scalars  complex[7]: pnp_over, pnp_ener, pnp_prot, pnp_neut, pnp_prot2,
 *                pnp_neut2, pnp_pari (0 unless iprint > 0)
 *   pnp_A, pnp_H20, pnp_AH20 complex ndim x ndim (may be NULL)
 *   field_H20v   real[ndim*ndim]
 *   imag_grad    (may be NULL) the largest Im part checked by :751-755
 * Returns TG_E_IMAG_GRADIENT where the reference stops.
 */
int tg_project_wavefunction(tg_ctx* ctx, int iprint, const double* U0, const double* V0,
                            const double* D0, int nocc0, int nemp0, double ovac0,
                            const double* voveru0, int Mphip, int Mphin, double facpi,
                            double valence_Z, double valence_N, double factor_delta,
                            double* pnp_scalars, double* pnp_A, double* pnp_H20,
                            double* pnp_AH20, double* field_H20v, double* imag_grad);

/*
 * tg_project_resident -- same computation on a state already uploaded with
 * tg_project_wavefunction (or tg_set_state); only the device work runs, results
 * stay on the device until tg_get_results.  Used for timing the hot path with
 * inputs resident in HBM.
 */
int tg_set_state(tg_ctx* ctx, const double* U0, const double* V0, const double* D0,
                 int nocc0, int nemp0, double ovac0, const double* voveru0, int Mphip,
                 int Mphin, double facpi, double valence_Z, double valence_N,
                 double factor_delta, int iprint);
int tg_project_resident(tg_ctx* ctx);
int tg_sync(tg_ctx* ctx);
int tg_get_results(tg_ctx* ctx, double* pnp_scalars, double* pnp_A, double* pnp_H20,
                   double* pnp_AH20, double* field_H20v, double* imag_grad);

/*
 * tg_set_states -- nstates independent states of this basis and Hamiltonian,
 * projected together (every kernel of the angle loop batches the nstates x A
 * items; the gauge sums, H20 and the gradient are per state).  The reference
 * runs one state per process; this is the many-solves case of the same
 * project_wavefunction(0, ...) (src/module_projection.f90:201-373).  State k:
 * U0 + k*ndim*ndim, V0 + ..., D0 + ..., nocc0[k], nemp0[k], ovac0[k], and
 * voveru0 + k*ndim holding its ndim - nocc0[k] values.  The angle grid,
 * valence numbers and factor_delta are common.  tg_set_state is nstates = 1.
 * Not with TG_FLAG_PLAIN.
 * tg_get_results_state / tg_get_gradient_state -- the results of state k.
 */
int tg_set_states(tg_ctx* ctx, int nstates, const double* U0, const double* V0,
                  const double* D0, const int* nocc0, const int* nemp0, const double* ovac0,
                  const double* voveru0, int Mphip, int Mphin, double facpi, double valence_Z,
                  double valence_N, double factor_delta, int iprint);
int tg_get_results_state(tg_ctx* ctx, int state, double* pnp_scalars, double* pnp_A,
                         double* pnp_H20, double* pnp_AH20, double* field_H20v,
                         double* imag_grad);

/*
 * tg_calculate_gradient -- calculate_gradient(ndim), src/module_gradient.f90:79-143.
 * opt_grad selection: gradient_type with the info_H11/info_hsp fall-back (:86-91);
 * eigen_H11/eigen_hsp real[ndim] (used by gradient_type 1/2, may be NULL for 0);
 * constraint_20 real[constraint_dim*ndim*ndim], lagrange_lambda1 real[constraint_dim];
 * gradient_Zim1 in, gradient_Zi out (real[ndim*ndim]); gradient_norm out.
 */
int tg_calculate_gradient(int ndim, int gradient_type, double gradient_eta, double gradient_mu,
                          int info_H11, int info_hsp, const double* eigen_H11,
                          const double* eigen_hsp, const double* field_H20v,
                          int constraint_dim, const double* constraint_20,
                          const double* lagrange_lambda1, const double* gradient_Zim1,
                          double* gradient_Zi, double* gradient_norm);

/*
 * tg_set_gradient / tg_get_gradient -- calculate_gradient with the step
 * parameters of gradient_type 0 (src/module_gradient.f90:97-141) run on the
 * device right after each tg_project_resident: Z = -eta (field_H20v -
 * sum_j lambda_j constraint_20_j) + mu Z_prev, with Z_prev the previous step's
 * Z (gradient_Zim1 = gradient_Zi, taurus_vap.f90:104); gradient_norm = ||Z||_F/eta.
 */
int tg_set_gradient(tg_ctx* ctx, double gradient_eta, double gradient_mu, int constraint_dim,
                    const double* constraint_20, const double* lagrange_lambda1);
int tg_get_gradient(tg_ctx* ctx, double* gradient_Zi, double* gradient_norm);
int tg_get_gradient_state(tg_ctx* ctx, int state, double* gradient_Zi, double* gradient_norm);

/*
 * tg_set_constraints -- the constraint operators of set_constraints
 * (src/module_constraints.f90:77-258) for the device iteration below:
 * constraint_HO real[constraint_dim*ndim*ndim] (column-major HO matrices in the
 * order of the sp arrays given to tg_init, one-body operators first and the
 * constraint_pair pair operators last, as src/module_constraints.f90:104-250
 * stores them), constraint_val real[nsets*constraint_dim] (constraint_val of
 * :77-258; nsets = 1: the same values for every state, or one set per state of
 * tg_set_states), constraint_eps the tolerance of adjust_constraints.
 * constraint_dim <= 64.
 */
int tg_set_constraints(tg_ctx* ctx, int constraint_dim, int constraint_pair,
                       const double* constraint_HO, int nsets, const double* constraint_val,
                       double constraint_eps);

/*
 * tg_iterate -- one pass of the gradient loop of src/taurus_vap.f90:77-113 for
 * gradient_type 0 with the Bogoliubov matrices kept in HBM:
 *   update_constraints_qpbasis      src/module_constraints.f90:381-397
 *   project_wavefunction(0, ...)    src/module_projection.f90 (as tg_project_resident)
 *   calculate_lagrange_multipliers  src/module_constraints.f90:413-484
 *   calculate_gradient              src/module_gradient.f90:79-143
 *   adjust_constraints              src/module_constraints.f90:270-373
 *   gradient_Zim1 = gradient_Zi     src/taurus_vap.f90:113
 * Every resident state takes the step at once (the adjustment loops run in
 * lockstep; a state whose constraints are met adds nothing more to its Z).
 * Needs tg_set_state(s) (states and canonical bases), tg_set_gradient (eta,
 * mu; its constraint arguments are replaced by the device's constraint_20)
 * and tg_set_constraints.  Outputs per state (each pointer may be NULL):
 * pnp_scalars complex[7*nstates] of the projected states (as tg_get_results),
 * gradient_norm[nstates], lagrange_lambda1 real[nstates*constraint_dim],
 * adjust_iterations[nstates], converged[nstates] (gradient_norm <=
 * gradient_eps: that state is left unchanged).  The other states become U1, V1
 * of adjust_constraints, and a new tg_set_state(s) (with the canonical bases,
 * construct_canonical_basis, src/module_wavefunctions.f90) must precede the
 * next tg_iterate.  A failed Cholesky factorisation returns TG_E_SINGULAR where
 * the reference stops; a singular multiplier system keeps the previous
 * multipliers (:471-477).  tg_canonical_basis sets the new states without
 * leaving the device.
 */
int tg_iterate(tg_ctx* ctx, double gradient_eps, double* pnp_scalars, double* gradient_norm,
               double* lagrange_lambda1, int* adjust_iterations, int* converged);

/*
 * tg_iterate_canonical -- tg_canonical_basis of the resident states followed by
 * tg_iterate: the loop body of src/taurus_vap.f90:77-123 with the canonical basis
 * of src/module_wavefunctions.f90:777-945 at its start, as project_wavefunction
 * builds it (src/module_projection.f90:212-234).  The results equal the two calls
 * in sequence; the canonical basis runs on the context's second stream beside the
 * projection's Thouless / field / H20 chain (only the overlap chain reads it).
 * Outputs as tg_iterate and tg_canonical_basis.  TG_E_UNSUPPORTED (a degenerate
 * kappa block wider than the device form) leaves the state as it was: the host
 * route (tg_get_state, construct_canonical_basis, tg_set_state) and tg_iterate
 * follow.
 */
int tg_iterate_canonical(tg_ctx* ctx, double canon_eps, double gradient_eps, double* pnp_scalars,
                         double* gradient_norm, double* lagrange_lambda1, int* adjust_iterations,
                         int* converged, int* nocc, int* nemp, double* ovacc, double* voveru);

/*
 * tg_set_gradient_type -- gradient_type of the device iteration (tg_iterate):
 * 0 = the eta, mu of tg_set_gradient; 1 / 2 = the heavy-ball recipes of
 * calculate_gradient (src/module_gradient.f90:97-120) from the H11 / h spectra
 * that diagonalize_hsp_and_H11(0, ndim) (:158-453) computes every iteration,
 * here on the device (h and Delta of calculate_fields_diag on rhoRR, kappaRR;
 * calculate_H11_real, module_fields.f90:327-353; the lambda_m C11 and lambda_m
 * constraint_HO terms; the eigenvalues of h and H11).  The eta, mu of
 * tg_set_gradient are not used by types 1, 2 (as in the reference, they are
 * the fall-back of a failed diagonalisation, which cannot happen here).
 */
int tg_set_gradient_type(tg_ctx* ctx, int gradient_type);

/*
 * tg_get_hsp_h11_eigenvalues -- after a tg_iterate with gradient_type 1 or 2:
 * the spectra of h and H11 of state `state` (eigen_hsp, eigen_H11 of
 * module_gradient.f90:35-36, real[ndim], ascending) and the eta, mu of its step.
 * Any pointer may be NULL.
 */
int tg_get_hsp_h11_eigenvalues(tg_ctx* ctx, int state, double* eigen_hsp, double* eigen_H11,
                               double* gradient_eta, double* gradient_mu);

/*
 * tg_diagonalize_hsp_and_H11 -- diagonalize_hsp_and_H11(0, ndim),
 * src/module_gradient.f90:158-453 (opt = 0: what calculate_gradient needs), for
 * the resident state (one state; tg_project_wavefunction or tg_set_state) with
 * the constraint operators of tg_set_constraints (constraint_dim must match)
 * and the multipliers lagrange_lambda1 real[constraint_dim]:
 *   eigen_hsp, eigen_H11   real[ndim], ascending (the DSYEV of :297, :416)
 *   info_hsp, info_H11     0 (the device solver has no failure mode)
 *   hsp_RR, field_H11      (may be NULL) real ndim x ndim: the matrices handed to
 *                          the eigenvalue solver, field_hspRR and field_H11 of
 *                          :181-213 (constraint terms removed)
 */
int tg_diagonalize_hsp_and_H11(tg_ctx* ctx, int constraint_dim, const double* lagrange_lambda1,
                               double* eigen_hsp, double* eigen_H11, int* info_hsp, int* info_H11,
                               double* hsp_RR, double* field_H11);

/*
 * tg_sym_eigenvalues -- eigenvalues (ascending) of `count` real symmetric
 * matrices of order n <= 512 at A + k*lda*n (upper triangles read): the
 * eigenvalues DSYEV('V','U') returns at src/module_gradient.f90:297 and :416.
 * w: real[count*n].  A is not modified.
 */
int tg_sym_eigenvalues(int n, int count, const double* A, int lda, double* w);

/*
 * tg_sym_eigensystem -- eigenvalues (ascending) and orthonormal eigenvectors of
 * count real symmetric n x n matrices (upper triangles read; host buffers: A
 * count x lda x n, w count x n, Z count x n x n, column-major), as DSYEV('V','U')
 * returns them for the canonical basis (construct_canonical_basis,
 * src/module_wavefunctions.f90:806): Householder tridiagonalization, divide and
 * conquer, back-transformation, all on the device.  n <= 1024.
 */
int tg_sym_eigensystem(int n, int count, const double* A, int lda, double* w, double* Z);

/*
 * tg_get_state -- bogo_U0, bogo_V0 of the device states (column-major
 * real[ndim*ndim] each, state k at + k*ndim*ndim, rows in the sp order given to
 * tg_init).
 */
int tg_get_state(tg_ctx* ctx, double* U0, double* V0);

/*
 * tg_canonical_basis -- construct_canonical_basis (src/module_wavefunctions.f90:
 * 777-945) of every resident state on the device, in place of tg_get_state +
 * the host routine + tg_set_states between two tg_iterate calls: rho = V V^T is
 * diagonalised (ascending eigenvalues), nocc / nemp count |rhoc_ii| >= 1 - eps /
 * <= eps, ovacc = sum log u over the others, kappac = D^T V U^T D gives the
 * signs of v/u.  D replaces the resident D0 and the states are set up as
 * tg_set_states does (the gauge grid, Mphip, Mphin, ... of the last
 * tg_set_states).  Outputs per state (each pointer may be NULL): nocc[nstates],
 * nemp[nstates], ovacc[nstates], voveru real[nstates*ndim] (zero past
 * ndim - nocc).  eps: the reference's cut, 1e-8 by default (seed_occeps).
 * A degenerate subspace of kappac wider than one canonical pair (the
 * reference's real Schur step, dgees :877-900) is brought to canonical form by
 * an orthogonal transformation of its block (any canonical form gives the same
 * projected quantities).  Returns TG_E_UNSUPPORTED for a subspace wider than 64:
 * the states are then left unset and the caller takes the host route.
 */
int tg_canonical_basis(tg_ctx* ctx, double eps, int* nocc, int* nemp, double* ovacc,
                       double* voveru);

/*
 * tg_project_observables -- the gauge-angle loop of project_wavefunction(1, 0,
 * Mphip, Mphin, ndim) (src/module_projection.f90:236-356, the final printing
 * projection of print_results :990-1062) for the resident state (tg_set_state,
 * one state, batched mode): every angle of the grid, with the parity overlap.
 * Per angle, in the reference's angle order, a record of R = 19 + 2*nshell +
 * 2*nobo + 4*nobos complex values:
 *   [0..6]   rot_over, rot_pari, rot_prot, rot_neut, rot_prot2, rot_neut2, rot_ener
 *   [7..18]  rot_ecomp (calculate_decompo_energy, module_hamiltonian.f90:1134-1260)
 *   then     rot_occnum per shell s: (s, p), (s, n) (calculate_occupation_number,
 *            module_particlenumber.f90:157-176; sp_shell[i] = 0-based shell of
 *            state i, the first ndim/2 states being protons)
 *   then     for each real one-body operator k of obo_ops (nobo x ndim x ndim):
 *            the proton and neutron parts of calculate_expectval_obo_cplx
 *            (module_operators.f90:146-164), without the operator's scalar factor
 *   then     for each pair (O, O2) of obos_ops (nobos x 2 x ndim x ndim): the traces
 *            T1 = Tr(O rho), T2 = Tr(O2 rho), T5 = Tr(O rho O rho),
 *            T6 = Tr(O^T kappaRL^T O kappaLR) of calculate_expectval_obos_cplx
 *            (:227-257); for factors c1 O, c2 O2 the expectation value is
 *            c1^2 T1^2 + c2 T2 - c1^2 T5 + c1^2 T6.
 * pnp_rho (complex ndim x ndim, may be NULL): sum over angles of the
 * weight x overlap x rho_LR term of sum_gauge (:425-458).  The caller
 * accumulates the other pnp_* with the reference's sum_gauge.
 */
int tg_project_observables(tg_ctx* ctx, int nobo, const double* obo_ops, int nobos,
                           const double* obos_ops, int nshell, const int* sp_shell,
                           double* per_angle, double* pnp_rho);

/*
 * Groups: the projections of several contexts (each with its own basis size,
 * Hamiltonian and states) issued together as one hipGraph launch per call
 * (use_graph = 1: captured the second time a configuration is seen, replayed
 * afterwards; 0: the same launches, eagerly, each context on its stream).
 * tg_group_project is tg_project_resident for every context; tg_group_iterate
 * is tg_iterate for every context around one group projection, with the
 * outputs of context `index` read by tg_group_iterate_results (arrays as in
 * tg_iterate).  The contexts stay owned by the caller and must outlive the
 * group; a context is in at most one group.
 */
int tg_group_create(tg_group** group, int nctx, tg_ctx** ctxs, int use_graph);
int tg_group_destroy(tg_group* group);
int tg_group_project(tg_group* group);
int tg_group_iterate(tg_group* group, double gradient_eps);
int tg_group_iterate_results(tg_group* group, int index, double* pnp_scalars,
                             double* gradient_norm, double* lagrange_lambda1,
                             int* adjust_iterations, int* converged);

/*
 * Device timing: with timing enabled, every stage and the dominant kernels of
 * tg_project_resident record HIP event pairs on the stream they run on (no
 * synchronisation is added); tg_get_timing returns the summed milliseconds and
 * the number of intervals recorded under `key` ("total", "kernel:gj_inverse",
 * "kernel:pfaffian", "kernel:grouped_ph", "kernel:grouped_pp", "kernel:zgemm"
 * (the complex batched GEMMs of the loop body on the context stream),
 * "kernel:k_parts" (the reduced overlap's angle-independent parts of its skew
 * matrix, on the second stream), "contraction", "gemm", "gemm_h20", "overlap",
 * "thouless_inverse", and under tg_iterate_canonical "canonical_basis_beside"
 * (the canonical basis on the second stream)).
 */
int tg_enable_timing(tg_ctx* ctx, int on);
int tg_reset_timing(tg_ctx* ctx);
int tg_get_timing(tg_ctx* ctx, const char* key, double* ms_sum, int* count);

#ifdef __cplusplus
}
#endif

#endif /* TAURUS_GPU_H */
