// Internal declarations of the gfx950 PNVAP library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/taurus_gpu.h"

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void tg_set_error(const std::string& msg);

#define TG_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      tg_set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
      return TG_E_HIP;                                                            \
    }                                                                             \
  } while (0)

#define TG_TRY(call)                                                              \
  do {                                                                            \
    int r_ = (call);                                                              \
    if (r_ != TG_OK) return r_;                                                   \
  } while (0)

// ---------------------------------------------------------------------------
// Complex arithmetic on double2 (x = re, y = im)
// ---------------------------------------------------------------------------
typedef double2 cplx;

__host__ __device__ inline cplx cmk(double r, double i) { cplx z; z.x = r; z.y = i; return z; }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return cmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return cmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx cmul(cplx a, cplx b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline cplx cconj(cplx a) { return cmk(a.x, -a.y); }
__host__ __device__ inline cplx cscale(cplx a, double s) { return cmk(a.x * s, a.y * s); }
__host__ __device__ inline double cabs1(cplx a) { return fabs(a.x) + fabs(a.y); }   // IZAMAX measure
__host__ __device__ inline cplx cinv(cplx a) {
  // 1/a with scaling (Smith's algorithm)
  if (fabs(a.x) >= fabs(a.y)) {
    double r = a.y / a.x, d = a.x + a.y * r;
    return cmk(1.0 / d, -r / d);
  } else {
    double r = a.x / a.y, d = a.y + a.x * r;
    return cmk(r / d, -1.0 / d);
  }
}
// a += b*c  (four FMAs)
__host__ __device__ inline void cfma(cplx& a, cplx b, cplx c) {
  a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}
// a -= b*c  (four FMAs)
__host__ __device__ inline void cfms(cplx& a, cplx b, cplx c) {
  a.x = fma(-b.x, c.x, fma(b.y, c.y, a.x));
  a.y = fma(-b.x, c.y, fma(-b.y, c.x, a.y));
}

// ---------------------------------------------------------------------------
// Device buffer helper
// ---------------------------------------------------------------------------
// Count of device (re)allocations made through DevBuf::ensure.
unsigned long long devbuf_epoch();

constexpr int TG_MAX_CONS = 64;   // Lagrange multiplier slots of the gradient buffer

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t nbytes);
  void release();
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// GEMM (tg_gemm.hip):  C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b]
// column-major; op: 'N', 'T', 'C'; element types selected by the entry point.
// A[b] = A + (b / divA) * sA, B[b] = B + (b / divB) * sB, C[b] = C + b * sC:
// divA > 1 shares one A among divA consecutive batch elements (the per-state
// operands of a batch of states x gauge angles).
// ---------------------------------------------------------------------------
enum TgOp { OP_N = 0, OP_T = 1, OP_C = 2 };

// complex = complex x complex
// variant >= 0 forces a matrix-core tile variant (tg_mfma.h) instead of the tuned one
int gemm_zz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const cplx* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA = 1,
            int divB = 1, int variant = -1);
// complex = real x complex
int gemm_dz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const double* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA = 1,
            int divB = 1);
// complex = complex x real
int gemm_zd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const cplx* A, int lda, long long sA, const double* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA = 1,
            int divB = 1);
// real = real x real
int gemm_dd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, double alpha,
            const double* A, int lda, long long sA, const double* B, int ldb, long long sB,
            double beta, double* C, int ldc, long long sC, int batch, int divA = 1,
            int divB = 1);

// Grouped GEMM for the contraction blocks:
//   C_g (M x s_g, ld M) = X_g (M x s_g, ld M) * W_g (s_g x s_g, ld s_g)
// X_g / C_g start at column offset off_g of X / C; W_g at woff_g (float or double).
struct GroupDesc {
  int off;        // first column of the group in X/C
  int size;       // s_g
  long long woff; // offset of W_g
};
// `tag` names the channel (0 = ph, 1 = pp) for the autotuner.
int gemm_grouped(hipStream_t s, int M, const cplx* X, cplx* C, const void* W, int w_is_fp32,
                 const GroupDesc* groups_dev, const int* tile_group_dev, const int* tile_n0_dev,
                 int tag, int ntiles_total);

// Autotuner: time every matrix-core tile variant of one GEMM signature on the
// given operands (C is overwritten) and keep the fastest for later calls with
// that signature.  No effect with TG_GEMM=vector or a forced variant.
int gemm_tune_zz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const cplx* A, int lda,
                 long long sA, const cplx* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch);
int gemm_tune_dz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const double* A,
                 int lda, long long sA, const cplx* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch);
int gemm_tune_zd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const cplx* A, int lda,
                 long long sA, const double* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch);
int gemm_tune_grouped(hipStream_t s, int M, const cplx* X, cplx* C, const void* W, int w_is_fp32,
                      const GroupDesc* groups_dev, const int* tile_group_dev,
                      const int* tile_n0_dev, int ntiles_total, int tag);
// true when gemm_tune_grouped has nothing left to time for this signature
bool gemm_grouped_tuned(int M, int ntiles_total, int tag);
// tile variant the dispatcher uses for a signature (kind: 0 zz, 1 dz, 2 zd, 3 grouped)
int gemm_tuned_variant(int kind, TgOp opa, TgOp opb, int M, int N, int K, int batch);

// ---------------------------------------------------------------------------
// Linear algebra (tg_linalg.hip)
// ---------------------------------------------------------------------------
// In-place Gauss-Jordan inverse with partial pivoting (pivot = largest |re|+|im|,
// first index on ties, as IZAMAX in ZGETRF).  det[b] = determinant (complex),
// info[b] = 0 or the first zero-pivot column (1-based).  perm: int workspace n*batch.
int batched_inverse(hipStream_t s, int n, cplx* A, int lda, long long sA, int batch,
                    cplx* det, int* info, int* perm);
// Pfaffian of skew matrices (upper triangle read) by the Parlett-Reid LTL^T
// reduction with the pivoting of ZSKTF2 (UPLO='U', MODE='P').  A is destroyed:
// it and work (n*n per matrix) hold the leading block in turn.  Output:
// logabs[b] = log|pf|, phase[b] =
// pf/|pf| (complex), zero[b] = 1 when a pivot column vanished (pf = 0).
int batched_pfaffian(hipStream_t s, int n, cplx* A, int lda, long long sA, int batch,
                     cplx* work, double* logabs, cplx* phase, int* zero);

// ---------------------------------------------------------------------------
// Contraction plan (tg_fields.hip)
// ---------------------------------------------------------------------------
struct Channel {
  int npos = 0;                 // number of positions (n*n for ph, n(n-1)/2 for pp)
  int ngroups = 0;
  std::vector<GroupDesc> groups;
  std::vector<int> perm;        // position -> linear index (x + y*n) of the element
  // linear index -> position: >= 0 for a stored element; pp: -2 - position of
  // the transposed element for x < y (the field there is minus it), -1 on the
  // diagonal
  std::vector<int> inv;
  long long wsize = 0;          // sum of s_g^2
  DevBuf d_inv, d_groups, d_W, d_tile_group, d_tile_n0;
  int ntiles_n = 0, ntiles_total = 0;
};

// Configuration a captured projection graph is valid for.
struct GraphKey {
  long long v[20];
  bool operator==(const GraphKey& o) const {
    for (int i = 0; i < 20; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
};

// Canonical-basis data of one resident state (construct_canonical_basis).
struct StateInfo {
  int nocc = 0, nemp = 0;
  double ovac = 0;
  std::vector<double> voveru;                     // n - nocc values
};

// Per-state strides of the device layouts (S states of one context):
//   d_U, d_V, d_D: state s at s * n^2;  d_real6: product k of state s at (k * S + s) * n^2
//   d_out: state s at s * tg_out_stride(n) complex: pnp_A | pnp_H20 | pnp_AH20 | field_H20v
//   d_Z: Zi [S n^2] | Zim1 [S n^2] | lambda [S * TG_MAX_CONS] | partial sums [S * tg_part_len(n)]
inline long long tg_out_stride(int n) {
  const long long nn = (long long)n * n;
  return 3 * nn + (nn + 1) / 2;
}
inline long long tg_part_len(int n) { return (long long)n * n / 256 + 2; }

struct tg_ctx {
  int n = 0;
  int flags = 0;
  int plain = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;                  // overlaps, concurrent with the Thouless chain
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipStream_t stream3 = nullptr;                  // the pp contraction beside the ph one
  hipEvent_t ev_fork3 = nullptr, ev_join3 = nullptr;
  // basis
  std::vector<int> sp_2mj, sp_2mt, sp_l, sp_tr;   // internal order; sp_tr 0-based here
  std::vector<int> sp_perm, sp_iperm;             // internal -> reference index, and back
  cplx H0;
  DevBuf d_H1;                                    // complex n*n
  // contraction channels
  Channel ph, pp;
  int w_fp32 = 0;
  // resident states: ns independent states of this basis, projected together
  int ns = 1;
  int A = 0;                                      // number of angles of each state
  int Mphip = 0, Mphin = 0;
  double facpi = 0, valZ = 0, valN = 0, factor_delta = 1.0;
  std::vector<StateInfo> st;                      // canonical-basis data per state
  std::vector<double> phip, phin;
  std::vector<cplx> weight;                       // weip * wein
  std::vector<double> mult;                       // 2 for a conjugate pair, else 1
  int conj_pairs = 0;                             // TG_FLAG_CONJ_PAIRS
  // device work areas
  DevBuf d_U, d_V, d_D, d_ang;                    // real n*n, angles
  DevBuf d_real6;                                 // Kp Kn Lp Ln Mp Mn (6 n*n real)
  DevBuf d_work;                                  // per-angle complex matrices
  DevBuf d_blk;                                   // contraction layouts
  DevBuf d_scal;                                  // per-angle scalars
  DevBuf d_out;                                   // pnp accumulators
  DevBuf d_ovl;                                   // overlap work
  DevBuf d_int;                                   // int work
  DevBuf d_pfw;                                   // pfaffian work
  int have_state = 0;
  int evolved = 0;   // the states are the last iteration's, their canonical basis not set
  int tuned_n = 0, tuned_chunk = 0;                // sizes the GEMM tiles were tuned for
  // items per launch sequence: batch_limit from tg_set_batch_limit (0: none),
  // mem_chunk chosen from the free device memory by the upload (0: all items)
  int batch_limit = 0, mem_chunk = 0;
  int state_dirty = 0;                            // set-up of a new state not uploaded yet
  // hipGraph replay of the projection (TG_FLAG_GRAPH)
  int use_graph = 0, graph_valid = 0, graph_seen = 0;
  GraphKey graph_key_v{}, graph_seen_key{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int iprint = 0;                                 // parity overlaps (rot_pari) wanted
  DevBuf d_psign;                                 // (-1)^l per single-particle state
  // on-device gradient step (calculate_gradient)
  int grad_on = 0;
  double grad_eta = 0, grad_mu = 0;
  // gradient_type of the device iteration: 0 = eta, mu as given; 1 / 2 = the
  // heavy-ball recipes from the H11 / h spectra of diagonalize_hsp_and_H11
  int grad_type = 0;
  std::vector<double> etamu;                      // eta, mu per state of the last step
  std::vector<double> eig_last;                   // S x (h, H11) spectra of the last step
  DevBuf d_etamu;
  int grad_ncons = 0;
  std::vector<double> grad_lambda;
  DevBuf d_c20, d_Z;                              // constraint_20; Zi, Zim1, lambda, norm
  DevBuf d_cq;                                    // products of the constraint_20 of one state
  int c20_shared = 1;                             // constraint_20 / lambda common to all states
  int z_states = 0;                               // states the Z buffers hold
  // constraints of the device iteration (tg_set_constraints / tg_iterate)
  int cons_dim = 0, cons_pair = 0;
  double cons_eps = 0;
  std::vector<double> cons_val;                   // cons_nsets x cons_dim
  int cons_nsets = 1;
  DevBuf d_cHO;                                   // constraint operators, internal order
  DevBuf d_it;                                    // iteration work (n*n real matrices)
  DevBuf d_obs;                                   // observables of the final projection
  DevBuf d_a19;                                   // h, H11 and their eigenvalues
  std::vector<double> lambda_prev;                // multipliers of the previous iteration (ns x dim)
  // canonical basis on the device (tg_canon.hip): work buffers
  DevBuf d_canon;
  // tg_iterate_canonical: the projection runs defer_hook(0) at its fork (the
  // canonical basis's device work enqueued on the second stream), enqueues its
  // Thouless / field / H20 chain, runs defer_hook(1) on the host (the rest of the
  // canonical basis and the overlap set-up of the new basis), then enqueues the
  // overlap chain
  int defer_overlap = 0;
  std::function<int(int)> defer_hook;
  // timing: event pairs recorded without synchronisation, summed at collection
  struct Pending { std::string key; hipEvent_t e0, e1; };
  int timing = 0;
  std::vector<Pending> pending;
  std::map<std::string, double> tsum;
  std::map<std::string, int> tcount;
};

// Records a [begin, end) event pair on the context stream under `key` when
// timing is enabled; the elapsed time is added at the next collection.
struct ScopedTiming {
  tg_ctx* c;
  hipEvent_t e0 = nullptr;
  const char* key;
  hipStream_t s;
  ScopedTiming(tg_ctx* c_, const char* k, hipStream_t st = nullptr);   // default: c->stream
  ~ScopedTiming();
};
int timing_collect(tg_ctx* c);

// Projection of the resident states (tg_project.hip) and download of the sums
// of state s.
int tg_project_device(tg_ctx* c);
// Allocations and set-up uploads of the resident states before a (captured)
// projection; project_run launches only.  Both for the group capture.
int tg_project_upload(tg_ctx* c);
int tg_project_run(tg_ctx* c);
// gradient_kernel for every state (calculate_gradient, gradient_type 0) with
// per-state constraint_20 (c20 + s * ncons * n^2) and lambda (lam + s * TG_MAX_CONS)
int tg_gradient_step(tg_ctx* c, const double* c20, const double* lam);
int tg_fetch_results(tg_ctx* c, int state, double* scal, double* pA, double* pH, double* pAH,
                     double* H20v, double* imag);
// ensures the Z buffers for c->ns states (zeroed when they are (re)sized)
int tg_ensure_z(tg_ctx* c);
// diagonalize_hsp_and_H11(0) of every resident state (tg_a19.hip): spectra of
// h and H11 (with the constraint terms of the device multipliers lam,
// S x TG_MAX_CONS) to eig (S x 2 x n); hmat / h11 receive the device matrices.
int tg_hsp_h11_device(tg_ctx* c, const double* lam, std::vector<double>& eig, double** hmat,
                      double** h11);
// eta, mu of calculate_gradient for gradient_type (module_gradient.f90:97-120)
void gradient_recipe(int type, int n, const double* eig_hsp, const double* eig_h11, double eta0,
                     double mu0, double* eta, double* mu);
// eigenvalues (ascending) of count symmetric matrices A + b sA (upper triangle
// read) to w + b ws; work: count n^2 doubles for orders above 112
// Eigenvalues (ascending, stride ws) and eigenvectors (n x n, stride sE) of count
// symmetric matrices (upper triangles read), n <= 1024; work of
// sym_eigensystem_work_bytes(n, count) bytes; info[k] != 0 when a secular root of
// matrix k did not converge (tg_eig.hip).
size_t sym_eigensystem_work_bytes(int n, int count);
int batched_sym_eigensystem(hipStream_t s, int n, const double* A, int lda, long long sA, int count,
                            void* work, double* w, long long ws, double* E, long long sE, int* info);
// tg_canonical_basis on a given stream (tg_canon.hip), also in two parts: the
// device work up to the first host read (canonical_launch, no host wait) and
// the rest (canonical_finish, same stream); the overlap set-up of the resident
// states' canonical data uploaded on stream s (tg_project.hip)
int canonical_basis_on(tg_ctx* c, hipStream_t s, double eps, int* nocc, int* nemp, double* ovacc,
                       double* voveru);
int canonical_launch(tg_ctx* c, hipStream_t s, double eps);
int canonical_finish(tg_ctx* c, hipStream_t s, double eps, int* nocc, int* nemp, double* ovacc,
                     double* voveru);
int upload_overlap_setup(tg_ctx* c, hipStream_t s);
int batched_sym_eigvals(hipStream_t s, int n, const double* A, int lda, long long sA, int count,
                        double* work, double* w, long long ws);
// The resident U, V, D and c->st of c->ns states are set: the tail of
// tg_set_states (Z buffers, gauge grid, state marked for the set-up upload).
int tg_states_ready(tg_ctx* c);
// releases the rocBLAS handle of the device canonical basis

int contraction_build(tg_ctx* c, int64_t ntb, const int16_t* abcd, const void* H2,
                      int h2_bytes, const int8_t* trperm);
// Gamma, DeltaLR, DeltaRL for `na` angles whose densities are stored per angle
// (complex n*n, stride n*n).  Writes gamma (without H1), deltaLR, deltaRL.
// hermitian != 0 reproduces calculate_fields_diag (Gamma made hermitian from its
// upper triangle, DeltaRL = conj(DeltaLR)).  With hsp != nullptr it also writes
// hsp = gamma + H1 (H1 one complex n*n matrix shared by the angles).
// The pp channel runs on the third stream.  defer_join (two pp right-hand sides,
// no hermitian completion, hsp given): gamma and hsp are complete on the context
// stream on return, deltaLR and deltaRL only after contraction_join.
int contraction_run(tg_ctx* c, int na, const cplx* rho, const cplx* kLR, const cplx* kRL,
                    cplx* gamma, cplx* dLR, cplx* dRL, int hermitian,
                    const cplx* H1 = nullptr, cplx* hsp = nullptr, bool defer_join = false);
int contraction_join(tg_ctx* c);
