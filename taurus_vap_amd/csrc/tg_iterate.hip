// One gradient iteration of the reference driver on the device
// (taurus_vap.f90:77-113), for every resident state of a context at once:
//   update_constraints_qpbasis          module_constraints.f90:381-397
//   project_wavefunction(0, ...)        (tg_project.hip)
//   calculate_lagrange_multipliers      module_constraints.f90:413-484
//   diagonalize_hsp_and_H11(0)          module_gradient.f90:158-453 (gradient_type
//                                       1, 2; tg_a19.hip)
//   calculate_gradient                  module_gradient.f90:79-143
//   adjust_constraints                  module_constraints.f90:270-373
//     cholesky_factorization            module_constraints.f90:497-543
//     evolve_wavefunction               module_constraints.f90:559-590
// The constraint operators and the Bogoliubov matrices stay in HBM; the host
// sees only the handful of constraint numbers that decide the adjustment loop
// and its dim x dim linear systems.  The states of a context run the
// adjustment in lockstep: a state whose constraints are met keeps its Z (the
// next evolutions reproduce the same U1, V1), so each state ends where its own
// loop would have.  The canonical basis of the new states is left to the
// caller (tg_get_state, then tg_set_state / tg_set_states).
//
// A group of contexts (tg_group_*) projects all its contexts, of any basis
// sizes, as one hipGraph launch per iteration.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tg_internal.h"

namespace {

constexpr int NT = 256;
constexpr int MAXPAIRS = TG_MAX_CONS * (TG_MAX_CONS + 1) / 2;

inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

// C = (A - A^T) / 2 with A = T1 - T2  (set_operator_qpbasis 'f20'/'g20', factor -1);
// state blockIdx.z: T1, T2 at + s * nn, C at + s * cs
__global__ void __launch_bounds__(256)
skew_half_kernel(int n, const double* __restrict__ T10, const double* __restrict__ T20,
                 double* __restrict__ C0, long long cs) {
  __shared__ double ta[32][33], tb[32][33];
  const long long nn = (long long)n * n;
  const double* T1 = T10 + blockIdx.z * nn;
  const double* T2 = T20 + blockIdx.z * nn;
  double* C = C0 + blockIdx.z * cs;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    ta[r][tx] = (i < n && j < n) ? T1[i + (long long)j * n] - T2[i + (long long)j * n] : 0.0;
    const int i2 = j0 + tx, j2 = i0 + r;
    tb[r][tx] = (i2 < n && j2 < n) ? T1[i2 + (long long)j2 * n] - T2[i2 + (long long)j2 * n] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < n && j < n) C[i + (long long)j * n] = 0.5 * (ta[r][tx] - tb[tx][r]);
  }
}

// out[p * gridDim.y + q] = sum_k X_p[k] Y_p[k] over slice q of the gridDim.y slices
// of [0, len), for the pointer pairs of the list (block (p, q))
__global__ void __launch_bounds__(NT)
dots_kernel(long long len, const double* const* __restrict__ xs, const double* const* __restrict__ ys,
            double* __restrict__ out) {
  const double* x = xs[blockIdx.x];
  const double* y = ys[blockIdx.x];
  const long long chunk = (len + gridDim.y - 1) / gridDim.y;
  const long long k0 = chunk * blockIdx.y, k1 = min(len, k0 + chunk);
  double s = 0.0, s1 = 0.0;
  long long k = k0 + threadIdx.x;
  for (; k + NT < k1; k += 2 * NT) {
    s = fma(x[k], y[k], s);
    s1 = fma(x[k + NT], y[k + NT], s1);
  }
  if (k < k1) s = fma(x[k], y[k], s);
  s += s1;
  __shared__ double red[NT];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x * gridDim.y + blockIdx.y] = red[0];
}

// S += 1 for the matrix of state blockIdx.y (stride nn)
__global__ void add_identity_kernel(int n, double* __restrict__ S0) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double* S = S0 + blockIdx.y * (long long)n * n;
  if (i < n) S[i + (long long)i * n] += 1.0;
}

__global__ void copy_kernel(long long len, const double* __restrict__ src, double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < len) dst[i] = src[i];
}

// dst = src for the states s (blockIdx.y, stride nn) with mask[s] != 0
__global__ void masked_copy_kernel(long long nn, const int* __restrict__ mask,
                                   const double* __restrict__ src, double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  const long long o = blockIdx.y * nn;
  if (i < nn && mask[blockIdx.y]) dst[o + i] = src[o + i];
}

// Z_s += sum_m x_s[m] C20_s,m   (state blockIdx.y: Z at + s nn, x at + s TG_MAX_CONS,
// C20 at + s ncons nn)
__global__ void axpy_cons_kernel(long long nn, int ncons, const double* __restrict__ x0,
                                 const double* __restrict__ c200, double* __restrict__ Z0) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= nn) return;
  const long long st = blockIdx.y;
  const double* x = x0 + st * TG_MAX_CONS;
  const double* c20 = c200 + st * ncons * nn;
  double* Z = Z0 + st * nn;
  double s = Z[i];
  for (int m = 0; m < ncons; ++m) s = fma(x[m], c20[(long long)m * nn + i], s);
  Z[i] = s;
}

__global__ void fill_kernel(long long len, double v, double* __restrict__ x) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < len) x[i] = v;
}

__global__ void fill_int_kernel(int len, int v, int* __restrict__ x) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < len) x[i] = v;
}

// Diagonal block of the blocked Cholesky factorisation (one workgroup per
// state, blockIdx.x; matrices at + s n^2): factors the b x b block A(k0:, k0:)
// (lower triangle read, b <= CB) in LDS, writes it to L (zeros above the
// diagonal) and its inverse to X, which is the diagonal block of L^-1.  A
// non-positive pivot is replaced by 1 and its column (1-based) recorded in
// info[s] if none was before.
// One thread per element (row ti, column tj).  Step k takes column k of the
// factor: each thread scales the two column-k entries its update needs, so
// column k of `a` is only read in that step and the factor column goes to lf.
// One step behind, the forward substitution L X = I takes row k-1 of X (final
// as x(k-1, :) / L(k-1, k-1)) off the rows below.  One barrier per step.
constexpr int CB = 32;
static_assert(CB == 32, "chol_block_kernel maps its 1024 threads onto 32 rows x 32 columns");
__global__ void __launch_bounds__(CB * CB)
chol_block_kernel(int n, int k0, int b, const double* __restrict__ A0, double* __restrict__ L0,
                  double* __restrict__ X0, int* __restrict__ info) {
  __shared__ double a[CB][CB + 1], lf[CB][CB + 1], x[CB][CB + 1];
  const long long nn = (long long)n * n;
  const double* A = A0 + blockIdx.x * nn;
  double* L = L0 + blockIdx.x * nn;
  double* X = X0 + blockIdx.x * nn;
  const int ti = threadIdx.x & (CB - 1), tj = threadIdx.x >> 5;
  const bool in = ti < b && tj < b;
  a[ti][tj] = (in && ti >= tj) ? A[(k0 + ti) + (long long)(k0 + tj) * n] : 0.0;
  lf[ti][tj] = 0.0;
  x[ti][tj] = (in && ti == tj) ? 1.0 : 0.0;
  __syncthreads();
  int bad = 0;
  for (int k = 0; k <= b; ++k) {
    if (k < b) {
      double d = a[k][k];
      if (!(d > 0.0)) {
        if (bad == 0) bad = k0 + k + 1;
        d = 1.0;
      }
      d = sqrt(d);
      if (tj > k && ti >= tj && ti < b) {
        const double li = a[ti][k] / d, lj = a[tj][k] / d;
        a[ti][tj] -= li * lj;
      }
      if (tj == k && ti >= k && ti < b) lf[ti][k] = (ti == k) ? d : a[ti][k] / d;
    }
    if (k > 0) {
      const int q = k - 1;
      if (tj <= q && ti > q && ti < b) x[ti][tj] -= lf[ti][q] * (x[q][tj] / lf[q][q]);
    }
    __syncthreads();
  }
  if (in) {
    const long long o = (k0 + ti) + (long long)(k0 + tj) * n;
    L[o] = (ti >= tj) ? lf[ti][tj] : 0.0;
    X[o] = (ti >= tj) ? x[ti][tj] / lf[ti][ti] : 0.0;
  }
  if (threadIdx.x == 0 && bad && info[blockIdx.x] == 0) info[blockIdx.x] = bad;
}

// Gaussian elimination with partial pivoting of a small dense system (host);
// false when singular.
bool small_solve(int m, std::vector<double> A, std::vector<double>& b) {
  for (int k = 0; k < m; ++k) {
    int p = k;
    for (int i = k + 1; i < m; ++i)
      if (fabs(A[i + k * m]) > fabs(A[p + k * m])) p = i;
    if (!(fabs(A[p + k * m]) > 0.0)) return false;
    if (p != k) {
      for (int j = 0; j < m; ++j) std::swap(A[k + j * m], A[p + j * m]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < m; ++i) {
      const double f = A[i + k * m] / A[k + k * m];
      for (int j = k; j < m; ++j) A[i + j * m] -= f * A[k + j * m];
      b[i] -= f * b[k];
    }
  }
  for (int k = m - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < m; ++j) s -= A[k + j * m] * b[j];
    b[k] = s / A[k + k * m];
  }
  return true;
}

// Work layout: IT_COUNT slots of S n^2 doubles (matrix of state s at + s n^2),
// then a tail of
//   [0, 3 * S * MAXPAIRS)          dot-product pointer lists and results
//   [TX, TX + S * TG_MAX_CONS)     constraint corrections x per state
//   [TI, ...)                      Cholesky info and copy masks (ints, S each)
//   [T2, T2 + 3 S MAXPAIRS)        a second dot-product area (two batches in flight)
enum { IT_S, IT_LINV, IT_A1, IT_A2, IT_U1, IT_V1, IT_R, IT_T1, IT_T2, IT_X, IT_COUNT };

struct Iter {
  tg_ctx* c;
  int n, S;
  long long nn;
  hipStream_t s;
  double* w;
  double* M(int k) const { return w + (size_t)k * S * nn; }
  double* tail() const { return w + (size_t)IT_COUNT * S * nn; }
  size_t tx() const { return 3 * (size_t)S * MAXPAIRS; }
  size_t ti() const { return tx() + (size_t)S * TG_MAX_CONS; }
  size_t t2() const { return ti() + (size_t)S + 64; }
  static size_t bytes(int n, int S) {
    return sizeof(double) * ((size_t)IT_COUNT * S * n * n + 6 * (size_t)S * MAXPAIRS +
                             (size_t)S * TG_MAX_CONS + (size_t)S + 64);
  }
  double* x() const { return tail() + tx(); }
  int* info() const { return reinterpret_cast<int*>(tail() + ti()); }
  int* mask() const { return info() + S; }
  double* tail2() const { return tail() + t2(); }   // a second dot-product area
};

// constraint_20 of every constraint at (U, V) of every state into c20
// (state s, constraint m at c20 + (s * dim + m) n^2)
int constraints_qp(const Iter& it, const double* U, const double* V, double* c20) {
  tg_ctx* c = it.c;
  const int n = it.n, S = it.S, dim = c->cons_dim;
  const long long nn = it.nn;
  hipStream_t s = it.s;
  if (dim == 0) return TG_OK;
  const double* O = c->d_cHO.as<double>();
  const int npair = c->cons_pair, nobo = dim - npair;
  if (S > dim) {
    // per constraint, every state in one launch
    double* X = it.M(IT_X);
    double* T1 = it.M(IT_T1);
    double* T2 = it.M(IT_T2);
    for (int m = 0; m < dim; ++m) {
      const double* Om = O + (size_t)m * nn;
      const bool pair = m >= nobo;
      // f20: T1 = U^T O V, T2 = V^T O^T U;  g20: T1 = U^T O U, T2 = V^T O V
      TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, Om, n, 0, pair ? U : V, n, nn, 0.0, X, n, nn, S));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, U, n, nn, X, n, nn, 0.0, T1, n, nn, S));
      TG_TRY(gemm_dd(s, pair ? OP_N : OP_T, OP_N, n, n, n, 1.0, Om, n, 0, pair ? V : U, n, nn, 0.0, X,
                     n, nn, S));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, V, n, nn, X, n, nn, 0.0, T2, n, nn, S));
      hipLaunchKernelGGL(skew_half_kernel, dim3((n + 31) / 32, (n + 31) / 32, S), dim3(256), 0, s,
                         n, T1, T2, c20 + (size_t)m * nn, (long long)dim * nn);
      TG_HIP(hipGetLastError());
    }
    return TG_OK;
  }
  // per state, the one-body and the pair constraints each in one launch
  TG_TRY(c->d_cq.ensure(sizeof(double) * 3 * (size_t)dim * nn));
  double* X = c->d_cq.as<double>();
  double* T1 = X + (size_t)dim * nn;
  double* T2 = T1 + (size_t)dim * nn;
  for (int st = 0; st < S; ++st) {
    const double* Us = U + (size_t)st * nn;
    const double* Vs = V + (size_t)st * nn;
    if (nobo > 0) {
      TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, O, n, nn, Vs, n, 0, 0.0, X, n, nn, nobo));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Us, n, 0, X, n, nn, 0.0, T1, n, nn, nobo));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, O, n, nn, Us, n, 0, 0.0, X, n, nn, nobo));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Vs, n, 0, X, n, nn, 0.0, T2, n, nn, nobo));
    }
    if (npair > 0) {
      const size_t o = (size_t)nobo * nn;
      TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, O + o, n, nn, Us, n, 0, 0.0, X + o, n, nn, npair));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Us, n, 0, X + o, n, nn, 0.0, T1 + o, n, nn, npair));
      TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, O + o, n, nn, Vs, n, 0, 0.0, X + o, n, nn, npair));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Vs, n, 0, X + o, n, nn, 0.0, T2 + o, n, nn, npair));
    }
    hipLaunchKernelGGL(skew_half_kernel, dim3((n + 31) / 32, (n + 31) / 32, dim), dim3(256), 0, s,
                       n, T1, T2, c20 + (size_t)st * dim * nn, nn);
    TG_HIP(hipGetLastError());
  }
  return TG_OK;
}

// A batch of dot products sum_k X_p[k] Y_p[k] in flight on the iteration's stream:
// its pointer lists (uploaded from h) and slice sums (read back into part) live in
// a device area of 3 S MAXPAIRS doubles; the sums are formed on the host after the
// caller has synchronised the stream (dots_finish).
struct DotsJob {
  int np = 0, ds = 1;
  std::vector<const double*> h;
  std::vector<double> part;
};

int dots_launch(const Iter& it, double* area, const std::vector<const double*>& xs,
                const std::vector<const double*>& ys, DotsJob& job) {
  const int np = (int)xs.size();
  job.np = np;
  if (np == 0) return TG_OK;
  if (np > it.S * MAXPAIRS) {
    tg_set_error("constraint dot products: too many pairs");
    return TG_E_ARG;
  }
  const double** pl = reinterpret_cast<const double**>(area);
  double* res = reinterpret_cast<double*>(pl + 2 * np);
  job.h.assign(2 * (size_t)np, nullptr);
  for (int p = 0; p < np; ++p) {
    job.h[p] = xs[p];
    job.h[np + p] = ys[p];
  }
  // each dot in up to 32 slices (one block each) as the result area allows;
  // the slices are summed on the host in a fixed order
  const long long room = 3LL * it.S * MAXPAIRS - 2LL * np;
  job.ds = (int)std::max(1LL, std::min(32LL, room / np));
  job.part.assign((size_t)np * job.ds, 0.0);
  TG_HIP(hipMemcpyAsync(pl, job.h.data(), sizeof(double*) * 2 * np, hipMemcpyHostToDevice, it.s));
  hipLaunchKernelGGL(dots_kernel, dim3(np, job.ds), dim3(NT), 0, it.s, it.nn, pl, pl + np, res);
  TG_HIP(hipGetLastError());
  TG_HIP(hipMemcpyAsync(job.part.data(), res, sizeof(double) * np * job.ds, hipMemcpyDeviceToHost,
                        it.s));
  return TG_OK;
}

void dots_finish(const DotsJob& job, std::vector<double>& out) {
  out.assign(job.np, 0.0);
  for (int p = 0; p < job.np; ++p) {
    double acc = 0.0;
    for (int q = 0; q < job.ds; ++q) acc += job.part[(size_t)p * job.ds + q];
    out[p] = acc;
  }
}

// host copies of sum_k X_p[k] Y_p[k] for the pairs (xs[p], ys[p])
int dots(const Iter& it, const std::vector<const double*>& xs, const std::vector<const double*>& ys,
         std::vector<double>& out) {
  DotsJob job;
  TG_TRY(dots_launch(it, it.tail(), xs, ys, job));
  TG_HIP(hipStreamSynchronize(it.s));
  dots_finish(job, out);
  return TG_OK;
}

// Gram matrices A_s,ij = Tr(C20_s,i^T C20_s,j) of the listed states (dim x dim each):
// the dot products, then (after a synchronisation) the matrices
int gram_launch(const Iter& it, double* area, const double* c20, const std::vector<int>& states,
                DotsJob& job) {
  const int m = it.c->cons_dim;
  std::vector<const double*> xs, ys;
  for (int s : states)
    for (int i = 0; i < m; ++i)
      for (int j = i; j < m; ++j) {
        xs.push_back(c20 + ((size_t)s * m + i) * it.nn);
        ys.push_back(c20 + ((size_t)s * m + j) * it.nn);
      }
  return dots_launch(it, area, xs, ys, job);
}

void gram_finish(const Iter& it, const DotsJob& job, size_t nstates,
                 std::vector<std::vector<double>>& A) {
  const int m = it.c->cons_dim;
  std::vector<double> d;
  dots_finish(job, d);
  A.assign(nstates, std::vector<double>((size_t)m * m, 0.0));
  int k = 0;
  for (size_t q = 0; q < nstates; ++q)
    for (int i = 0; i < m; ++i)
      for (int j = i; j < m; ++j, ++k) A[q][i + j * m] = A[q][j + i * m] = d[k];
}

int gram(const Iter& it, const double* c20, const std::vector<int>& states,
         std::vector<std::vector<double>>& A) {
  DotsJob job;
  TG_TRY(gram_launch(it, it.tail(), c20, states, job));
  TG_HIP(hipStreamSynchronize(it.s));
  gram_finish(it, job, states.size(), A);
  return TG_OK;
}

// cholesky_factorization + dtrtri of every state: A (n x n per state, destroyed)
// = L L^T, X = L^-1 (dense, zero above the diagonal).  Right-looking blocked
// Cholesky with CB-wide panels (diagonal block in LDS, panel and trailing
// update as GEMMs), then the off-diagonal blocks of L^-1 row block by row
// block: X(i, :i) = -X(i, i) L(i, :i) X(:i, :i).  L and T are scratch (T:
// CB x n per state).  info[s] = 0 or the first non-positive pivot column.
// The block rows of L^-1 go to stream s2 (T is its scratch), each after the
// diagonal block it needs (event ev), beside the Cholesky's later steps on the
// context stream; s2 is not joined here.
int cholesky_inverse(const Iter& it, double* A, double* L, double* X, double* T, int* info,
                     hipStream_t s2, hipEvent_t ev) {
  const int n = it.n, S = it.S;
  const long long nn = it.nn;
  hipStream_t s = it.s;
  hipLaunchKernelGGL(fill_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s, S * nn, 0.0, L);
  hipLaunchKernelGGL(fill_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s, S * nn, 0.0, X);
  hipLaunchKernelGGL(fill_int_kernel, dim3(nblk(S)), dim3(NT), 0, s, S, 0, info);
  TG_HIP(hipGetLastError());
  for (int k0 = 0; k0 < n; k0 += CB) {
    const int b = std::min(CB, n - k0);
    hipLaunchKernelGGL(chol_block_kernel, dim3(S), dim3(CB * CB), 0, s, n, k0, b, A, L, X, info);
    TG_HIP(hipGetLastError());
    if (k0 > 0) {
      // X(k0:k0+b, :k0) = -X(k0, k0) L(k0:k0+b, :k0) X(:k0, :k0): L's rows of this block
      // are final (earlier panels), X(:k0, :k0) is the diagonal blocks and the earlier
      // block rows (this stream)
      TG_HIP(hipEventRecord(ev, s));
      TG_HIP(hipStreamWaitEvent(s2, ev, 0));
      TG_TRY(gemm_dd(s2, OP_N, OP_N, b, k0, k0, 1.0, L + k0, n, nn, X, n, nn, 0.0, T, CB, nn, S));
      TG_TRY(gemm_dd(s2, OP_N, OP_N, b, k0, b, -1.0, X + k0 + (long long)k0 * n, n, nn, T, CB, nn,
                     0.0, X + k0, n, nn, S));
    }
    const int m = n - k0 - b;
    if (m == 0) break;
    const long long pk = k0 + (long long)k0 * n, pb = (k0 + b) + (long long)k0 * n;
    // panel L(k0+b:, k0:k0+b) = A(k0+b:, k0:k0+b) L(k0:k0+b, k0:k0+b)^-T
    TG_TRY(gemm_dd(s, OP_N, OP_T, m, b, b, 1.0, A + pb, n, nn, X + pk, n, nn, 0.0, L + pb, n, nn, S));
    // trailing update A(k0+b:, k0+b:) -= L_panel L_panel^T
    TG_TRY(gemm_dd(s, OP_N, OP_T, m, m, b, -1.0, L + pb, n, nn, L + pb, n, nn, 1.0,
                   A + (k0 + b) + (long long)(k0 + b) * n, n, nn, S));
  }
  return TG_OK;
}

// Outputs of the last iteration of a context (per state)
struct IterOut {
  std::vector<double> scal, gnorm, lambda;
  std::vector<int> nadj, conv;
};

// ---- the iteration in two halves around the projection
int iterate_pre(tg_ctx* c) {
  if (!c->have_state || !c->grad_on) {
    tg_set_error("tg_iterate: tg_set_state and tg_set_gradient must be called first");
    return TG_E_ARG;
  }
  if (c->cons_nsets != 1 && c->cons_nsets != c->ns) {
    tg_set_error("tg_iterate: constraint values given for " + std::to_string(c->cons_nsets) +
                 " states, the context holds " + std::to_string(c->ns));
    return TG_E_ARG;
  }
  const int n = c->n, S = c->ns, dim = c->cons_dim;
  const long long nn = (long long)n * n;
  TG_TRY(tg_ensure_z(c));
  TG_TRY(c->d_it.ensure(Iter::bytes(n, S)));
  TG_TRY(c->d_c20.ensure(sizeof(double) * std::max<size_t>(1, (size_t)S * dim * nn)));
  Iter it{c, n, S, nn, c->stream, c->d_it.as<double>()};
  // update_constraints_qpbasis
  TG_TRY(constraints_qp(it, c->d_U.as<double>(), c->d_V.as<double>(), c->d_c20.as<double>()));
  c->grad_ncons = dim;
  c->c20_shared = 0;
  return TG_OK;
}

int iterate_post(tg_ctx* c, double gradient_eps, IterOut& out) {
  const int n = c->n, S = c->ns, dim = c->cons_dim;
  const long long nn = (long long)n * n;
  hipStream_t s = c->stream;
  Iter it{c, n, S, nn, s, c->d_it.as<double>()};
  double* U0 = c->d_U.as<double>();
  double* V0 = c->d_V.as<double>();
  double* c20 = c->d_c20.as<double>();
  double* Zb = c->d_Z.as<double>();
  double* Z = Zb;                                  // gradient_Zi, S states
  double* Zm1 = Zb + S * nn;                       // gradient_Zim1
  double* lam = Zb + 2 * S * nn;                   // S x TG_MAX_CONS
  double* part = lam + (size_t)S * TG_MAX_CONS;    // S x tg_part_len(n)
  const long long ps = tg_part_len(n);
  const long long outs = tg_out_stride(n);
  const cplx* out0 = c->d_out.as<cplx>();
  auto H20v = [&](int st) {
    return reinterpret_cast<const double*>(out0 + (size_t)st * outs + 3 * nn);
  };
  auto cval = [&](int st, int m) { return c->cons_val[(c->cons_nsets == 1 ? 0 : st) * dim + m]; };

  // projected scalars, and the reference's stop on a complex gradient
  // (calculate_projected_H20, module_projection.f90:751-755)
  out.scal.assign((size_t)S * 14, 0.0);
  for (int st = 0; st < S; ++st)
    TG_TRY(tg_fetch_results(c, st, out.scal.data() + 14 * st, nullptr, nullptr, nullptr, nullptr,
                            nullptr));
  // calculate_lagrange_multipliers: A_ij = Tr(C_i C_j^T), b_i = Tr(H20 C_i^T)
  out.lambda.assign((size_t)S * dim, 0.0);
  if (c->lambda_prev.size() != (size_t)S * dim) c->lambda_prev.assign((size_t)S * dim, 0.0);
  std::vector<double> lamh((size_t)S * TG_MAX_CONS, 0.0);
  if (dim > 0) {
    std::vector<int> all(S);
    for (int st = 0; st < S; ++st) all[st] = st;
    std::vector<std::vector<double>> A;
    TG_TRY(gram(it, c20, all, A));
    std::vector<const double*> xs, ys;
    for (int st = 0; st < S; ++st)
      for (int i = 0; i < dim; ++i) {
        xs.push_back(H20v(st));
        ys.push_back(c20 + ((size_t)st * dim + i) * nn);
      }
    std::vector<double> b;
    TG_TRY(dots(it, xs, ys, b));
    for (int st = 0; st < S; ++st) {
      std::vector<double> bs(b.begin() + st * dim, b.begin() + (st + 1) * dim);
      double* prev = c->lambda_prev.data() + (size_t)st * dim;
      // the reference keeps the multipliers of the previous iteration when the
      // system is singular (module_constraints.f90:471-477)
      if (small_solve(dim, A[st], bs)) std::copy(bs.begin(), bs.end(), prev);
      for (int i = 0; i < dim; ++i) {
        lamh[(size_t)st * TG_MAX_CONS + i] = prev[i];
        out.lambda[(size_t)st * dim + i] = prev[i];
      }
    }
    TG_HIP(hipMemcpyAsync(lam, lamh.data(), sizeof(double) * lamh.size(), hipMemcpyHostToDevice, s));
  }
  // gradient_type 1, 2: diagonalize_hsp_and_H11(0) with this iteration's
  // multipliers, then eta, mu of the recipe per state (module_gradient.f90:97-120)
  if (c->grad_type > 0) {
    TG_TRY(tg_hsp_h11_device(c, lam, c->eig_last, nullptr, nullptr));
    c->etamu.assign(2 * (size_t)S, 0.0);
    for (int st = 0; st < S; ++st)
      gradient_recipe(c->grad_type, n, c->eig_last.data() + 2 * (size_t)st * n,
                      c->eig_last.data() + (2 * (size_t)st + 1) * n, c->grad_eta, c->grad_mu,
                      &c->etamu[2 * st], &c->etamu[2 * st + 1]);
  }
  // calculate_gradient: Z = -eta (H20v - sum_j lambda_j C20_j) + mu Zim1
  TG_TRY(tg_gradient_step(c, c20, lam));
  std::vector<double> ph((size_t)S * ps);
  TG_HIP(hipMemcpyAsync(ph.data(), part, sizeof(double) * ph.size(), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  const long long nb = (nn + NT - 1) / NT;
  out.gnorm.assign(S, 0.0);
  out.conv.assign(S, 0);
  out.nadj.assign(S, 0);
  std::vector<int> evolve;
  for (int st = 0; st < S; ++st) {
    double sum = 0;
    for (long long k = 0; k < nb; ++k) sum += ph[(size_t)st * ps + k];
    out.gnorm[st] = sqrt(sum) / (c->grad_type > 0 ? c->etamu[2 * st] : c->grad_eta);
    out.conv[st] = out.gnorm[st] <= gradient_eps;
    if (!out.conv[st]) evolve.push_back(st);
  }
  if (evolve.empty()) return TG_OK;

  // adjust_constraints: evolve with Z, then fix the constraints with the C20,
  // every state in lockstep (a state whose constraints are met adds nothing to Z)
  const int maxit = 100;
  std::vector<int> open_state(S, 0);
  for (int st : evolve) open_state[st] = 1;
  double* xd = it.x();
  int* dinfo = it.info();
  std::vector<int> hinfo(S, 0);
  std::vector<double> xh((size_t)S * TG_MAX_CONS, 0.0);
  for (int k = 1; k <= maxit; ++k) {
    for (int st = 0; st < S; ++st)
      if (open_state[st]) out.nadj[st] = k;
    // cholesky_factorization: L L^T = 1 + Z^T Z
    TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Z, n, nn, Z, n, nn, 0.0, it.M(IT_S), n, nn, S));
    hipLaunchKernelGGL(add_identity_kernel, dim3(nblk(n), S), dim3(NT), 0, s, n, it.M(IT_S));
    TG_HIP(hipGetLastError());
    // evolve_wavefunction's U0 + V0 Z, V0 + U0 Z on the second stream beside the
    // Cholesky (and then the block rows of L^-1 there)
    hipStream_t s2 = c->stream2;
    TG_HIP(hipEventRecord(c->ev_fork, s));
    TG_HIP(hipStreamWaitEvent(s2, c->ev_fork, 0));
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s2, S * nn, U0, it.M(IT_A1));
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s2, S * nn, V0, it.M(IT_A2));
    TG_HIP(hipGetLastError());
    TG_TRY(gemm_dd(s2, OP_N, OP_N, n, n, n, 1.0, V0, n, nn, Z, n, nn, 1.0, it.M(IT_A1), n, nn, S));
    TG_TRY(gemm_dd(s2, OP_N, OP_N, n, n, n, 1.0, U0, n, nn, Z, n, nn, 1.0, it.M(IT_A2), n, nn, S));
    TG_TRY(cholesky_inverse(it, it.M(IT_S), it.M(IT_T2), it.M(IT_LINV), it.M(IT_X), dinfo, s2,
                            c->ev_fork));
    TG_HIP(hipEventRecord(c->ev_join, s2));
    TG_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
    TG_HIP(hipMemcpyAsync(hinfo.data(), dinfo, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    // evolve_wavefunction: U1 = (U0 + V0 Z) L^-T, V1 = (V0 + U0 Z) L^-T
    TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_A1), n, nn, it.M(IT_LINV), n, nn, 0.0,
                   it.M(IT_U1), n, nn, S));
    TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_A2), n, nn, it.M(IT_LINV), n, nn, 0.0,
                   it.M(IT_V1), n, nn, S));
    auto cholesky_ok = [&]() -> int {
      for (int st : evolve)
        if (hinfo[st] != 0) {
          tg_set_error("Cholesky factorization has failed, info = " + std::to_string(hinfo[st]) +
                       " (state " + std::to_string(st) + ")");
          return TG_E_SINGULAR;
        }
      return TG_OK;
    };
    if (dim == 0) {
      TG_HIP(hipStreamSynchronize(s));
      TG_TRY(cholesky_ok());
      break;
    }
    // expectation values <O_m>: Tr(O rho) (one-body), Tr(O kappa^T) (pair), the
    // constraint_20 at the evolved states and their Gram matrices, read back with the
    // Cholesky info at one synchronisation
    const bool any_pair = c->cons_pair > 0, any_obo = dim > c->cons_pair;
    double* rho = it.M(IT_R);
    double* kap = it.M(IT_T1);
    if (any_obo)
      TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_V1), n, nn, it.M(IT_V1), n, nn, 0.0,
                     rho, n, nn, S));
    if (any_pair)
      TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_V1), n, nn, it.M(IT_U1), n, nn, 0.0,
                     kap, n, nn, S));
    std::vector<int> act;
    for (int st = 0; st < S; ++st)
      if (open_state[st]) act.push_back(st);
    std::vector<const double*> xs, ys;
    for (int st : act)
      for (int m = 0; m < dim; ++m) {
        xs.push_back(c->d_cHO.as<double>() + (size_t)m * nn);
        ys.push_back((m >= dim - c->cons_pair ? kap : rho) + (size_t)st * nn);
      }
    DotsJob evj, grj;
    TG_TRY(dots_launch(it, it.tail(), xs, ys, evj));
    TG_TRY(constraints_qp(it, it.M(IT_U1), it.M(IT_V1), c20));
    TG_TRY(gram_launch(it, it.tail2(), c20, act, grj));
    TG_HIP(hipStreamSynchronize(s));
    TG_TRY(cholesky_ok());
    std::vector<double> ev;
    dots_finish(evj, ev);
    std::vector<std::vector<double>> Bv(act.size(), std::vector<double>(dim));
    std::vector<int> still;
    for (size_t q = 0; q < act.size(); ++q) {
      int open = 0;
      for (int m = 0; m < dim; ++m) {
        Bv[q][m] = cval(act[q], m) - ev[q * dim + m];
        if (fabs(Bv[q][m]) > c->cons_eps) ++open;
      }
      if (open == 0) open_state[act[q]] = 0;
      else still.push_back((int)q);
    }
    if (still.empty()) break;
    std::vector<std::vector<double>> A;
    gram_finish(it, grj, act.size(), A);
    std::fill(xh.begin(), xh.end(), 0.0);
    for (int q : still) {
      std::vector<double> b = Bv[q];
      if (!small_solve(dim, A[q], b)) {
        tg_set_error("adjust_constraints: singular constraint system (state " +
                     std::to_string(act[q]) + ")");
        return TG_E_SINGULAR;
      }
      std::copy(b.begin(), b.end(), xh.begin() + (size_t)act[q] * TG_MAX_CONS);
    }
    // xh is rewritten only after the next synchronisation of the stream
    TG_HIP(hipMemcpyAsync(xd, xh.data(), sizeof(double) * xh.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(axpy_cons_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, dim, xd, c20, Z);
    TG_HIP(hipGetLastError());
  }
  // the evolved states and the momentum term of their next step
  std::vector<int> mask(S, 0);
  for (int st : evolve) mask[st] = 1;
  int* dmask = it.mask();
  TG_HIP(hipMemcpyAsync(dmask, mask.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(masked_copy_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, dmask,
                     it.M(IT_U1), U0);
  hipLaunchKernelGGL(masked_copy_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, dmask,
                     it.M(IT_V1), V0);
  hipLaunchKernelGGL(masked_copy_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, dmask, Z, Zm1);
  TG_HIP(hipGetLastError());
  TG_HIP(hipStreamSynchronize(s));
  c->have_state = 0;   // the canonical basis of the new states is still to be set
  c->evolved = 1;
  return TG_OK;
}

void copy_out(const tg_ctx* c, const IterOut& o, double* pnp_scalars, double* gradient_norm,
              double* lagrange_lambda1, int* adjust_iterations, int* converged) {
  const int S = c->ns;
  if (pnp_scalars) std::copy(o.scal.begin(), o.scal.end(), pnp_scalars);
  if (gradient_norm) std::copy(o.gnorm.begin(), o.gnorm.end(), gradient_norm);
  if (lagrange_lambda1) std::copy(o.lambda.begin(), o.lambda.end(), lagrange_lambda1);
  for (int st = 0; st < S; ++st) {
    if (adjust_iterations) adjust_iterations[st] = o.nadj.empty() ? 0 : o.nadj[st];
    if (converged) converged[st] = o.conv.empty() ? 0 : o.conv[st];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// group of contexts projected as one graph launch
// ---------------------------------------------------------------------------
struct tg_group {
  std::vector<tg_ctx*> ctx;
  hipEvent_t ev_start = nullptr;
  std::vector<hipEvent_t> ev_done;
  int use_graph = 1, graph_valid = 0, graph_seen = 0;
  std::vector<GraphKey> key_v, seen_key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<IterOut> last;
};

namespace {

GraphKey ctx_key(tg_ctx* c) {
  GraphKey k{};
  k.v[0] = (long long)(intptr_t)c;
  k.v[1] = c->n;
  k.v[2] = c->A;
  k.v[3] = c->ns;
  k.v[4] = c->iprint;
  k.v[5] = c->grad_on;
  k.v[6] = c->grad_ncons;
  k.v[7] = c->c20_shared;
  k.v[8] = (long long)devbuf_epoch();
  k.v[9] = c->Mphip;
  k.v[10] = c->Mphin;
  unsigned long long h = 1469598103934665603ULL;
  for (const StateInfo& si : c->st)
    for (int x : {si.nocc, si.nemp}) {
      h ^= (unsigned long long)(unsigned)x;
      h *= 1099511628211ULL;
    }
  k.v[11] = (long long)h;
  double d[3] = {c->factor_delta, c->grad_eta, c->grad_mu};
  std::memcpy(&k.v[12], d, sizeof(d));
  k.v[15] = c->plain;
  k.v[16] = c->w_fp32;
  k.v[17] = c->batch_limit;
  k.v[18] = c->mem_chunk;
  return k;
}

void group_release_graph(tg_group* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
  g->graph_valid = 0;
}

// Launches of every context: context 0 on its own stream, the others forked
// from it and joined back, so one capture on context 0's stream holds them all.
int group_launch(tg_group* g) {
  tg_ctx* c0 = g->ctx[0];
  const size_t k = g->ctx.size();
  TG_HIP(hipEventRecord(g->ev_start, c0->stream));
  for (size_t i = 1; i < k; ++i) TG_HIP(hipStreamWaitEvent(g->ctx[i]->stream, g->ev_start, 0));
  for (size_t i = 1; i < k; ++i) TG_TRY(tg_project_run(g->ctx[i]));
  TG_TRY(tg_project_run(c0));
  for (size_t i = 1; i < k; ++i) {
    TG_HIP(hipEventRecord(g->ev_done[i], g->ctx[i]->stream));
    TG_HIP(hipStreamWaitEvent(c0->stream, g->ev_done[i], 0));
  }
  return TG_OK;
}

// the projections of all contexts; graph capture on the second call of a
// configuration, replay afterwards.  Every context's stream is ordered after
// the work when this returns.
int group_project(tg_group* g) {
  for (tg_ctx* c : g->ctx) {
    if (!c->have_state) {
      tg_set_error("tg_group_project: a context has no state");
      return TG_E_ARG;
    }
    TG_TRY(tg_project_upload(c));
  }
  tg_ctx* c0 = g->ctx[0];
  std::vector<GraphKey> keys;
  for (tg_ctx* c : g->ctx) keys.push_back(ctx_key(c));
  bool timing = false;
  for (tg_ctx* c : g->ctx) timing = timing || c->timing;
  int rc = TG_OK;
  if (!g->use_graph || timing) {   // event pairs of the timing are not captured
    rc = group_launch(g);
  } else if (g->graph_valid && keys == g->key_v) {
    TG_HIP(hipGraphLaunch(g->exec, c0->stream));
  } else if (!(g->graph_seen && keys == g->seen_key)) {
    g->graph_seen = 1;
    g->seen_key = keys;
    rc = group_launch(g);
  } else {
    group_release_graph(g);
    TG_HIP(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
    const int r = group_launch(g);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c0->stream, &gr);
    if (r != TG_OK || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (r != TG_OK) return r;
      tg_set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return TG_E_HIP;
    }
    g->graph = gr;
    if (hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0) != hipSuccess) {
      g->exec = nullptr;
      group_release_graph(g);
      tg_set_error("hipGraphInstantiate failed");
      return TG_E_HIP;
    }
    g->graph_valid = 1;
    g->key_v = keys;
    if (std::getenv("TG_GRAPH_VERBOSE")) {
      size_t nodes = 0;
      (void)hipGraphGetNodes(gr, nullptr, &nodes);
      fprintf(stderr, "[taurus_gpu] group of %zu contexts captured: %zu graph nodes\n",
              g->ctx.size(), nodes);
    }
    TG_HIP(hipGraphLaunch(g->exec, c0->stream));
  }
  TG_TRY(rc);
  // later work and synchronisations on any context's stream follow the launch
  TG_HIP(hipEventRecord(g->ev_start, c0->stream));
  for (size_t i = 1; i < g->ctx.size(); ++i)
    TG_HIP(hipStreamWaitEvent(g->ctx[i]->stream, g->ev_start, 0));
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_set_constraints(tg_ctx* c, int constraint_dim, int constraint_pair,
                       const double* constraint_HO, int nsets, const double* constraint_val,
                       double constraint_eps) {
  if (!c || constraint_dim < 0 || constraint_dim > TG_MAX_CONS || constraint_pair < 0 ||
      constraint_pair > constraint_dim || nsets < 1 ||
      (constraint_dim > 0 && (!constraint_HO || !constraint_val))) {
    tg_set_error("tg_set_constraints: invalid argument");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  c->cons_dim = constraint_dim;
  c->cons_pair = constraint_pair;
  c->cons_eps = constraint_eps;
  c->cons_nsets = nsets;
  if (constraint_dim > 0)
    c->cons_val.assign(constraint_val, constraint_val + (size_t)nsets * constraint_dim);
  else
    c->cons_val.clear();
  c->lambda_prev.clear();
  if (constraint_dim == 0) return TG_OK;
  std::vector<double> h((size_t)constraint_dim * nn);
  for (int m = 0; m < constraint_dim; ++m)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        h[m * nn + i + (size_t)j * n] =
            constraint_HO[m * nn + c->sp_perm[i] + (size_t)c->sp_perm[j] * n];
  TG_TRY(c->d_cHO.ensure(sizeof(double) * h.size()));
  TG_HIP(hipMemcpy(c->d_cHO.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  return TG_OK;
}

int tg_iterate(tg_ctx* c, double gradient_eps, double* pnp_scalars, double* gradient_norm,
               double* lagrange_lambda1, int* adjust_iterations, int* converged) {
  if (!c) {
    tg_set_error("tg_iterate: invalid argument");
    return TG_E_ARG;
  }
  TG_TRY(iterate_pre(c));
  // project_wavefunction (the gradient step is taken after, with new multipliers)
  const int grad_was = c->grad_on;
  c->grad_on = 0;
  const int r = tg_project_device(c);
  c->grad_on = grad_was;
  TG_TRY(r);
  IterOut o;
  TG_TRY(iterate_post(c, gradient_eps, o));
  copy_out(c, o, pnp_scalars, gradient_norm, lagrange_lambda1, adjust_iterations, converged);
  return TG_OK;
}

int tg_iterate_canonical(tg_ctx* c, double canon_eps, double gradient_eps, double* pnp_scalars,
                         double* gradient_norm, double* lagrange_lambda1, int* adjust_iterations,
                         int* converged, int* nocc, int* nemp, double* ovacc, double* voveru) {
  if (!c) {
    tg_set_error("tg_iterate_canonical: invalid argument");
    return TG_E_ARG;
  }
  // The canonical basis runs beside the projection when the states are the last
  // iteration's (the projection's set-up current but for the canonical data) and
  // the projection runs eagerly in one batch; otherwise one after the other.
  bool overlap = c->evolved && !c->state_dirty && !c->plain && !c->use_graph && c->grad_on;
  if (overlap) {
    const int B = c->ns * c->A;
    overlap = (c->batch_limit <= 0 || c->batch_limit >= B) && c->mem_chunk <= 0 &&
              std::max(c->Mphip, c->Mphin) > 1;
  }
  if (!overlap) {
    TG_TRY(tg_canonical_basis(c, canon_eps, nocc, nemp, ovacc, voveru));
    return tg_iterate(c, gradient_eps, pnp_scalars, gradient_norm, lagrange_lambda1,
                      adjust_iterations, converged);
  }
  // update_constraints_qpbasis and the projection up to its overlaps read U, V
  // only; the hook runs on the second stream after the projection's fork (the
  // U, V products are done), before its join
  c->have_state = 1;
  if (const int r = iterate_pre(c)) {
    c->have_state = 0;
    return r;
  }
  c->defer_overlap = 1;
  hipEvent_t cb0 = nullptr;   // start of the canonical basis on the second stream (timing)
  c->defer_hook = [&](int phase) -> int {
    if (phase == 0) {
      if (c->timing && hipEventCreate(&cb0) == hipSuccess) (void)hipEventRecord(cb0, c->stream2);
      return canonical_launch(c, c->stream2, canon_eps);
    }
    TG_TRY(canonical_finish(c, c->stream2, canon_eps, nocc, nemp, ovacc, voveru));
    TG_TRY(upload_overlap_setup(c, c->stream2));
    if (cb0) {
      hipEvent_t cb1 = nullptr;
      if (hipEventCreate(&cb1) == hipSuccess) {
        (void)hipEventRecord(cb1, c->stream2);
        c->pending.push_back({std::string("canonical_basis_beside"), cb0, cb1});
      } else {
        (void)hipEventDestroy(cb0);
      }
      cb0 = nullptr;
    }
    c->state_dirty = 0;   // the rest of the projection's set-up did not change
    return TG_OK;
  };
  const int grad_was = c->grad_on;
  c->grad_on = 0;
  const int r = tg_project_run(c);
  c->grad_on = grad_was;
  c->defer_overlap = 0;
  c->defer_hook = nullptr;
  if (cb0) (void)hipEventDestroy(cb0);
  if (r != TG_OK) {
    // leave nothing in flight (a wide degenerate block, TG_E_UNSUPPORTED, leaves
    // the state as it was: the caller takes the host route and tg_iterate)
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream3);
    (void)hipStreamSynchronize(c->stream);
    c->have_state = 0;
    c->evolved = 1;
    c->state_dirty = 1;
    return r;
  }
  IterOut o;
  TG_TRY(iterate_post(c, gradient_eps, o));
  copy_out(c, o, pnp_scalars, gradient_norm, lagrange_lambda1, adjust_iterations, converged);
  return TG_OK;
}

int tg_set_gradient_type(tg_ctx* c, int gradient_type) {
  if (!c || gradient_type < 0 || gradient_type > 2) {
    tg_set_error("tg_set_gradient_type: gradient_type must be 0, 1 or 2");
    return TG_E_ARG;
  }
  c->grad_type = gradient_type;
  c->etamu.clear();
  return TG_OK;
}

int tg_diagonalize_hsp_and_H11(tg_ctx* c, int constraint_dim, const double* lagrange_lambda1,
                               double* eigen_hsp, double* eigen_H11, int* info_hsp, int* info_H11,
                               double* hsp_RR, double* field_H11) {
  if (!c || !c->have_state || c->ns != 1 || constraint_dim != c->cons_dim ||
      (constraint_dim > 0 && !lagrange_lambda1)) {
    tg_set_error("tg_diagonalize_hsp_and_H11: one resident state and the constraints of "
                 "tg_set_constraints are needed");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  TG_TRY(tg_ensure_z(c));
  // the multipliers go to the device slot of the gradient step
  double* lam = c->d_Z.as<double>() + 2 * nn;
  std::vector<double> lamh(TG_MAX_CONS, 0.0);
  for (int m = 0; m < constraint_dim; ++m) lamh[m] = lagrange_lambda1[m];
  TG_HIP(hipMemcpyAsync(lam, lamh.data(), sizeof(double) * TG_MAX_CONS, hipMemcpyHostToDevice,
                        c->stream));
  double *hd = nullptr, *h11d = nullptr;
  std::vector<double> eig;
  TG_TRY(tg_hsp_h11_device(c, lam, eig, &hd, &h11d));
  if (eigen_hsp) std::copy(eig.begin(), eig.begin() + n, eigen_hsp);
  if (eigen_H11) std::copy(eig.begin() + n, eig.begin() + 2 * n, eigen_H11);
  if (info_hsp) *info_hsp = 0;
  if (info_H11) *info_H11 = 0;
  // the matrices handed to the eigenvalue solver: h in the single-particle basis
  // (rows and columns back to the reference order), H11 in the quasiparticle
  // basis (the columns of U, V, which keep the reference order)
  if (hsp_RR) {
    std::vector<double> hm(nn);
    TG_HIP(hipMemcpy(hm.data(), hd, sizeof(double) * nn, hipMemcpyDeviceToHost));
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        hsp_RR[c->sp_perm[i] + (size_t)c->sp_perm[j] * n] = hm[i + (size_t)j * n];
  }
  if (field_H11) TG_HIP(hipMemcpy(field_H11, h11d, sizeof(double) * nn, hipMemcpyDeviceToHost));
  return TG_OK;
}

int tg_get_hsp_h11_eigenvalues(tg_ctx* c, int state, double* eigen_hsp, double* eigen_H11,
                               double* gradient_eta, double* gradient_mu) {
  if (!c || state < 0 || state >= c->ns || c->eig_last.size() != 2 * (size_t)c->ns * c->n) {
    tg_set_error("tg_get_hsp_h11_eigenvalues: no spectra (gradient_type 1 or 2 iteration)");
    return TG_E_ARG;
  }
  const int n = c->n;
  const double* e = c->eig_last.data() + 2 * (size_t)state * n;
  if (eigen_hsp) std::copy(e, e + n, eigen_hsp);
  if (eigen_H11) std::copy(e + n, e + 2 * n, eigen_H11);
  if (gradient_eta) *gradient_eta = c->etamu[2 * state];
  if (gradient_mu) *gradient_mu = c->etamu[2 * state + 1];
  return TG_OK;
}

int tg_sym_eigenvalues(int n, int count, const double* A, int lda, double* w) {
  if (n < 0 || count < 0 || lda < std::max(1, n) || n > 512 || (count > 0 && (!A || !w))) {
    tg_set_error("tg_sym_eigenvalues: invalid argument (n <= 512)");
    return TG_E_ARG;
  }
  if (count == 0 || n == 0) return TG_OK;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) {
    tg_set_error("no HIP device available");
    return TG_E_NODEVICE;
  }
  const size_t m = (size_t)lda * n;
  double *dA = nullptr, *dW = nullptr, *dw = nullptr;
  int r = TG_OK;
  if (hipMalloc(&dA, sizeof(double) * m * count) != hipSuccess ||
      hipMalloc(&dW, sizeof(double) * (size_t)n * n * count) != hipSuccess ||
      hipMalloc(&dw, sizeof(double) * (size_t)n * count) != hipSuccess) {
    tg_set_error("tg_sym_eigenvalues: hipMalloc failed");
    r = TG_E_NOMEM;
  }
  if (r == TG_OK && hipMemcpy(dA, A, sizeof(double) * m * count, hipMemcpyHostToDevice) != hipSuccess)
    r = TG_E_HIP;
  if (r == TG_OK) r = batched_sym_eigvals(0, n, dA, lda, (long long)m, count, dW, dw, n);
  if (r == TG_OK && (hipDeviceSynchronize() != hipSuccess ||
                     hipMemcpy(w, dw, sizeof(double) * (size_t)n * count, hipMemcpyDeviceToHost) !=
                         hipSuccess))
    r = TG_E_HIP;
  if (dA) (void)hipFree(dA);
  if (dW) (void)hipFree(dW);
  if (dw) (void)hipFree(dw);
  return r;
}

int tg_get_state(tg_ctx* c, double* U, double* V) {
  if (!c || !U || !V || !c->d_U.p) {
    tg_set_error("tg_get_state: invalid argument");
    return TG_E_ARG;
  }
  const int n = c->n, S = c->ns;
  const size_t nn = (size_t)n * n;
  std::vector<double> h(2 * nn * S);
  TG_HIP(hipStreamSynchronize(c->stream));
  TG_HIP(hipMemcpy(h.data(), c->d_U.p, sizeof(double) * nn * S, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(h.data() + nn * S, c->d_V.p, sizeof(double) * nn * S, hipMemcpyDeviceToHost));
  // rows back to the reference single-particle order
  for (int k = 0; k < S; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t dst = k * nn + c->sp_perm[i] + (size_t)j * n, src = k * nn + i + (size_t)j * n;
        U[dst] = h[src];
        V[dst] = h[nn * S + src];
      }
  return TG_OK;
}

int tg_group_create(tg_group** out, int nctx, tg_ctx** ctxs, int use_graph) {
  if (!out || nctx < 1 || !ctxs) {
    tg_set_error("tg_group_create: invalid argument");
    return TG_E_ARG;
  }
  for (int i = 0; i < nctx; ++i) {
    bool bad = !ctxs[i];
    for (int j = 0; j < i && !bad; ++j) bad = ctxs[i] == ctxs[j];
    if (bad) {
      tg_set_error("tg_group_create: contexts must be distinct and non-null");
      return TG_E_ARG;
    }
  }
  tg_group* g = new tg_group();
  g->ctx.assign(ctxs, ctxs + nctx);
  g->use_graph = use_graph ? 1 : 0;
  g->ev_done.assign(nctx, nullptr);
  bool ok = hipEventCreateWithFlags(&g->ev_start, hipEventDisableTiming) == hipSuccess;
  for (int i = 1; i < nctx && ok; ++i)
    ok = hipEventCreateWithFlags(&g->ev_done[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    tg_group_destroy(g);
    tg_set_error("tg_group_create: event creation failed");
    return TG_E_HIP;
  }
  *out = g;
  return TG_OK;
}

int tg_group_destroy(tg_group* g) {
  if (!g) return TG_OK;
  for (tg_ctx* c : g->ctx)
    if (c && c->stream) (void)hipStreamSynchronize(c->stream);
  group_release_graph(g);
  if (g->ev_start) (void)hipEventDestroy(g->ev_start);
  for (hipEvent_t e : g->ev_done)
    if (e) (void)hipEventDestroy(e);
  delete g;
  return TG_OK;
}

int tg_group_project(tg_group* g) {
  if (!g) return TG_E_ARG;
  return group_project(g);
}

int tg_group_iterate(tg_group* g, double gradient_eps) {
  if (!g) return TG_E_ARG;
  for (tg_ctx* c : g->ctx) TG_TRY(iterate_pre(c));
  std::vector<int> was;
  for (tg_ctx* c : g->ctx) {
    was.push_back(c->grad_on);
    c->grad_on = 0;
  }
  const int r = group_project(g);
  for (size_t i = 0; i < g->ctx.size(); ++i) g->ctx[i]->grad_on = was[i];
  TG_TRY(r);
  g->last.assign(g->ctx.size(), IterOut());
  for (size_t i = 0; i < g->ctx.size(); ++i) TG_TRY(iterate_post(g->ctx[i], gradient_eps, g->last[i]));
  return TG_OK;
}

int tg_group_iterate_results(tg_group* g, int index, double* pnp_scalars, double* gradient_norm,
                             double* lagrange_lambda1, int* adjust_iterations, int* converged) {
  if (!g || index < 0 || index >= (int)g->last.size()) {
    tg_set_error("tg_group_iterate_results: no such context or no iteration yet");
    return TG_E_ARG;
  }
  copy_out(g->ctx[index], g->last[index], pnp_scalars, gradient_norm, lagrange_lambda1,
           adjust_iterations, converged);
  return TG_OK;
}

}  // extern "C"
