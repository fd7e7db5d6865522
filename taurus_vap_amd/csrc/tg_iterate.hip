// One gradient iteration of the reference driver on the device
// (taurus_vap.f90:79-117 with gradient_type 0):
//   update_constraints_qpbasis          module_constraints.f90:381-397
//   project_wavefunction(0, ...)        (tg_project.hip)
//   calculate_lagrange_multipliers      module_constraints.f90:413-484
//   calculate_gradient                  module_gradient.f90:79-143
//   adjust_constraints                  module_constraints.f90:270-373
//     cholesky_factorization            module_constraints.f90:497-543
//     evolve_wavefunction               module_constraints.f90:559-590
// The constraint operators and the Bogoliubov matrices stay in HBM; the host
// sees only the handful of constraint numbers that decide the adjustment loop
// and its dim x dim linear systems.  The canonical basis of the new state is
// left to the caller (tg_get_state, then tg_set_state).
#include <cmath>
#include <string>
#include <vector>

#include "tg_internal.h"

namespace {

constexpr int NT = 256;
constexpr int NT1 = 1024;   // single-workgroup factorisations

inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

// C = (A - A^T) / 2 with A = T1 - T2  (set_operator_qpbasis 'f20'/'g20', factor -1)
__global__ void __launch_bounds__(256)
skew_half_kernel(int n, const double* __restrict__ T1, const double* __restrict__ T2,
                 double* __restrict__ C) {
  __shared__ double ta[32][33], tb[32][33];
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

// out[p] = sum_k X_p[k] Y_p[k] for the pointer pairs of the list (one block per pair)
__global__ void __launch_bounds__(NT)
dots_kernel(long long len, const double* const* __restrict__ xs, const double* const* __restrict__ ys,
            double* __restrict__ out) {
  const double* x = xs[blockIdx.x];
  const double* y = ys[blockIdx.x];
  double s = 0.0;
  for (long long k = threadIdx.x; k < len; k += NT) s = fma(x[k], y[k], s);
  __shared__ double red[NT];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void add_identity_kernel(int n, double* __restrict__ S) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) S[i + (long long)i * n] += 1.0;
}

__global__ void copy_kernel(long long len, const double* __restrict__ src, double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < len) dst[i] = src[i];
}

// Z += sum_m x_m C20_m
__global__ void axpy_cons_kernel(long long nn, int ncons, const double* __restrict__ x,
                                 const double* __restrict__ c20, double* __restrict__ Z) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= nn) return;
  double s = Z[i];
  for (int m = 0; m < ncons; ++m) s = fma(x[m], c20[(long long)m * nn + i], s);
  Z[i] = s;
}

// In-place lower Cholesky factor of a symmetric positive matrix (column-major,
// right-looking, one workgroup); the strict upper triangle is zeroed.
// info = 0, or the first column with a non-positive pivot (1-based).
__global__ void __launch_bounds__(NT1)
cholesky_kernel(int n, double* __restrict__ A, int* __restrict__ info) {
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      const double d = A[k + (long long)k * n];
      if (!(d > 0.0)) {
        if (bad == 0) bad = k + 1;
        A[k + (long long)k * n] = 1.0;
      } else {
        A[k + (long long)k * n] = sqrt(d);
      }
    }
    __syncthreads();
    const double dk = A[k + (long long)k * n];
    for (int i = k + 1 + t; i < n; i += NT1) A[i + (long long)k * n] /= dk;
    __syncthreads();
    const int m = n - k - 1;
    for (long long e = t; e < (long long)m * m; e += NT1) {
      const int i = k + 1 + (int)(e % m), j = k + 1 + (int)(e / m);
      if (i >= j) A[i + (long long)j * n] -= A[i + (long long)k * n] * A[j + (long long)k * n];
    }
    __syncthreads();
  }
  for (long long e = t; e < (long long)n * n; e += NT1) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i < j) A[e] = 0.0;
  }
  if (t == 0) *info = bad;
}

// X = L^-1 for lower-triangular L (dtrtri 'L','N'): thread j solves column j
__global__ void __launch_bounds__(NT1)
trtri_lower_kernel(int n, const double* __restrict__ L, double* __restrict__ X) {
  for (int j = threadIdx.x; j < n; j += NT1) {
    for (int i = 0; i < j; ++i) X[i + (long long)j * n] = 0.0;
    X[j + (long long)j * n] = 1.0 / L[j + (long long)j * n];
    for (int i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (int k = j; k < i; ++k) s = fma(L[i + (long long)k * n], X[k + (long long)j * n], s);
      X[i + (long long)j * n] = -s / L[i + (long long)i * n];
    }
  }
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

// work layout: IT_COUNT matrices of n*n doubles, then a tail of
//   [0, TAIL_X)      dot-product pointer lists and results (<= 3 * 2080 words)
//   [TAIL_X, +64)    constraint corrections x
//   [TAIL_INFO]      Cholesky info
enum { IT_S, IT_LINV, IT_A1, IT_A2, IT_U1, IT_V1, IT_R, IT_T1, IT_T2, IT_X, IT_COUNT };
constexpr size_t TAIL_X = 2 * TG_MAX_CONS * TG_MAX_CONS;
constexpr size_t TAIL_INFO = 4 * TG_MAX_CONS * TG_MAX_CONS;
constexpr size_t TAIL_SIZE = TAIL_INFO + 64;

struct Iter {
  tg_ctx* c;
  int n;
  long long nn;
  hipStream_t s;
  double* w;
  double* M(int k) const { return w + (size_t)k * nn; }
};

// constraint_20 of every constraint at (U, V) into c20 (dim x n*n)
int constraints_qp(const Iter& it, const double* U, const double* V, double* c20) {
  tg_ctx* c = it.c;
  const int n = it.n;
  const long long nn = it.nn;
  for (int m = 0; m < c->cons_dim; ++m) {
    const double* O = c->d_cHO.as<double>() + (size_t)m * nn;
    const bool pair = m >= c->cons_dim - c->cons_pair;
    double* X = it.M(IT_X);
    // f20: T1 = U^T O V, T2 = V^T O^T U;  g20: T1 = U^T O U, T2 = V^T O V
    TG_TRY(gemm_dd(it.s, OP_N, OP_N, n, n, n, 1.0, O, n, 0, pair ? U : V, n, 0, 0.0, X, n, 0, 1));
    TG_TRY(gemm_dd(it.s, OP_T, OP_N, n, n, n, 1.0, U, n, 0, X, n, 0, 0.0, it.M(IT_T1), n, 0, 1));
    if (pair)
      TG_TRY(gemm_dd(it.s, OP_N, OP_N, n, n, n, 1.0, O, n, 0, V, n, 0, 0.0, X, n, 0, 1));
    else
      TG_TRY(gemm_dd(it.s, OP_T, OP_N, n, n, n, 1.0, O, n, 0, U, n, 0, 0.0, X, n, 0, 1));
    TG_TRY(gemm_dd(it.s, OP_T, OP_N, n, n, n, 1.0, V, n, 0, X, n, 0, 0.0, it.M(IT_T2), n, 0, 1));
    hipLaunchKernelGGL(skew_half_kernel, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, it.s, n,
                       it.M(IT_T1), it.M(IT_T2), c20 + (size_t)m * nn);
    TG_HIP(hipGetLastError());
  }
  return TG_OK;
}

// host copies of sum_k X_p[k] Y_p[k] for the pairs (xs[p], ys[p])
int dots(const Iter& it, const std::vector<const double*>& xs, const std::vector<const double*>& ys,
         std::vector<double>& out) {
  const int np = (int)xs.size();
  out.assign(np, 0.0);
  if (np == 0) return TG_OK;
  // pointer lists and results in the tail of the work area (TAIL_DOTS)
  const double** pl = reinterpret_cast<const double**>(it.w + (size_t)IT_COUNT * it.nn);
  double* res = reinterpret_cast<double*>(pl + 2 * np);
  if (np > TG_MAX_CONS * (TG_MAX_CONS + 1) / 2) {
    tg_set_error("constraint dot products: too many pairs");
    return TG_E_ARG;
  }
  std::vector<const double*> h(2 * np);
  for (int p = 0; p < np; ++p) {
    h[p] = xs[p];
    h[np + p] = ys[p];
  }
  TG_HIP(hipMemcpyAsync(pl, h.data(), sizeof(double*) * 2 * np, hipMemcpyHostToDevice, it.s));
  hipLaunchKernelGGL(dots_kernel, dim3(np), dim3(NT), 0, it.s, it.nn, pl, pl + np, res);
  TG_HIP(hipGetLastError());
  TG_HIP(hipMemcpyAsync(out.data(), res, sizeof(double) * np, hipMemcpyDeviceToHost, it.s));
  TG_HIP(hipStreamSynchronize(it.s));
  return TG_OK;
}

// Gram matrix A_ij = Tr(C20_i^T C20_j) (symmetric, dim x dim, host)
int gram(const Iter& it, const double* c20, std::vector<double>& A) {
  const int m = it.c->cons_dim;
  std::vector<const double*> xs, ys;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) {
      xs.push_back(c20 + (size_t)i * it.nn);
      ys.push_back(c20 + (size_t)j * it.nn);
    }
  std::vector<double> d;
  TG_TRY(dots(it, xs, ys, d));
  A.assign((size_t)m * m, 0.0);
  int k = 0;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j, ++k) A[i + j * m] = A[j + i * m] = d[k];
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_set_constraints(tg_ctx* c, int constraint_dim, int constraint_pair,
                       const double* constraint_HO, const double* constraint_val,
                       double constraint_eps) {
  if (!c || constraint_dim < 0 || constraint_dim > TG_MAX_CONS || constraint_pair < 0 ||
      constraint_pair > constraint_dim ||
      (constraint_dim > 0 && (!constraint_HO || !constraint_val))) {
    tg_set_error("tg_set_constraints: invalid argument");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  c->cons_dim = constraint_dim;
  c->cons_pair = constraint_pair;
  c->cons_eps = constraint_eps;
  c->cons_val.assign(constraint_val, constraint_val + constraint_dim);
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
  if (!c || !c->have_state || !c->grad_on) {
    tg_set_error("tg_iterate: tg_set_state and tg_set_gradient must be called first");
    return TG_E_ARG;
  }
  const int n = c->n, dim = c->cons_dim;
  const long long nn = (long long)n * n;
  hipStream_t s = c->stream;
  TG_TRY(c->d_it.ensure(sizeof(double) * ((size_t)IT_COUNT * nn + TAIL_SIZE)));
  TG_TRY(c->d_c20.ensure(sizeof(double) * std::max<size_t>(1, (size_t)dim * nn)));
  Iter it{c, n, nn, s, c->d_it.as<double>()};
  double* U0 = c->d_U.as<double>();
  double* V0 = c->d_V.as<double>();
  double* c20 = c->d_c20.as<double>();
  double* Zb = c->d_Z.as<double>();
  double* Z = Zb;                 // gradient_Zi
  double* Zm1 = Zb + nn;          // gradient_Zim1
  double* lam = Zb + 2 * nn;
  double* part = lam + TG_MAX_CONS;

  // update_constraints_qpbasis
  TG_TRY(constraints_qp(it, U0, V0, c20));
  // project_wavefunction (the gradient step is taken below, with new multipliers)
  const int grad_was = c->grad_on;
  c->grad_on = 0;
  int r = tg_project_device(c);
  c->grad_on = grad_was;
  TG_TRY(r);
  // projected scalars, and the reference's stop on a complex gradient
  // (calculate_projected_H20, module_projection.f90:751-755)
  TG_TRY(tg_fetch_results(c, pnp_scalars, nullptr, nullptr, nullptr, nullptr, nullptr));
  double* H20v = reinterpret_cast<double*>(c->d_out.as<double>() + 6 * nn);
  // calculate_lagrange_multipliers: A_ij = Tr(C_i C_j^T), b_i = Tr(H20 C_i^T)
  std::vector<double> lamh(dim, 0.0);
  if (dim > 0) {
    std::vector<double> A, b;
    TG_TRY(gram(it, c20, A));
    std::vector<const double*> xs, ys;
    for (int i = 0; i < dim; ++i) {
      xs.push_back(H20v);
      ys.push_back(c20 + (size_t)i * nn);
    }
    TG_TRY(dots(it, xs, ys, b));
    if (small_solve(dim, A, b)) {
      lamh = b;
      c->lambda_prev = lamh;
    } else {
      // the reference keeps the multipliers of the previous iteration
      lamh = c->lambda_prev;
      lamh.resize(dim, 0.0);
    }
    TG_HIP(hipMemcpyAsync(lam, lamh.data(), sizeof(double) * dim, hipMemcpyHostToDevice, s));
  }
  // calculate_gradient (gradient_type 0)
  c->grad_ncons = dim;
  {
    const long long nb = (nn + NT - 1) / NT;
    // Z = -eta (H20v - sum_j lambda_j C20_j) + mu Zim1; partial sums of Z^2
    TG_TRY(tg_gradient_step(c, H20v, Z, Zm1, lam, c20, part));
    std::vector<double> ph(nb);
    TG_HIP(hipMemcpyAsync(ph.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    double sum = 0;
    for (double v : ph) sum += v;
    const double gn = sqrt(sum) / c->grad_eta;
    if (gradient_norm) *gradient_norm = gn;
    if (converged) *converged = gn <= gradient_eps;
    if (lagrange_lambda1)
      for (int i = 0; i < dim; ++i) lagrange_lambda1[i] = lamh[i];
    if (gn <= gradient_eps) {
      if (adjust_iterations) *adjust_iterations = 0;
      return TG_OK;
    }
  }
  // adjust_constraints: evolve with Z, then fix the constraints with the C20
  int iters = 0;
  const int maxit = 100;
  int* dinfo = reinterpret_cast<int*>(it.w + (size_t)IT_COUNT * nn + TAIL_INFO);
  for (int k = 1; k <= maxit; ++k) {
    iters = k;
    // cholesky_factorization: L L^T = 1 + Z^T Z
    TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, Z, n, 0, Z, n, 0, 0.0, it.M(IT_S), n, 0, 1));
    hipLaunchKernelGGL(add_identity_kernel, dim3(nblk(n)), dim3(NT), 0, s, n, it.M(IT_S));
    hipLaunchKernelGGL(cholesky_kernel, dim3(1), dim3(NT1), 0, s, n, it.M(IT_S), dinfo);
    hipLaunchKernelGGL(trtri_lower_kernel, dim3(1), dim3(NT1), 0, s, n, it.M(IT_S),
                       it.M(IT_LINV));
    TG_HIP(hipGetLastError());
    int info = 0;
    TG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    // evolve_wavefunction: U1 = (U0 + V0 Z) L^-T, V1 = (V0 + U0 Z) L^-T
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, U0, it.M(IT_A1));
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, V0, it.M(IT_A2));
    TG_HIP(hipGetLastError());
    TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, V0, n, 0, Z, n, 0, 1.0, it.M(IT_A1), n, 0, 1));
    TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, U0, n, 0, Z, n, 0, 1.0, it.M(IT_A2), n, 0, 1));
    TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_A1), n, 0, it.M(IT_LINV), n, 0, 0.0,
                   it.M(IT_U1), n, 0, 1));
    TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_A2), n, 0, it.M(IT_LINV), n, 0, 0.0,
                   it.M(IT_V1), n, 0, 1));
    TG_HIP(hipStreamSynchronize(s));
    if (info != 0) {
      tg_set_error("Cholesky factorization has failed, info = " + std::to_string(info));
      return TG_E_SINGULAR;
    }
    if (dim == 0) break;
    // expectation values <O_m>: Tr(O rho) (one-body), Tr(O kappa^T) (pair)
    std::vector<const double*> xs, ys;
    std::vector<double> ev;
    const bool any_pair = c->cons_pair > 0, any_obo = dim > c->cons_pair;
    double* rho = it.M(IT_R);
    double* kap = it.M(IT_T1);
    if (any_obo)
      TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_V1), n, 0, it.M(IT_V1), n, 0, 0.0, rho,
                     n, 0, 1));
    if (any_pair)
      TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, it.M(IT_V1), n, 0, it.M(IT_U1), n, 0, 0.0, kap,
                     n, 0, 1));
    for (int m = 0; m < dim; ++m) {
      xs.push_back(c->d_cHO.as<double>() + (size_t)m * nn);
      ys.push_back(m >= dim - c->cons_pair ? kap : rho);
    }
    TG_TRY(dots(it, xs, ys, ev));
    std::vector<double> B(dim);
    int open = 0;
    for (int m = 0; m < dim; ++m) {
      B[m] = c->cons_val[m] - ev[m];
      if (fabs(B[m]) > c->cons_eps) ++open;
    }
    // constraint_20 at the evolved state
    TG_TRY(constraints_qp(it, it.M(IT_U1), it.M(IT_V1), c20));
    if (open == 0) break;
    std::vector<double> A;
    TG_TRY(gram(it, c20, A));
    if (!small_solve(dim, A, B)) {
      tg_set_error("adjust_constraints: singular constraint system");
      return TG_E_SINGULAR;
    }
    double* xd = it.w + (size_t)IT_COUNT * nn + TAIL_X;
    TG_HIP(hipMemcpyAsync(xd, B.data(), sizeof(double) * dim, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(axpy_cons_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, dim, xd, c20, Z);
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(s));
  }
  // the evolved state and the momentum term of the next step
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, it.M(IT_U1), U0);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, it.M(IT_V1), V0);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, Z, Zm1);
  TG_HIP(hipGetLastError());
  TG_HIP(hipStreamSynchronize(s));
  if (adjust_iterations) *adjust_iterations = iters;
  c->have_state = 0;   // the canonical basis of the new state is still to be set
  return TG_OK;
}

int tg_get_state(tg_ctx* c, double* U, double* V) {
  if (!c || !U || !V || !c->d_U.p) {
    tg_set_error("tg_get_state: invalid argument");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  std::vector<double> h(2 * nn);
  TG_HIP(hipStreamSynchronize(c->stream));
  TG_HIP(hipMemcpy(h.data(), c->d_U.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(h.data() + nn, c->d_V.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
  // rows back to the reference single-particle order
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      U[c->sp_perm[i] + (size_t)j * n] = h[i + (size_t)j * n];
      V[c->sp_perm[i] + (size_t)j * n] = h[nn + i + (size_t)j * n];
    }
  return TG_OK;
}

}  // extern "C"
