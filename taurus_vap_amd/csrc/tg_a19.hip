// diagonalize_hsp_and_H11(opt = 0) on the device (src/module_gradient.f90:158-453)
// and the eigenvalue solver behind it.
//
// For every resident state (U, V real):
//   rho = V V^T, kappa = V U^T                       (calculate_densities_real)
//   h, Delta of calculate_fields_diag(rho, kappa)   (the block contraction, :181)
//   H11 = U^T h U - V^T h^T V + fd (U^T Delta V + (U^T Delta V)^T)
//                                                    (calculate_H11_real,
//                                                     module_fields.f90:327-353)
//   H11 -= sum_m lambda_m C11_m  (one-body constraints, 'f11' of
//          set_operator_qpbasis: 1/2 (A1 + A1^T), A1 = U^T O U - V^T O^T V);
//   h   -= sum_m lambda_m O_m    (every constraint, :194-212)
// and the eigenvalues of h and H11 (the two DSYEV('V','U') of :297 and :416,
// upper triangles read), ascending.  The gradient recipes of calculate_gradient
// (:97-120) turn the extreme eigenvalues into the step parameters eta, mu.
//
// Eigenvalues: one workgroup per symmetric matrix.  Householder reduction to
// tridiagonal form (the reflector of each column from its sub-diagonal part,
// p = tau A v, w = p - (tau/2)(p.v) v, A -= v w^T + w v^T), with the matrix in
// LDS up to order EIG_NL and in global memory (L2) beyond; then every
// eigenvalue by bisection on Sturm sequence counts of the tridiagonal matrix,
// one thread per eigenvalue index, to the resolution of the floating-point
// numbers.
#include <algorithm>
#include <cmath>
#include <cfloat>
#include <vector>

#include "tg_internal.h"

namespace {

constexpr int NT = 256;
constexpr int EIG_TH = 256;
constexpr int EIG_NL = 112;          // matrices up to this order stay in LDS

inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int TH>
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < TH / 64; ++q) s += red[q];
  return s;
}

// number of eigenvalues of the tridiagonal (d, e2 = e^2) below x
__device__ inline int sturm_count(int n, const double* d, const double* e2, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (q <= pivmin) { ++cnt; q = fmin(q, -pivmin); }
  for (int j = 1; j < n; ++j) {
    q = d[j] - x - e2[j - 1] / q;
    if (q <= pivmin) { ++cnt; q = fmin(q, -pivmin); }
  }
  return cnt;
}

// M: n x n symmetric matrix (leading dimension ldm; LDS or global), destroyed.
// d, e2, v, p: LDS vectors of n.  Writes the n eigenvalues, ascending, to w.
template <int TH>
__device__ void sym_eigvals(int n, double* M, int ldm, double* d, double* e2, double* v, double* p,
                            double* red, double* __restrict__ w) {
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  constexpr int NW = TH / 64;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double s2 = 0.0;
    for (int i = t; i < m; i += TH) {
      const double x = M[(k + 1 + i) + (long long)k * ldm];
      v[i] = x;
      if (i > 0) s2 = fma(x, x, s2);
    }
    s2 = block_sum<TH>(s2, red);      // (barrier inside: v visible)
    const double alpha = v[0];
    if (t == 0) d[k] = M[k + (long long)k * ldm];
    if (s2 == 0.0) {
      if (t == 0) e2[k] = alpha * alpha;
      __syncthreads();
      continue;
    }
    const double beta = -copysign(sqrt(fma(alpha, alpha, s2)), alpha);
    const double tau = (beta - alpha) / beta;
    const double scale = 1.0 / (alpha - beta);
    if (t == 0) e2[k] = beta * beta;
    __syncthreads();
    for (int i = t; i < m; i += TH) v[i] = (i == 0) ? 1.0 : v[i] * scale;
    __syncthreads();
    // p = tau A22 v (A22 symmetric: column j dotted with v, one wave per column)
    for (int j = wid; j < m; j += NW) {
      const double* col = M + (k + 1) + (long long)(k + 1 + j) * ldm;
      double acc = 0.0;
      for (int i = lane; i < m; i += 64) acc = fma(col[i], v[i], acc);
      acc = wave_sum(acc);
      if (lane == 0) p[j] = tau * acc;
    }
    __syncthreads();
    double dot = 0.0;
    for (int i = t; i < m; i += TH) dot = fma(p[i], v[i], dot);
    dot = block_sum<TH>(dot, red);
    const double K = 0.5 * tau * dot;
    for (int i = t; i < m; i += TH) p[i] = fma(-K, v[i], p[i]);
    __syncthreads();
    // A22 -= v w^T + w v^T
    const long long mm = (long long)m * m;
    for (long long idx = t; idx < mm; idx += TH) {
      const int i = (int)(idx % m), j = (int)(idx / m);
      double* a = M + (k + 1 + i) + (long long)(k + 1 + j) * ldm;
      *a = fma(-v[i], p[j], fma(-p[i], v[j], *a));
    }
    __syncthreads();
  }
  if (t == 0) {
    if (n >= 2) {
      d[n - 2] = M[(n - 2) + (long long)(n - 2) * ldm];
      const double en = M[(n - 1) + (long long)(n - 2) * ldm];
      e2[n - 2] = en * en;
    }
    d[n - 1] = M[(n - 1) + (long long)(n - 1) * ldm];
  }
  __syncthreads();
  // Gershgorin interval and the pivot guard of the Sturm counts
  double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double el = (i > 0) ? sqrt(e2[i - 1]) : 0.0, er = (i + 1 < n) ? sqrt(e2[i]) : 0.0;
    gl = fmin(gl, d[i] - el - er);
    gu = fmax(gu, d[i] + el + er);
    if (i + 1 < n) emax2 = fmax(emax2, e2[i]);
  }
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double span = fmax(fabs(gl), fabs(gu));
  gl -= 2.0 * DBL_EPSILON * span * n + pivmin;
  gu += 2.0 * DBL_EPSILON * span * n + pivmin;
  for (int i = t; i < n; i += TH) {
    double lo = gl, hi = gu;
    for (int it = 0; it < 256; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (!(mid > lo && mid < hi)) break;
      if (sturm_count(n, d, e2, mid, pivmin) > i) hi = mid;
      else lo = mid;
    }
    w[i] = 0.5 * (lo + hi);
  }
}

// count matrices: A + b * sA (upper triangle read, ld lda), work + b * n * n
// (global work for n > EIG_NL), eigenvalues at w + b * ws.
template <int TH, bool LDSM>
__global__ void __launch_bounds__(TH)
sym_eigvals_kernel(int n, const double* __restrict__ A0, int lda, long long sA,
                   double* __restrict__ work, double* __restrict__ w0, long long ws) {
  __shared__ double Ml[LDSM ? EIG_NL * EIG_NL : 1];
  __shared__ double vecs[4 * (LDSM ? EIG_NL : 512)];
  __shared__ double red[TH / 64];
  const double* A = A0 + blockIdx.x * sA;
  double* M = LDSM ? Ml : work + (long long)blockIdx.x * n * n;
  const long long nn = (long long)n * n;
  for (long long idx = threadIdx.x; idx < nn; idx += TH) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    M[idx] = (i <= j) ? A[i + (long long)j * lda] : A[j + (long long)i * lda];
  }
  __syncthreads();
  const int cap = LDSM ? EIG_NL : 512;
  sym_eigvals<TH>(n, M, n, vecs, vecs + cap, vecs + 2 * cap, vecs + 3 * cap, red,
                  w0 + blockIdx.x * ws);
}

// ---- pieces of H11 and of the constraint terms ----------------------------

__global__ void real_to_cplx_kernel(long long len, const double* __restrict__ x, cplx* __restrict__ z) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < len) z[i] = cmk(x[i], 0.0);
}

// h = Re(hsp) - sum_m lambda_m O_m ; Dl = Re(DeltaLR)   (state blockIdx.y)
__global__ void h_delta_kernel(long long nn, const cplx* __restrict__ hsp0, const cplx* __restrict__ dl0,
                               int ncons, const double* __restrict__ lam0, const double* __restrict__ O,
                               double* __restrict__ h0, double* __restrict__ D0) {
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= nn) return;
  const long long st = blockIdx.y;
  const double* lam = lam0 + st * TG_MAX_CONS;
  double x = hsp0[st * nn + i].x;
  for (int m = 0; m < ncons; ++m) x = fma(-lam[m], O[(long long)m * nn + i], x);
  h0[st * nn + i] = x;
  D0[st * nn + i] = dl0[st * nn + i].x;
}

// H11 = A2 - A3 + fd (A4 + A4^T) (calculate_H11_real), 32 x 32 LDS tiles
__global__ void __launch_bounds__(256)
h11_combine_kernel(int n, double fd, const double* __restrict__ A20, const double* __restrict__ A30,
                   const double* __restrict__ A40, double* __restrict__ H0) {
  __shared__ double tb[32][33];
  const long long nn = (long long)n * n, o = blockIdx.z * nn;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i2 = j0 + tx, j2 = i0 + r;      // A4 at (j, i), staged transposed
    tb[r][tx] = (i2 < n && j2 < n) ? A40[o + i2 + (long long)j2 * n] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < n && j < n) {
      const long long q = o + i + (long long)j * n;
      H0[q] = A20[q] - A30[q] + (A40[q] + tb[tx][r]) * fd;
    }
  }
}

// H11 -= lambda_m / 2 (A1 + A1^T), A1 = T1 - T2 (the f11 part of a one-body
// constraint, set_operator_qpbasis 'f11')
__global__ void __launch_bounds__(256)
sub_f11_kernel(int n, int m, const double* __restrict__ lam0, const double* __restrict__ T10,
               const double* __restrict__ T20, double* __restrict__ H0) {
  __shared__ double tb[32][33];
  const long long nn = (long long)n * n, o = blockIdx.z * nn;
  const double lam = lam0[blockIdx.z * TG_MAX_CONS + m];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i2 = j0 + tx, j2 = i0 + r;
    tb[r][tx] = (i2 < n && j2 < n) ? T10[o + i2 + (long long)j2 * n] - T20[o + i2 + (long long)j2 * n] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < n && j < n) {
      const long long q = o + i + (long long)j * n;
      const double a1 = T10[q] - T20[q];
      H0[q] -= lam * (0.5 * (a1 + tb[tx][r]));
    }
  }
}

}  // namespace

// batched eigenvalues of symmetric matrices (upper triangles read); work: count
// n*n doubles when n > EIG_NL (else unused, may be null)
int batched_sym_eigvals(hipStream_t s, int n, const double* A, int lda, long long sA, int count,
                        double* work, double* w, long long ws) {
  if (count <= 0 || n <= 0) return TG_OK;
  if (n > 512) {
    tg_set_error("batched_sym_eigvals: order above 512");
    return TG_E_UNSUPPORTED;
  }
  if (n <= EIG_NL)
    hipLaunchKernelGGL((sym_eigvals_kernel<EIG_TH, true>), dim3(count), dim3(EIG_TH), 0, s, n, A,
                       lda, sA, work, w, ws);
  else
    hipLaunchKernelGGL((sym_eigvals_kernel<EIG_TH, false>), dim3(count), dim3(EIG_TH), 0, s, n, A,
                       lda, sA, work, w, ws);
  TG_HIP(hipGetLastError());
  return TG_OK;
}

// Step parameters of calculate_gradient for gradient_type 1 (from the H11
// spectrum) and 2 (from the h spectrum), module_gradient.f90:97-120; type 0 or
// a failed diagonalisation keep the input eta, mu.
void gradient_recipe(int type, int n, const double* eig_hsp, const double* eig_h11, double eta0,
                     double mu0, double* eta, double* mu) {
  *eta = eta0;
  *mu = mu0;
  if (type != 1 && type != 2) return;
  const double* e = (type == 1) ? eig_h11 : eig_hsp;
  double amax = 0.0, amin = INFINITY;
  for (int i = 0; i < n; ++i) {
    amax = std::max(amax, fabs(e[i]));
    amin = std::min(amin, fabs(e[i]));
  }
  const double emax = (type == 1 ? 4.0 : 2.0) * amax;
  const double emin = 2.0 * amin;
  const double cond = emax / emin;
  const double se = sqrt(emax) + sqrt(emin);
  *eta = (2.0 / se) * (2.0 / se);
  const double r = (sqrt(cond) - 1.0) / (sqrt(cond) + 1.0);
  *mu = r * r;
}

// h and H11 of every resident state (with the constraint terms of the
// multipliers lam, S x TG_MAX_CONS on the device) and their eigenvalues, to
// the host: eig (S x 2 x n: h then H11 per state).  hmat / h11 (device, S n^2
// each, internal order) are left with the two matrices (may be null).
int tg_hsp_h11_device(tg_ctx* c, const double* lam, std::vector<double>& eig, double** hmat,
                      double** h11) {
  const int n = c->n, S = c->ns, dim = c->cons_dim, pair = c->cons_pair;
  const long long nn = (long long)n * n;
  hipStream_t s = c->stream;
  const double* U = c->d_U.as<double>();
  const double* V = c->d_V.as<double>();
  // complex: rho, kappa, Gamma, DeltaLR, DeltaRL, h (S nn each); real: h, Delta,
  // H11, X, A2, A3 (S nn each), eigen work (2 S nn), eigenvalues (2 S n)
  const size_t cb = sizeof(cplx) * 6 * S * nn;
  const size_t rb = sizeof(double) * (8 * (size_t)S * nn + 2 * (size_t)S * n);
  TG_TRY(c->d_a19.ensure(cb + rb));
  cplx* zr = c->d_a19.as<cplx>();
  cplx* zk = zr + S * nn;
  cplx* zG = zk + S * nn;
  cplx* zdl = zG + S * nn;
  cplx* zdr = zdl + S * nn;
  cplx* zh = zdr + S * nn;
  double* rh = reinterpret_cast<double*>(zh + S * nn);
  double* rH = rh + S * nn;          // h then H11 of every state: one eigen launch
  double* rD = rH + S * nn;
  double* X = rD + S * nn;
  double* A2 = X + S * nn;
  double* A3 = A2 + S * nn;
  double* ew = A3 + S * nn;          // 2 S nn
  double* ev = ew + 2 * S * nn;      // 2 S n
  // rho = V V^T, kappa = V U^T, as complex inputs of the contraction
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, V, n, nn, 0.0, X, n, nn, S));
  hipLaunchKernelGGL(real_to_cplx_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s, S * nn, X, zr);
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, U, n, nn, 0.0, X, n, nn, S));
  hipLaunchKernelGGL(real_to_cplx_kernel, dim3(nblk(S * nn)), dim3(NT), 0, s, S * nn, X, zk);
  TG_HIP(hipGetLastError());
  // calculate_fields_diag: Gamma, Delta, h = Gamma + H1
  TG_TRY(contraction_run(c, S, zr, zk, nullptr, zG, zdl, zdr, 1, c->d_H1.as<cplx>(), zh));
  // h and Delta (real parts), h minus the constraint operators
  hipLaunchKernelGGL(h_delta_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, zh, zdl, dim, lam,
                     c->d_cHO.as<double>(), rh, rD);
  TG_HIP(hipGetLastError());
  // H11 = U^T h U - V^T h^T V + fd (U^T D V + (U^T D V)^T): the reference takes
  // h with the constraint terms removed only after H11, so rebuild h + lam O
  // here from Re(hsp) -- H11 uses the field before the constraint terms.
  hipLaunchKernelGGL(h_delta_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, zh, zdl, 0, lam,
                     c->d_cHO.as<double>(), A3, rD);
  TG_HIP(hipGetLastError());
  double* hf = A3;                       // h of the fields (no constraint terms)
  double* A4 = ew;                       // scratch before the eigen work
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, hf, n, nn, U, n, nn, 0.0, X, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, U, n, nn, X, n, nn, 0.0, A2, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, hf, n, nn, V, n, nn, 0.0, X, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, V, n, nn, X, n, nn, 0.0, A4 + S * nn, n, nn, S));
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, rD, n, nn, V, n, nn, 0.0, X, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, U, n, nn, X, n, nn, 0.0, A4, n, nn, S));
  const dim3 tiles((n + 31) / 32, (n + 31) / 32, S);
  hipLaunchKernelGGL(h11_combine_kernel, tiles, dim3(256), 0, s, n, c->factor_delta, A2,
                     A4 + S * nn, A4, rH);
  TG_HIP(hipGetLastError());
  // one-body constraints: H11 -= lambda_m f11(O_m)
  for (int m = 0; m < dim - pair; ++m) {
    const double* O = c->d_cHO.as<double>() + (size_t)m * nn;
    TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, O, n, 0, U, n, nn, 0.0, X, n, nn, S));
    TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, U, n, nn, X, n, nn, 0.0, A2, n, nn, S));
    TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, O, n, 0, V, n, nn, 0.0, X, n, nn, S));
    TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, V, n, nn, X, n, nn, 0.0, A3, n, nn, S));
    hipLaunchKernelGGL(sub_f11_kernel, tiles, dim3(256), 0, s, n, m, lam, A2, A3, rH);
    TG_HIP(hipGetLastError());
  }
  // eigenvalues of h (with the constraint terms) and of H11, all states in one
  // launch: matrix b < S is h of state b, matrix S + b its H11
  TG_TRY(batched_sym_eigvals(s, n, rh, n, nn, 2 * S, ew, ev, n));
  std::vector<double> hv(2 * (size_t)S * n);
  TG_HIP(hipMemcpyAsync(hv.data(), ev, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  eig.assign(2 * (size_t)S * n, 0.0);
  for (int st = 0; st < S; ++st) {
    std::copy(hv.begin() + (size_t)st * n, hv.begin() + (size_t)(st + 1) * n,
              eig.begin() + 2 * (size_t)st * n);
    std::copy(hv.begin() + (size_t)(S + st) * n, hv.begin() + (size_t)(S + st + 1) * n,
              eig.begin() + (2 * (size_t)st + 1) * n);
  }
  if (hmat) *hmat = rh;
  if (h11) *h11 = rH;
  return TG_OK;
}
