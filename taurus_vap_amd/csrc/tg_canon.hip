// Canonical (BMZ) basis of the resident states on the device
// (construct_canonical_basis, src/module_wavefunctions.f90:777-945).
//
// For every state s of the context (real U, V in the internal single-particle
// order):
//   rho = V V^T, D = eigenvectors of rho (ascending eigenvalues; rocSOLVER
//   dsyevd, batched over the states), rhoc = D^T rho D,
//   v2_i = |rhoc_ii|: nocc counts v2 >= 1 - eps, nemp v2 <= eps (the
//   reference's seed_allemp = 0), ovac = sum log sqrt|1 - v2| over the rest,
//   kappac = D^T (V U^T) D, and the degenerate-subspace analysis of kappac
//   (:843-873): a subspace wider than one canonical pair needs the real Schur
//   form of its block (:877-900) and is refused here (TG_E_UNSUPPORTED); the
//   caller then takes the host route.  Otherwise voveru_i = sign(v, kappac(i,
//   i +- 1)) / u for the paired states, alternating partners as zVc (:921-941).
// D replaces the resident D0 and the state is set up as tg_set_states would,
// with U and V never leaving the device.
#include "tg_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <vector>

namespace {

constexpr int NT = 256;
inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

// diag_i = sum_j D(j, i) T(j, i) with T = rho D: the diagonal of D^T rho D
__global__ void diag_kernel(int n, long long nn, const double* __restrict__ D,
                            const double* __restrict__ T, double* __restrict__ diag) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= n) return;
  const double* d = D + s * nn + (long long)i * n;
  const double* t = T + s * nn + (long long)i * n;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc = fma(d[j], t[j], acc);
  diag[(long long)s * n + i] = acc;
}

// Column i of kappac in [lo_s, hi_s): reach = last row j in (i, hi_s) with
// |kappac(j, i)| > thr, else i (the reference's submax); the partner entries
// kappac(i, i+1) and kappac(i, i-1).
__global__ void reach_kernel(int n, long long nn, const double* __restrict__ K,
                             const int* __restrict__ lohi, double thr, int* __restrict__ reach,
                             double* __restrict__ kd) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= n) return;
  const double* k = K + s * nn;
  const int lo = lohi[2 * s], hi = lohi[2 * s + 1];
  int r = i;
  if (i >= lo && i < hi)
    for (int j = hi - 1; j > i; --j)
      if (fabs(k[j + (long long)i * n]) > thr) {
        r = j;
        break;
      }
  const long long o = (long long)s * n + i;
  reach[o] = r;
  kd[2 * o] = (i + 1 < n) ? k[i + (long long)(i + 1) * n] : 0.0;
  kd[2 * o + 1] = (i > 0) ? k[i + (long long)(i - 1) * n] : 0.0;
}

rocblas_handle handle_of(tg_ctx* c) {
  if (!c->rb) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    c->rb = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(c->rb);
  if (rocblas_set_stream(h, c->stream) != rocblas_status_success) return nullptr;
  return h;
}

// Subspace dimensions of the reference's analysis (:851-873), 0-based, on the
// columns [lo, hi); returns the largest one (odd ones above 2 count as 2, as the
// reference continues with them after its warning).
int widest_subspace(std::vector<int> sub, int lo, int hi) {
  int k = lo;
  while (k < hi - 1) {
    const int j = sub[k];
    int m = j;
    for (int q = k; q <= j; ++q) m = std::max(m, sub[q]);
    for (int q = k; q <= j; ++q) sub[q] = m;
    if (m == j) k = j + 1;
  }
  int widest = 0;
  k = lo;
  while (k < hi - 1) {
    int d = sub[k] - k + 1;
    if (d > 2 && d % 2 != 0) d = 2;
    widest = std::max(widest, d);
    k += d;
  }
  return widest;
}

}  // namespace

void tg_canon_release(tg_ctx* c) {
  if (c->rb) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c->rb));
  c->rb = nullptr;
}

extern "C" int tg_canonical_basis(tg_ctx* c, double eps, int* nocc_out, int* nemp_out,
                                  double* ovac_out, double* voveru_out) {
  if (!c || !c->d_U.p || !c->d_V.p || c->ns < 1 || c->Mphip < 1 || !(eps > 0.0 && eps < 0.5)) {
    tg_set_error("tg_canonical_basis: invalid argument (states set once with tg_set_states?)");
    return TG_E_ARG;
  }
  const int n = c->n, S = c->ns;
  const long long nn = (long long)n * n;
  // work: R (rho, then rho D), E (eigenvectors D), T1, T2, K (kappac), W, F, diag,
  // lohi, reach, kd, info
  const size_t mats = (size_t)5 * S * nn;
  const size_t vecs = (size_t)3 * S * n;
  const size_t need = sizeof(double) * (mats + vecs + 2 * (size_t)S * n) +
                      sizeof(int) * (2 * (size_t)S + (size_t)S * n + (size_t)S);
  TG_TRY(c->d_canon.ensure(need));
  TG_TRY(c->d_D.ensure(sizeof(double) * nn * S));
  double* R = c->d_canon.as<double>();
  double* E = R + S * nn;
  double* T1 = E + S * nn;
  double* T2 = T1 + S * nn;
  double* K = T2 + S * nn;
  double* W = K + S * nn;             // eigenvalues
  double* F = W + (size_t)S * n;      // dsyevd off-diagonal workspace
  double* dg = F + (size_t)S * n;     // diag(D^T rho D)
  double* kd = dg + (size_t)S * n;    // partner entries, 2 per column
  int* lohi = reinterpret_cast<int*>(kd + 2 * (size_t)S * n);
  int* reach = lohi + 2 * S;
  int* info = reach + (size_t)S * n;
  const double* U = c->d_U.as<double>();
  const double* V = c->d_V.as<double>();
  hipStream_t s = c->stream;

  // rho = V V^T; D from dsyevd on a copy (upper triangle, as dsyev('V','U'))
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, V, n, nn, 0.0, R, n, nn, S));
  TG_HIP(hipMemcpyAsync(E, R, sizeof(double) * nn * S, hipMemcpyDeviceToDevice, s));
  rocblas_handle h = handle_of(c);
  if (!h) {
    tg_set_error("tg_canonical_basis: rocBLAS handle");
    return TG_E_HIP;
  }
  if (rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n, E, n, nn,
                                       W, n, F, n, info, S) != rocblas_status_success) {
    tg_set_error("tg_canonical_basis: rocsolver_dsyevd failed");
    return TG_E_HIP;
  }
  // rhoc diagonal, kappac = D^T (V U^T) D
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, R, n, nn, E, n, nn, 0.0, T1, n, nn, S));
  hipLaunchKernelGGL(diag_kernel, dim3(nblk(n), S), dim3(NT), 0, s, n, nn, E, T1, dg);
  TG_HIP(hipGetLastError());
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, U, n, nn, 0.0, T1, n, nn, S));
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, T1, n, nn, E, n, nn, 0.0, T2, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, E, n, nn, T2, n, nn, 0.0, K, n, nn, S));

  std::vector<double> v2((size_t)S * n);
  std::vector<int> hinfo(S);
  TG_HIP(hipMemcpyAsync(v2.data(), dg, sizeof(double) * S * n, hipMemcpyDeviceToHost, s));
  TG_HIP(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * S, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < S; ++k)
    if (hinfo[k] != 0) {
      tg_set_error("tg_canonical_basis: dsyevd did not converge (state " + std::to_string(k) + ")");
      return TG_E_SINGULAR;
    }
  std::vector<int> nocc(S, 0), nemp(S, 0), hlohi(2 * S);
  std::vector<double> ovac(S, 0.0);
  for (int k = 0; k < S; ++k) {
    for (int i = 0; i < n; ++i) {
      const double x = fabs(v2[(size_t)k * n + i]);
      v2[(size_t)k * n + i] = x;
      if (x >= 1.0 - eps) ++nocc[k];
      else if (x <= eps) ++nemp[k];
      else ovac[k] += log(sqrt(fabs(1.0 - x)));
    }
    hlohi[2 * k] = nemp[k];
    hlohi[2 * k + 1] = n - nocc[k];
  }
  TG_HIP(hipMemcpyAsync(lohi, hlohi.data(), sizeof(int) * 2 * S, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(reach_kernel, dim3(nblk(n), S), dim3(NT), 0, s, n, nn, K, lohi,
                     sqrt(eps) * 1.0e-2, reach, kd);
  TG_HIP(hipGetLastError());
  std::vector<int> hreach((size_t)S * n);
  std::vector<double> hkd(2 * (size_t)S * n);
  TG_HIP(hipMemcpyAsync(hreach.data(), reach, sizeof(int) * S * n, hipMemcpyDeviceToHost, s));
  TG_HIP(hipMemcpyAsync(hkd.data(), kd, sizeof(double) * 2 * S * n, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < S; ++k) {
    std::vector<int> sub(hreach.begin() + (size_t)k * n, hreach.begin() + (size_t)(k + 1) * n);
    const int wide = widest_subspace(sub, nemp[k], n - nocc[k]);
    if (wide > 2) {
      tg_set_error("tg_canonical_basis: state " + std::to_string(k) + " has a degenerate kappa "
                   "subspace of dimension " + std::to_string(wide) +
                   " (needs its real Schur form): use the host canonical basis");
      return TG_E_UNSUPPORTED;
    }
  }
  // the new canonical data; D into the resident D
  std::vector<StateInfo> st(S);
  for (int k = 0; k < S; ++k) {
    StateInfo& si = st[k];
    si.nocc = nocc[k];
    si.nemp = nemp[k];
    si.ovac = (n - nocc[k] - nemp[k] == 0) ? 0.0 : ovac[k];
    si.voveru.assign(n - nocc[k], 0.0);
    int par = 0;
    for (int i = nemp[k]; i < n - nocc[k]; ++i) {
      const double x = v2[(size_t)k * n + i];
      const double u = sqrt(fabs(1.0 - x)), v = sqrt(x);
      const double partner = hkd[2 * ((size_t)k * n + i) + (par % 2 == 0 ? 0 : 1)];
      si.voveru[i] = copysign(v, partner) / u;
      ++par;
    }
  }
  TG_HIP(hipMemcpyAsync(c->d_D.p, E, sizeof(double) * nn * S, hipMemcpyDeviceToDevice, s));
  TG_HIP(hipStreamSynchronize(s));
  c->st = st;
  for (int k = 0; k < S; ++k) {
    if (nocc_out) nocc_out[k] = st[k].nocc;
    if (nemp_out) nemp_out[k] = st[k].nemp;
    if (ovac_out) ovac_out[k] = st[k].ovac;
    if (voveru_out)
      for (int i = 0; i < n; ++i)
        voveru_out[(size_t)k * n + i] = i < n - st[k].nocc ? st[k].voveru[i] : 0.0;
  }
  return tg_states_ready(c);
}
