// Canonical (BMZ) basis of the resident states on the device
// (construct_canonical_basis, src/module_wavefunctions.f90:777-945).
//
// For every state s of the context (real U, V in the internal single-particle
// order):
//   rho = V V^T, D = eigenvectors of rho (ascending eigenvalues, upper triangle
//   read as DSYEV('V','U') of :806; batched_sym_eigensystem: Householder
//   tridiagonalization, divide and conquer, back-transformation, all on the
//   device), rhoc = D^T rho D,
//   v2_i = |rhoc_ii|: nocc counts v2 >= 1 - eps, nemp v2 <= eps (the
//   reference's seed_allemp = 0), ovac = sum log sqrt|1 - v2| over the rest,
//   kappac = D^T (V U^T) D, and the degenerate-subspace analysis of kappac
//   (:843-873).  A subspace wider than one canonical pair (even dimension m > 2)
//   is brought to canonical form as the reference's real Schur step (:877-900)
//   does, by an orthogonal T of the block: T pairs each eigenvector x of the
//   symmetric K^T K (K the real skew block) with K x / |K x|, so T^T K T is
//   block diagonal with 2 x 2 skew blocks; D's columns of the block become D T.
//   Then voveru_i = sign(v, kappac(i, i +- 1)) / u for the paired states,
//   alternating partners as zVc (:921-941).
// D replaces the resident D0 and the state is set up as tg_set_states would,
// with U and V never leaving the device.  The m x m blocks (m <= 64) are small
// dense problems solved on the host, as the dim x dim multiplier systems of the
// device iteration are.
#include "tg_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int NT = 256;
constexpr int MAX_BLOCK = 64;   // widest degenerate subspace taken on
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

// D(:, k0:k0+m) <- D(:, k0:k0+m) T, T m x m column-major (m <= MAX_BLOCK); one
// thread per row of D
__global__ void block_rotate_kernel(int n, int k0, int m, double* __restrict__ D,
                                    const double* __restrict__ T) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r >= n) return;
  double d[MAX_BLOCK];
  for (int q = 0; q < m; ++q) d[q] = D[r + (long long)(k0 + q) * n];
  for (int c = 0; c < m; ++c) {
    double acc = 0.0;
    for (int q = 0; q < m; ++q) acc = fma(d[q], T[q + c * m], acc);
    D[r + (long long)(k0 + c) * n] = acc;
  }
}

// Subspaces of the reference's analysis (:851-873), 0-based, on the columns
// [lo, hi): (start, dimension) of every subspace, odd ones above 2 counted as
// 2 (the reference continues with them after its warning).
std::vector<std::pair<int, int>> subspaces(std::vector<int> sub, int lo, int hi) {
  int k = lo;
  while (k < hi - 1) {
    const int j = sub[k];
    int m = j;
    for (int q = k; q <= j; ++q) m = std::max(m, sub[q]);
    for (int q = k; q <= j; ++q) sub[q] = m;
    if (m == j) k = j + 1;
  }
  std::vector<std::pair<int, int>> out;
  k = lo;
  while (k < hi - 1) {
    int d = sub[k] - k + 1;
    if (d > 2 && d % 2 != 0) d = 2;
    out.push_back({k, d});
    k += d;
  }
  return out;
}

// Cyclic Jacobi on the symmetric m x m S (column-major, overwritten): returns
// the eigenvectors (columns of Q) and eigenvalues in w.
void jacobi_eigen(int m, std::vector<double>& S, std::vector<double>& Q, std::vector<double>& w) {
  Q.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) Q[i + (size_t)i * m] = 1.0;
  auto at = [&](int i, int j) -> double& { return S[i + (size_t)j * m]; };
  double norm = 0.0;
  for (double x : S) norm += x * x;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < j; ++i) off += at(i, j) * at(i, j);
    if (off <= 1e-30 * norm) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = at(p, q);
        if (apq == 0.0) continue;
        const double theta = (at(q, q) - at(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < m; ++k) {   // S <- S J (columns p, q)
          const double a = at(k, p), b = at(k, q);
          at(k, p) = cs * a - sn * b;
          at(k, q) = sn * a + cs * b;
        }
        for (int k = 0; k < m; ++k) {   // S <- J^T S (rows p, q)
          const double a = at(p, k), b = at(q, k);
          at(p, k) = cs * a - sn * b;
          at(q, k) = sn * a + cs * b;
        }
        for (int k = 0; k < m; ++k) {   // Q <- Q J
          const double a = Q[k + (size_t)p * m], b = Q[k + (size_t)q * m];
          Q[k + (size_t)p * m] = cs * a - sn * b;
          Q[k + (size_t)q * m] = sn * a + cs * b;
        }
      }
  }
  w.resize(m);
  for (int i = 0; i < m; ++i) w[i] = at(i, i);
}

// Orthogonal T (m x m, m even) with T^T K T block diagonal in 2 x 2 skew blocks,
// for the real skew K (column-major): eigenvectors x of K^T K taken in order of
// their eigenvalue, each not yet spanned paired with K x / |K x| (K^T K maps
// both to the same eigenvalue, and x^T K x = 0).  False if a pair degenerates.
bool skew_canonical(int m, const std::vector<double>& K, std::vector<double>& T) {
  std::vector<double> S((size_t)m * m, 0.0), Q, w;
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int q = 0; q < m; ++q) acc += K[q + (size_t)i * m] * K[q + (size_t)j * m];
      S[i + (size_t)j * m] = acc;
    }
  jacobi_eigen(m, S, Q, w);
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] < w[b]; });
  T.assign((size_t)m * m, 0.0);
  int cols = 0;
  auto orth = [&](std::vector<double>& x) {
    for (int pass = 0; pass < 2; ++pass)
      for (int c = 0; c < cols; ++c) {
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += T[i + (size_t)c * m] * x[i];
        for (int i = 0; i < m; ++i) x[i] -= d * T[i + (size_t)c * m];
      }
    double s = 0.0;
    for (double v : x) s += v * v;
    return sqrt(s);
  };
  for (int e = 0; e < m && cols < m; ++e) {
    std::vector<double> x(Q.begin() + (size_t)ord[e] * m, Q.begin() + (size_t)(ord[e] + 1) * m);
    const double nx = orth(x);
    if (nx < 0.5) continue;   // already spanned by the pairs taken
    for (double& v : x) v /= nx;
    std::vector<double> y(m, 0.0);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) y[i] += K[i + (size_t)j * m] * x[j];
    for (int i = 0; i < m; ++i) T[i + (size_t)cols * m] = x[i];
    ++cols;
    const double ny = orth(y);
    if (!(ny > 0.0)) return false;
    for (int i = 0; i < m; ++i) T[i + (size_t)cols * m] = y[i] / ny;
    ++cols;
  }
  return cols == m;
}

}  // namespace

extern "C" int tg_canonical_basis(tg_ctx* c, double eps, int* nocc_out, int* nemp_out,
                                  double* ovac_out, double* voveru_out) {
  if (!c) {
    tg_set_error("tg_canonical_basis: invalid argument");
    return TG_E_ARG;
  }
  return canonical_basis_on(c, c->stream, eps, nocc_out, nemp_out, ovac_out, voveru_out);
}

namespace {

// Work areas of the canonical basis in c->d_canon: R (rho, then rho D), E
// (eigenvectors D), T1, T2, K (kappac), W, diag, kd, then ints lohi, reach,
// info, then the block transformations, then the eigen-solver's workspace.
struct CanonWork {
  double *R, *E, *T1, *T2, *K, *W, *dg, *kd, *Tdev;
  int *lohi, *reach, *info;
  void* eig;
};

int canon_work(tg_ctx* c, CanonWork& w) {
  const int n = c->n, S = c->ns;
  const long long nn = (long long)n * n;
  const size_t mats = (size_t)5 * S * nn;
  const size_t vecs = (size_t)2 * S * n;
  const size_t ints = 2 * (size_t)S + (size_t)S * n + (size_t)S;
  const size_t base = sizeof(double) * (mats + vecs + 2 * (size_t)S * n) +
                      sizeof(int) * (ints + 2);   // pad: the block area stays 8-byte aligned
  const size_t tmax = (size_t)n * MAX_BLOCK;     // block transformations of one state
  const size_t eoff = ((base + 8 + sizeof(double) * tmax + 255) / 256) * 256;
  TG_TRY(c->d_canon.ensure(eoff + sym_eigensystem_work_bytes(n, S)));
  TG_TRY(c->d_D.ensure(sizeof(double) * nn * S));
  w.R = c->d_canon.as<double>();
  w.E = w.R + S * nn;
  w.T1 = w.E + S * nn;
  w.T2 = w.T1 + S * nn;
  w.K = w.T2 + S * nn;
  w.W = w.K + S * nn;                    // eigenvalues
  w.dg = w.W + (size_t)S * n;            // diag(D^T rho D)
  w.kd = w.dg + (size_t)S * n;           // partner entries, 2 per column
  w.lohi = reinterpret_cast<int*>(w.kd + 2 * (size_t)S * n);
  w.reach = w.lohi + 2 * S;
  w.info = w.reach + (size_t)S * n;
  w.Tdev = reinterpret_cast<double*>(reinterpret_cast<char*>(c->d_canon.p) + ((base + 7) / 8) * 8);
  w.eig = static_cast<char*>(c->d_canon.p) + eoff;
  return TG_OK;
}

}  // namespace

// The device part of the canonical basis up to its first host read, on stream
// s: rho, its eigensystem, diag(D^T rho D) and kappac = D^T (V U^T) D.
int canonical_launch(tg_ctx* c, hipStream_t s, double eps) {
  if (!c || !c->d_U.p || !c->d_V.p || c->ns < 1 || c->Mphip < 1 || !(eps > 0.0 && eps < 0.5)) {
    tg_set_error("tg_canonical_basis: invalid argument (states set once with tg_set_states?)");
    return TG_E_ARG;
  }
  const int n = c->n, S = c->ns;
  const long long nn = (long long)n * n;
  CanonWork w;
  TG_TRY(canon_work(c, w));
  const double* U = c->d_U.as<double>();
  const double* V = c->d_V.as<double>();
  // rho = V V^T; D its eigenvectors (upper triangle read, as dsyev('V','U'))
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, V, n, nn, 0.0, w.R, n, nn, S));
  TG_TRY(batched_sym_eigensystem(s, n, w.R, n, nn, S, w.eig, w.W, n, w.E, nn, w.info));
  // rhoc diagonal, kappac = D^T (V U^T) D
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, w.R, n, nn, w.E, n, nn, 0.0, w.T1, n, nn, S));
  hipLaunchKernelGGL(diag_kernel, dim3(nblk(n), S), dim3(NT), 0, s, n, nn, w.E, w.T1, w.dg);
  TG_HIP(hipGetLastError());
  TG_TRY(gemm_dd(s, OP_N, OP_T, n, n, n, 1.0, V, n, nn, U, n, nn, 0.0, w.T1, n, nn, S));
  TG_TRY(gemm_dd(s, OP_N, OP_N, n, n, n, 1.0, w.T1, n, nn, w.E, n, nn, 0.0, w.T2, n, nn, S));
  TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, n, 1.0, w.E, n, nn, w.T2, n, nn, 0.0, w.K, n, nn, S));
  return TG_OK;
}

int canonical_basis_on(tg_ctx* c, hipStream_t s, double eps, int* nocc_out, int* nemp_out,
                       double* ovac_out, double* voveru_out) {
  TG_TRY(canonical_launch(c, s, eps));
  return canonical_finish(c, s, eps, nocc_out, nemp_out, ovac_out, voveru_out);
}

// The rest, after canonical_launch on the same stream s: occupations, the
// degenerate blocks' canonical form (host), partner entries and v/u; the new D
// into the resident D.  The host waits on stream s only.
int canonical_finish(tg_ctx* c, hipStream_t s, double eps, int* nocc_out, int* nemp_out,
                     double* ovac_out, double* voveru_out) {
  const int n = c->n, S = c->ns;
  const long long nn = (long long)n * n;
  CanonWork w;
  TG_TRY(canon_work(c, w));
  double* E = w.E;
  double* K = w.K;
  double* dg = w.dg;
  double* kd = w.kd;
  int* lohi = w.lohi;
  int* reach = w.reach;
  int* info = w.info;
  double* Tdev = w.Tdev;
  std::vector<double> v2((size_t)S * n);
  std::vector<int> hinfo(S);
  TG_HIP(hipMemcpyAsync(v2.data(), dg, sizeof(double) * S * n, hipMemcpyDeviceToHost, s));
  TG_HIP(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * S, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < S; ++k)
    if (hinfo[k] != 0) {
      tg_set_error("tg_canonical_basis: eigen-solver did not converge (state " + std::to_string(k) + ")");
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

  // degenerate subspaces wider than one pair: canonical form of their kappac
  // block, D's columns rotated on the device, the partner entries from T^T K T
  for (int k = 0; k < S; ++k) {
    std::vector<int> sub(hreach.begin() + (size_t)k * n, hreach.begin() + (size_t)(k + 1) * n);
    std::vector<std::pair<int, int>> blocks;
    size_t tsum = 0;
    for (const auto& b : subspaces(sub, nemp[k], n - nocc[k]))
      if (b.second > 2) {
        if (b.second > MAX_BLOCK) {
          tg_set_error("tg_canonical_basis: state " + std::to_string(k) + " has a degenerate kappa "
                       "subspace of dimension " + std::to_string(b.second) + " (above " +
                       std::to_string(MAX_BLOCK) + "): use the host canonical basis");
          return TG_E_UNSUPPORTED;
        }
        blocks.push_back(b);
        tsum += (size_t)b.second * b.second;
      }
    if (blocks.empty()) continue;
    std::vector<double> Tall;
    Tall.reserve(tsum);
    for (const auto& b : blocks) {
      const int k0 = b.first, m = b.second;
      std::vector<double> Kb((size_t)m * m), Tb;
      TG_HIP(hipMemcpy2DAsync(Kb.data(), sizeof(double) * m, K + (size_t)k * nn + k0 + (size_t)k0 * n,
                              sizeof(double) * n, sizeof(double) * m, m, hipMemcpyDeviceToHost, s));
      TG_HIP(hipStreamSynchronize(s));
      for (int j = 0; j < m; ++j)   // the block is skew up to rounding: take its skew part
        for (int i = 0; i < j; ++i) {
          const double a = 0.5 * (Kb[i + (size_t)j * m] - Kb[j + (size_t)i * m]);
          Kb[i + (size_t)j * m] = a;
          Kb[j + (size_t)i * m] = -a;
        }
      for (int i = 0; i < m; ++i) Kb[i + (size_t)i * m] = 0.0;
      if (!skew_canonical(m, Kb, Tb)) {
        tg_set_error("tg_canonical_basis: degenerate kappa block of state " + std::to_string(k) +
                     " has no canonical pairing: use the host canonical basis");
        return TG_E_UNSUPPORTED;
      }
      // partner entries of T^T K T
      for (int p = 0; p < m; ++p) {
        auto kc = [&](int r, int q) {
          double acc = 0.0;
          for (int a = 0; a < m; ++a) {
            double t = 0.0;
            for (int bb = 0; bb < m; ++bb) t += Kb[a + (size_t)bb * m] * Tb[bb + (size_t)q * m];
            acc += Tb[a + (size_t)r * m] * t;
          }
          return acc;
        };
        const size_t o = (size_t)k * n + k0 + p;
        hkd[2 * o] = (p + 1 < m) ? kc(p, p + 1) : 0.0;
        hkd[2 * o + 1] = (p > 0) ? kc(p, p - 1) : 0.0;
      }
      Tall.insert(Tall.end(), Tb.begin(), Tb.end());
    }
    TG_HIP(hipMemcpyAsync(Tdev, Tall.data(), sizeof(double) * Tall.size(), hipMemcpyHostToDevice, s));
    size_t off = 0;
    for (const auto& b : blocks) {
      hipLaunchKernelGGL(block_rotate_kernel, dim3(nblk(n)), dim3(NT), 0, s, n, b.first, b.second,
                         E + (size_t)k * nn, Tdev + off);
      off += (size_t)b.second * b.second;
    }
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(s));   // Tdev is reused by the next state
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
