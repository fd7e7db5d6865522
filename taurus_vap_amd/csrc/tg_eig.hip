// Symmetric eigen-decomposition on the device: eigenvalues (ascending) and
// eigenvectors of a batch of real symmetric matrices, the upper triangle read, as
// DSYEV('V','U') gives them (construct_canonical_basis,
// src/module_wavefunctions.f90:806).
//
//   1. Householder reduction to tridiagonal form, one workgroup per matrix, in
//      panels of 32 columns: per column the reflector of its sub-diagonal part and
//      w = tau (A22 v - V (W^T v) - W (V^T v)) - (tau/2)(w.v) v with A22 the
//      trailing matrix as stored at the start of the panel; after the panel one
//      rank-64 update A22 -= V W^T + W V^T.  The trailing matrix is read once per
//      column instead of three times.
//   2. The tridiagonal eigenproblem by divide and conquer.  Every off-diagonal
//      element e_i tears the matrix into two blocks plus |e_i| v v^T
//      (v = e_i + sign(e_i) e_{i+1}), down to 1 x 1 blocks; the blocks are merged
//      bottom-up, all merges of one height at once, one workgroup each:
//        * the coupling vector z from the last / first rows of the children's
//          eigenvectors, the children's eigenvalues sorted;
//        * deflation: components with rho |z_i| below 8 eps max(|d|, |z|) keep
//          their eigenpair, and a close pair d_p, d_i is rotated (Givens) so that
//          one component vanishes when c s (d_i - d_p) is below the same bound;
//        * the roots of the secular equation 1 + rho sum z_i^2 / (d_i - l) = 0,
//          each in its interval, in coordinates of the nearer pole (l = d_o + tau,
//          the differences d_i - d_o formed once) by a two-pole rational model
//          iteration safeguarded by the bracket;
//        * z recomputed from the roots (Loewner's formula), so that the secular
//          eigenvectors z_i / (d_i - l_j) are orthogonal to working precision;
//        * the merged eigenvectors: the children's, gathered and rotated, times the
//          secular ones (in the workgroup for small merges, on the matrix cores
//          for the wide ones), their columns in ascending order of eigenvalue.
//   3. The back-transformation by the reflectors in blocks of 32 as
//      I - V T V^T (T upper triangular from V^T V and tau), each workgroup
//      carrying 32 (or 16) columns of the eigenvector matrix, in LDS, through
//      every block.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "tg_internal.h"

namespace {

constexpr int EG_TH = 256;      // threads of every kernel here
constexpr int EG_NB = 32;       // panel width of the reduction / reflector block
constexpr int EG_NMAX = 1024;   // largest order
constexpr int EG_SMALL = 64;    // merges up to this order multiply in their workgroup

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
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

template <int TH>
__device__ inline double block_max(double v, double* red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < TH / 64; ++q) s = fmax(s, red[q]);
  return s;
}

// ---------------------------------------------------------------------------
// 1. Householder tridiagonalization (lower), one workgroup per matrix.
// A: n x n working copy (ld n), Vm: reflectors by column (v_i(i+1) = 1, zero
// above), Wm: the panel's w vectors (n x EG_NB).  d, e (n - 1), tau (n - 2).
// Dynamic LDS: 3 n + 2 EG_NB + TH / 64 doubles.
template <int TH>
__global__ void __launch_bounds__(TH)
tridiag_kernel(int n, const double* __restrict__ Ain, int lda, long long sA, double* __restrict__ A,
               double* __restrict__ Vm, double* __restrict__ Wm, double* __restrict__ d,
               double* __restrict__ e, double* __restrict__ tau) {
  const int b = blockIdx.x;
  const long long nn = (long long)n * n;
  const double* src = Ain + (long long)b * sA;
  A += b * nn;
  Vm += b * nn;
  Wm += (long long)b * n * EG_NB;
  d += (long long)b * n;
  e += (long long)b * n;
  tau += (long long)b * n;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  constexpr int NW = TH / 64;
  extern __shared__ double sh[];
  double* col = sh;
  double* v = col + n;
  double* p = v + n;
  double* cf = p + n;             // 2 EG_NB coefficients
  double* red = cf + 2 * EG_NB;
  // symmetric working copy from the upper triangle
  for (long long idx = t; idx < nn; idx += TH) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    A[idx] = (i <= j) ? src[i + (long long)j * lda] : src[j + (long long)i * lda];
    Vm[idx] = 0.0;
  }
  __syncthreads();
  const int nref = n - 2;
  for (int k0 = 0; k0 < nref; k0 += EG_NB) {
    const int nbp = min(EG_NB, nref - k0);
    for (int j = 0; j < nbp; ++j) {
      const int i = k0 + j;
      const int m = n - i;   // rows i..n-1
      // column i with the panel's earlier reflectors applied
      if (t < j) {
        cf[t] = Wm[i + (long long)t * n];
        cf[EG_NB + t] = Vm[i + (long long)(k0 + t) * n];
      }
      __syncthreads();
      for (int r = t; r < m; r += TH) {
        const int row = i + r;
        double a = A[row + (long long)i * n];
        for (int jj = 0; jj < j; ++jj)
          a -= Vm[row + (long long)(k0 + jj) * n] * cf[jj] + Wm[row + (long long)jj * n] * cf[EG_NB + jj];
        col[r] = a;
      }
      __syncthreads();
      // the reflector of col[1 .. m-1]
      double s2 = 0.0;
      for (int r = 2 + t; r < m; r += TH) s2 = fma(col[r], col[r], s2);
      s2 = block_sum<TH>(s2, red);
      const double alpha = col[1];
      double tv = 0.0, beta = alpha, scale = 0.0;
      if (s2 != 0.0) {
        beta = -copysign(sqrt(fma(alpha, alpha, s2)), alpha);
        tv = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      if (t == 0) {
        d[i] = col[0];
        e[i] = beta;
        tau[i] = tv;
      }
      for (int r = 1 + t; r < m; r += TH) {
        const double vr = (r == 1) ? 1.0 : col[r] * scale;
        v[r - 1] = vr;
        Vm[(i + r) + (long long)i * n] = vr;
      }
      __syncthreads();
      // p = A22 v (A22 symmetric: column c dotted with v), t1 = W^T v, t2 = V^T v
      const int mv = m - 1;
      for (int c = wid; c < mv; c += NW) {
        const double* ac = A + (i + 1) + (long long)(i + 1 + c) * n;
        double acc = 0.0;
        for (int r = lane; r < mv; r += 64) acc = fma(ac[r], v[r], acc);
        acc = wave_sum(acc);
        if (lane == 0) p[c] = acc;
      }
      for (int q = wid; q < 2 * j; q += NW) {
        const double* x = (q < j) ? Wm + (i + 1) + (long long)q * n
                                  : Vm + (i + 1) + (long long)(k0 + q - j) * n;
        double acc = 0.0;
        for (int r = lane; r < mv; r += 64) acc = fma(x[r], v[r], acc);
        acc = wave_sum(acc);
        if (lane == 0) cf[q] = acc;   // cf[0, j) = t1, cf[j, 2j) = t2
      }
      __syncthreads();
      for (int r = t; r < mv; r += TH) {
        double a = p[r];
        for (int jj = 0; jj < j; ++jj)
          a -= Vm[(i + 1 + r) + (long long)(k0 + jj) * n] * cf[jj] +
               Wm[(i + 1 + r) + (long long)jj * n] * cf[j + jj];
        p[r] = tv * a;
      }
      __syncthreads();
      double dot = 0.0;
      for (int r = t; r < mv; r += TH) dot = fma(p[r], v[r], dot);
      dot = block_sum<TH>(dot, red);
      const double al = -0.5 * tv * dot;
      for (int r = t; r < mv; r += TH) Wm[(i + 1 + r) + (long long)j * n] = fma(al, v[r], p[r]);
      __syncthreads();
    }
    // trailing update A(s0:, s0:) -= V W^T + W V^T
    const int s0 = k0 + nbp, ms = n - s0;
    for (long long idx = t; idx < (long long)ms * ms; idx += TH) {
      const int r = s0 + (int)(idx % ms), c = s0 + (int)(idx / ms);
      double a = A[r + (long long)c * n];
      for (int jj = 0; jj < nbp; ++jj)
        a -= Vm[r + (long long)(k0 + jj) * n] * Wm[c + (long long)jj * n] +
             Wm[r + (long long)jj * n] * Vm[c + (long long)(k0 + jj) * n];
      A[r + (long long)c * n] = a;
    }
    __syncthreads();
  }
  if (t == 0) {
    if (n >= 2) {
      d[n - 2] = A[(n - 2) + (long long)(n - 2) * n];
      e[n - 2] = A[(n - 1) + (long long)(n - 2) * n];
    }
    d[n - 1] = A[(n - 1) + (long long)(n - 1) * n];
  }
}

// ---------------------------------------------------------------------------
// 2. Divide and conquer.  Leaves are 1 x 1: w_i = d_i - |e_{i-1}| - |e_i|, Z = I.
__global__ void dc_init_kernel(int n, const double* __restrict__ d, const double* __restrict__ e,
                               double* __restrict__ w, double* __restrict__ Z, int* __restrict__ info) {
  const int st = blockIdx.y;
  if (blockIdx.x == 0 && threadIdx.x == 0) info[st] = 0;
  const long long nn = (long long)n * n;
  d += (long long)st * n;
  e += (long long)st * n;
  w += (long long)st * n;
  Z += st * nn;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nn;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    Z[idx] = (i == j) ? 1.0 : 0.0;
    if (i == j) w[i] = d[i] - (i > 0 ? fabs(e[i - 1]) : 0.0) - (i + 1 < n ? fabs(e[i]) : 0.0);
  }
}

// One merge per workgroup: desc holds (r0, n1, n2) of each merge of this height.
// Z: eigenvectors (block diagonal, ld n), w: eigenvalues of the blocks; Qr, Uf:
// the gathered, rotated children's eigenvectors and the merge's eigenvector
// matrix (m x m blocks at (r0, r0)).  Merges wider than EG_SMALL leave the product
// Z = Qr Uf to the caller.
template <int TH, int NM>
__global__ void __launch_bounds__(TH)
dc_merge_kernel(int n, const int* __restrict__ desc, const double* __restrict__ e,
                double* __restrict__ w, double* __restrict__ Z, double* __restrict__ Qr,
                double* __restrict__ Uf, int* __restrict__ info) {
  const int b = blockIdx.x, st = blockIdx.y;
  const long long nn = (long long)n * n;
  const int r0 = desc[3 * b], n1 = desc[3 * b + 1], n2 = desc[3 * b + 2], m = n1 + n2;
  e += (long long)st * n;
  w += (long long)st * n;
  Z += st * nn;
  Qr += st * nn;
  Uf += st * nn;
  __shared__ double ds[NM], zs[NM], dd[NM], zz[NM], tz[NM], zh[NM], cn[NM], ov[NM], rc[NM], rs[NM];
  __shared__ int perm[NM], flag[NM], nd[NM], og[NM], pos[NM], fp[NM], ra[NM], rb[NM];
  __shared__ double red[TH / 64];
  __shared__ int s_k, s_nrot;
  const int t = threadIdx.x;
  const double beta = e[r0 + n1 - 1];
  const double rho = 2.0 * fabs(beta);
  const double sg = (beta >= 0.0) ? 1.0 : -1.0;
  const double isq2 = 0.70710678118654752440;
  // the children's eigenvalues and the coupling vector (z / sqrt 2, rho = 2 |beta|)
  for (int u = t; u < m; u += TH) {
    ov[u] = w[r0 + u];
    const double zu = (u < n1) ? Z[(r0 + n1 - 1) + (long long)(r0 + u) * n]
                               : sg * Z[(r0 + n1) + (long long)(r0 + u) * n];
    rc[u] = zu * isq2;
  }
  __syncthreads();
  for (int u = t; u < m; u += TH) {
    const double x = ov[u];
    int r = 0;
    for (int q = 0; q < m; ++q) {
      const double y = ov[q];
      r += (y < x) || (y == x && q < u);
    }
    perm[r] = u;
  }
  __syncthreads();
  double mx = 0.0;
  for (int p = t; p < m; p += TH) {
    ds[p] = ov[perm[p]];
    zs[p] = rc[perm[p]];
    mx = fmax(mx, fmax(fabs(ds[p]), fabs(zs[p])));
  }
  mx = block_max<TH>(mx, red);
  // deflation, in order of the sorted eigenvalues
  if (t == 0) {
    const double tol = 8.0 * DBL_EPSILON * mx;
    int prev = -1, nrot = 0, k = 0;
    for (int p = 0; p < m; ++p) flag[p] = 0;
    for (int p = 0; p < m; ++p) {
      if (rho * fabs(zs[p]) <= tol) {
        flag[p] = 1;
        continue;
      }
      if (prev >= 0) {
        const double tt = hypot(zs[prev], zs[p]);
        const double c = zs[p] / tt, s = -zs[prev] / tt;
        if (fabs(c * s * (ds[p] - ds[prev])) <= tol) {
          ra[nrot] = prev;
          rb[nrot] = p;
          rc[nrot] = c;
          rs[nrot] = s;
          ++nrot;
          const double dp = ds[prev] * c * c + ds[p] * s * s;
          ds[p] = ds[prev] * s * s + ds[p] * c * c;
          ds[prev] = dp;
          zs[p] = tt;
          zs[prev] = 0.0;
          flag[prev] = 1;
        }
      }
      prev = p;
    }
    for (int p = 0; p < m; ++p)
      if (!flag[p]) nd[k++] = p;
    s_k = k;
    s_nrot = nrot;
  }
  __syncthreads();
  const int k = s_k, nrot = s_nrot;
  // the non-deflated ascending (a rotation can perturb the order)
  for (int j = t; j < k; j += TH) ov[j] = ds[nd[j]];
  __syncthreads();
  for (int j = t; j < k; j += TH) {
    const double x = ov[j];
    int r = 0;
    for (int q = 0; q < k; ++q) {
      const double y = ov[q];
      r += (y < x) || (y == x && q < j);
    }
    pos[r] = nd[j];
  }
  __syncthreads();
  for (int j = t; j < k; j += TH) {
    nd[j] = pos[j];
    dd[j] = ds[pos[j]];
    zz[j] = zs[pos[j]];
  }
  __syncthreads();
  for (int j = t; j < k; j += TH) pos[nd[j]] = j;   // position -> root index
  // secular roots: l_j = dd[og_j] + tz_j
  for (int j = t; j < k; j += TH) {
    int o;
    double a, bnd;
    if (j < k - 1) {
      const double mid = 0.5 * (dd[j + 1] - dd[j]);
      double fm = 1.0;
      for (int i = 0; i < k; ++i) fm += rho * zz[i] * zz[i] / ((dd[i] - dd[j]) - mid);
      if (fm > 0.0) {
        o = j;
        a = 0.0;
        bnd = mid;
      } else {
        o = j + 1;
        a = -mid;
        bnd = 0.0;
      }
    } else {
      o = k - 1;
      a = 0.0;
      double zn = 0.0;
      for (int i = 0; i < k; ++i) zn = fma(zz[i], zz[i], zn);
      bnd = rho * zn;
    }
    const int jh = (j < k - 1) ? j + 1 : -1;
    double tau = 0.5 * (a + bnd);
    bool ok = false;
    for (int it = 0; it < 100 && !ok; ++it) {
      double psi = 0.0, phi = 0.0, dpsi = 0.0, dphi = 0.0;
      for (int i = 0; i < k; ++i) {
        const double den = (dd[i] - dd[o]) - tau;
        const double term = rho * zz[i] * zz[i] / den;
        if (jh < 0 || i < jh) {
          psi += term;
          dpsi += term / den;
        } else {
          phi += term;
          dphi += term / den;
        }
      }
      const double f = 1.0 + psi + phi;
      const double err = 4.0 * DBL_EPSILON * (1.0 + fabs(psi) + fabs(phi) + fabs(tau) * (dpsi + dphi));
      if (fabs(f) <= err) {
        ok = true;
        break;
      }
      if (f < 0.0) a = tau;
      else bnd = tau;
      const double Dl = (dd[j] - dd[o]) - tau;
      double eta;
      if (jh >= 0) {
        const double Dh = (dd[jh] - dd[o]) - tau;
        const double c = f - dpsi * Dl - dphi * Dh;
        const double sl = dpsi * Dl * Dl, shh = dphi * Dh * Dh;
        const double AA = c * (Dl + Dh) + sl + shh;
        const double BB = c * Dl * Dh + sl * Dh + shh * Dl;
        if (c == 0.0) {
          eta = BB / AA;
        } else {
          const double disc = sqrt(fabs(AA * AA - 4.0 * BB * c));
          eta = (AA <= 0.0) ? (AA - disc) / (2.0 * c) : 2.0 * BB / (AA + disc);
        }
      } else {
        const double c = f - dpsi * Dl;
        eta = (c != 0.0) ? Dl + dpsi * Dl * Dl / c : -f / dpsi;
      }
      if (!(f * eta < 0.0)) eta = -f / (dpsi + dphi);
      double t2 = tau + eta;
      if (!(t2 > a && t2 < bnd)) t2 = 0.5 * (a + bnd);
      if (fabs(t2 - tau) <= 2.0 * DBL_EPSILON * fmax(fabs(tau), DBL_MIN) || !(bnd > a)) ok = true;
      tau = t2;
    }
    if (!ok) info[st] = 1;
    og[j] = o;
    tz[j] = tau;
  }
  __syncthreads();
  // z from the roots (Loewner), column norms of the secular eigenvectors
  for (int i = t; i < k; i += TH) {
    double pr = (dd[og[k - 1]] - dd[i]) + tz[k - 1];
    for (int j = 0; j < k - 1; ++j) {
      const double num = (dd[og[j]] - dd[i]) + tz[j];
      const double den = (j < i) ? (dd[j] - dd[i]) : (dd[j + 1] - dd[i]);
      pr *= num / den;
    }
    pr /= rho;
    zh[i] = copysign(sqrt(fabs(pr)), zz[i]);
  }
  __syncthreads();
  for (int j = t; j < k; j += TH) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) {
      const double x = zh[i] / ((dd[i] - dd[og[j]]) - tz[j]);
      s = fma(x, x, s);
    }
    cn[j] = 1.0 / sqrt(s);
  }
  // merged eigenvalues in the sorted-index space and their final positions
  for (int p = t; p < m; p += TH) ov[p] = flag[p] ? ds[p] : dd[og[pos[p]]] + tz[pos[p]];
  __syncthreads();
  for (int p = t; p < m; p += TH) {
    const double x = ov[p];
    int r = 0;
    for (int q = 0; q < m; ++q) {
      const double y = ov[q];
      r += (y < x) || (y == x && q < p);
    }
    fp[p] = r;
  }
  for (long long idx = t; idx < (long long)m * m; idx += TH)
    Uf[(r0 + (int)(idx % m)) + (long long)(r0 + (int)(idx / m)) * n] = 0.0;
  __syncthreads();
  // the merge's eigenvector matrix: column fp[p] is e_p (deflated p) or the
  // secular eigenvector of root j (p = nd[j]) over the non-deflated rows
  for (int p = t; p < m; p += TH) {
    double* uc = Uf + r0 + (long long)(r0 + fp[p]) * n;
    if (flag[p]) {
      uc[p] = 1.0;
    } else {
      const int j = pos[p];
      const double dj = dd[og[j]], tj = tz[j], sj = cn[j];
      for (int i = 0; i < k; ++i) uc[nd[i]] = zh[i] / ((dd[i] - dj) - tj) * sj;
    }
  }
  // the children's eigenvectors gathered in sorted order, then the deflation
  // rotations, row by row
  for (int row = t; row < m; row += TH) {
    const double* zr = Z + (r0 + row);
    double* qr = Qr + (r0 + row);
    for (int p = 0; p < m; ++p) qr[(long long)(r0 + p) * n] = zr[(long long)(r0 + perm[p]) * n];
    for (int q = 0; q < nrot; ++q) {
      const long long ca = (long long)(r0 + ra[q]) * n, cb = (long long)(r0 + rb[q]) * n;
      const double xa = qr[ca], xb = qr[cb];
      qr[ca] = rc[q] * xa + rs[q] * xb;
      qr[cb] = -rs[q] * xa + rc[q] * xb;
    }
  }
  __syncthreads();
  if (m <= EG_SMALL) {
    for (long long idx = t; idx < (long long)m * m; idx += TH) {
      const int row = (int)(idx % m), c = (int)(idx / m);
      double acc = 0.0;
      for (int q = 0; q < m; ++q)
        acc = fma(Qr[(r0 + row) + (long long)(r0 + q) * n], Uf[(r0 + q) + (long long)(r0 + c) * n], acc);
      Z[(r0 + row) + (long long)(r0 + c) * n] = acc;
    }
  }
  for (int p = t; p < m; p += TH) w[r0 + fp[p]] = ov[p];
}

// ---------------------------------------------------------------------------
// 3. Back-transformation.  T of reflector block p (columns k0 .. k0+nbp-1):
// T(a, a) = tau_a, T(0:a, a) = -tau_a T(0:a, 0:a) (V^T v_a)(0:a).
template <int TH>
__global__ void __launch_bounds__(TH)
wy_t_kernel(int n, const double* __restrict__ Vm, const double* __restrict__ tau,
            double* __restrict__ Tm) {
  const int pnl = blockIdx.x, st = blockIdx.y, npan = gridDim.x;
  const long long nn = (long long)n * n;
  Vm += st * nn;
  tau += (long long)st * n;
  Tm += ((long long)st * npan + pnl) * EG_NB * EG_NB;
  const int nref = n - 2, k0 = pnl * EG_NB, nbp = min(EG_NB, nref - k0);
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  constexpr int NW = TH / 64;
  __shared__ double G[EG_NB][EG_NB + 1], T[EG_NB][EG_NB + 1];
  // G(a, c) = v_a . v_c for a < c (v_c vanishes above row k0 + c + 1)
  for (int q = wid; q < nbp * nbp; q += NW) {
    const int a = q % nbp, c = q / nbp;
    if (a >= c) continue;
    const double* va = Vm + (long long)(k0 + a) * n;
    const double* vc = Vm + (long long)(k0 + c) * n;
    double acc = 0.0;
    for (int r = k0 + c + 1 + lane; r < n; r += 64) acc = fma(va[r], vc[r], acc);
    acc = wave_sum(acc);
    if (lane == 0) G[a][c] = acc;
  }
  for (int q = t; q < EG_NB * EG_NB; q += TH) T[q % EG_NB][q / EG_NB] = 0.0;
  __syncthreads();
  if (t < nbp) T[t][t] = tau[k0 + t];
  __syncthreads();
  for (int c = 1; c < nbp; ++c) {
    if (t < c) {
      double acc = 0.0;
      for (int l = t; l < c; ++l) acc = fma(T[t][l], G[l][c], acc);
      T[t][c] = -tau[k0 + c] * acc;
    }
    __syncthreads();
  }
  for (int q = t; q < EG_NB * EG_NB; q += TH) Tm[q] = T[q % EG_NB][q / EG_NB];
}

// E(:, c0:c0+CB) = (I - V_0 T_0 V_0^T) ... (I - V_P T_P V_P^T) Z(:, c0:c0+CB).
// Dynamic LDS: n (CB + 1) + EG_NB CB + EG_NB (EG_NB + 1) doubles.
template <int TH, int CB>
__global__ void __launch_bounds__(TH)
wy_apply_kernel(int n, int npan, const double* __restrict__ Vm, const double* __restrict__ Tm,
                const double* __restrict__ Z, double* __restrict__ E, long long sE) {
  const int st = blockIdx.y, c0 = blockIdx.x * CB, nc = min(CB, n - c0);
  const long long nn = (long long)n * n;
  Vm += st * nn;
  Z += st * nn;
  E += st * sE;
  Tm += (long long)st * npan * EG_NB * EG_NB;
  const int t = threadIdx.x;
  extern __shared__ double sh[];
  double* Zb = sh;                               // Zb[r * (CB + 1) + c]
  double* X = Zb + (size_t)n * (CB + 1);         // X[a * CB + c]
  double* T = X + EG_NB * CB;                    // T[a * (EG_NB + 1) + b]
  constexpr int LDZ = CB + 1;
  for (long long idx = t; idx < (long long)n * CB; idx += TH) {
    const int r = (int)(idx % n), c = (int)(idx / n);
    Zb[r * LDZ + c] = (c < nc) ? Z[r + (long long)(c0 + c) * n] : 0.0;
  }
  const int nref = n - 2;
  const int c = t % CB, g = t / CB;
  constexpr int NG = TH / CB;
  for (int pnl = npan - 1; pnl >= 0; --pnl) {
    const int k0 = pnl * EG_NB, nbp = min(EG_NB, nref - k0);
    for (int q = t; q < EG_NB * EG_NB; q += TH)
      T[(q % EG_NB) * (EG_NB + 1) + q / EG_NB] = Tm[(long long)pnl * EG_NB * EG_NB + q];
    __syncthreads();
    // X = V^T Zb
    for (int a = g; a < nbp; a += NG) {
      const double* va = Vm + (long long)(k0 + a) * n;
      double acc = 0.0;
      for (int r = k0 + a + 1; r < n; ++r) acc = fma(va[r], Zb[r * LDZ + c], acc);
      X[a * CB + c] = acc;
    }
    __syncthreads();
    // Y = T X (in place, row a from rows >= a: top-down reads rows not yet written)
    double y[EG_NB / NG + 1];
    int ny = 0;
    for (int a = g; a < nbp; a += NG) {
      double acc = 0.0;
      for (int l = a; l < nbp; ++l) acc = fma(T[a * (EG_NB + 1) + l], X[l * CB + c], acc);
      y[ny++] = acc;
    }
    __syncthreads();
    ny = 0;
    for (int a = g; a < nbp; a += NG) X[a * CB + c] = y[ny++];
    __syncthreads();
    // Zb -= V Y
    for (int r = k0 + 1 + g; r < n; r += NG) {
      double acc = Zb[r * LDZ + c];
      const int amax = min(nbp, r - k0);   // v_a(r) = 0 for r <= k0 + a
      for (int a = 0; a < amax; ++a) acc = fma(-Vm[r + (long long)(k0 + a) * n], X[a * CB + c], acc);
      Zb[r * LDZ + c] = acc;
    }
    __syncthreads();
  }
  for (long long idx = t; idx < (long long)n * nc; idx += TH) {
    const int r = (int)(idx % n), cc = (int)(idx / n);
    E[r + (long long)(c0 + cc) * n] = Zb[r * LDZ + cc];
  }
}

// merges grouped by height (children first): (r0, n1, n2) triples
struct MergePlan {
  std::vector<int> desc;                 // all levels, concatenated
  std::vector<int> start, count;         // per height
};

int plan_node(int r0, int m, std::vector<std::vector<int>>& byh) {
  if (m <= 1) return 0;
  const int n1 = m / 2, n2 = m - n1;
  const int h = 1 + std::max(plan_node(r0, n1, byh), plan_node(r0 + n1, n2, byh));
  if ((int)byh.size() < h) byh.resize(h);
  byh[h - 1].insert(byh[h - 1].end(), {r0, n1, n2});
  return h;
}

const MergePlan& merge_plan(int n) {
  static std::mutex mu;
  static std::map<int, MergePlan> plans;
  std::lock_guard<std::mutex> lk(mu);
  auto it = plans.find(n);
  if (it != plans.end()) return it->second;
  std::vector<std::vector<int>> byh;
  plan_node(0, n, byh);
  MergePlan P;
  for (auto& lv : byh) {
    P.start.push_back((int)P.desc.size() / 3);
    P.count.push_back((int)lv.size() / 3);
    P.desc.insert(P.desc.end(), lv.begin(), lv.end());
  }
  return plans.emplace(n, std::move(P)).first->second;
}

}  // namespace

size_t sym_eigensystem_work_bytes(int n, int count) {
  const size_t nn = (size_t)n * n;
  const int npan = n > 2 ? (n - 2 + EG_NB - 1) / EG_NB : 0;
  const size_t doubles = (size_t)count * (5 * nn + (size_t)n * EG_NB + 4 * (size_t)n +
                                          (size_t)std::max(1, npan) * EG_NB * EG_NB);
  const size_t ints = (size_t)3 * n + (size_t)count;
  return sizeof(double) * doubles + sizeof(int) * ints + 256;
}

int batched_sym_eigensystem(hipStream_t s, int n, const double* A, int lda, long long sA, int count,
                            void* work, double* w, long long ws, double* E, long long sE, int* info) {
  if (count <= 0 || n <= 0) return TG_OK;
  if (n > EG_NMAX || lda < n || ws < n || sE < (long long)n * n) {
    tg_set_error("batched_sym_eigensystem: order above 1024 or bad strides");
    return TG_E_UNSUPPORTED;
  }
  const long long nn = (long long)n * n;
  const int npan = n > 2 ? (n - 2 + EG_NB - 1) / EG_NB : 0;
  double* Aw = static_cast<double*>(work);
  double* Vm = Aw + count * nn;
  double* Z = Vm + count * nn;
  double* Qr = Z + count * nn;
  double* Uf = Qr + count * nn;
  double* Wm = Uf + count * nn;
  double* d = Wm + (size_t)count * n * EG_NB;
  double* e = d + (size_t)count * n;
  double* tau = e + (size_t)count * n;
  double* wi = tau + (size_t)count * n;
  double* Tm = wi + (size_t)count * n;
  int* idesc = reinterpret_cast<int*>(Tm + (size_t)count * std::max(1, npan) * EG_NB * EG_NB);
  int* dinfo = idesc + 3 * n;
  const MergePlan& P = merge_plan(n);
  if (!P.desc.empty())
    TG_HIP(hipMemcpyAsync(idesc, P.desc.data(), sizeof(int) * P.desc.size(), hipMemcpyHostToDevice, s));
  const size_t lds1 = sizeof(double) * (3 * (size_t)n + 2 * EG_NB + EG_TH / 64);
  hipLaunchKernelGGL(tridiag_kernel<EG_TH>, dim3(count), dim3(EG_TH), lds1, s, n, A, lda, sA, Aw, Vm,
                     Wm, d, e, tau);
  TG_HIP(hipGetLastError());
  const unsigned gi = (unsigned)std::min<long long>((nn + EG_TH - 1) / EG_TH, 1024);
  hipLaunchKernelGGL(dc_init_kernel, dim3(gi, count), dim3(EG_TH), 0, s, n, d, e, wi, Z, dinfo);
  TG_HIP(hipGetLastError());
  for (size_t h = 0; h < P.count.size(); ++h) {
    const int* dh = idesc + 3 * P.start[h];
    hipLaunchKernelGGL((dc_merge_kernel<EG_TH, EG_NMAX>), dim3(P.count[h], count), dim3(EG_TH), 0, s, n,
                       dh, e, wi, Z, Qr, Uf, dinfo);
    TG_HIP(hipGetLastError());
    for (int q = 0; q < P.count[h]; ++q) {
      const int* g = &P.desc[3 * (P.start[h] + q)];
      const int r0 = g[0], m = g[1] + g[2];
      if (m <= EG_SMALL) continue;
      const long long off = r0 + (long long)r0 * n;
      TG_TRY(gemm_dd(s, OP_N, OP_N, m, m, m, 1.0, Qr + off, n, nn, Uf + off, n, nn, 0.0, Z + off, n,
                     nn, count));
    }
  }
  if (npan > 0) {
    hipLaunchKernelGGL(wy_t_kernel<EG_TH>, dim3(npan, count), dim3(EG_TH), 0, s, n, Vm, tau, Tm);
    TG_HIP(hipGetLastError());
  }
  if (n <= 512) {
    constexpr int CB = 32;
    const size_t lds = sizeof(double) * ((size_t)n * (CB + 1) + EG_NB * CB + EG_NB * (EG_NB + 1));
    hipLaunchKernelGGL((wy_apply_kernel<EG_TH, CB>), dim3((n + CB - 1) / CB, count), dim3(EG_TH), lds, s,
                       n, npan, Vm, Tm, Z, E, sE);
  } else {
    constexpr int CB = 16;
    const size_t lds = sizeof(double) * ((size_t)n * (CB + 1) + EG_NB * CB + EG_NB * (EG_NB + 1));
    hipLaunchKernelGGL((wy_apply_kernel<EG_TH, CB>), dim3((n + CB - 1) / CB, count), dim3(EG_TH), lds, s,
                       n, npan, Vm, Tm, Z, E, sE);
  }
  TG_HIP(hipGetLastError());
  TG_HIP(hipMemcpy2DAsync(w, sizeof(double) * ws, wi, sizeof(double) * n, sizeof(double) * n, count,
                          hipMemcpyDeviceToDevice, s));
  if (info) TG_HIP(hipMemcpyAsync(info, dinfo, sizeof(int) * count, hipMemcpyDeviceToDevice, s));
  return TG_OK;
}

extern "C" int tg_sym_eigensystem(int n, int count, const double* A, int lda, double* w, double* Z) {
  if (n < 0 || count < 0 || n > EG_NMAX || lda < std::max(1, n) ||
      (count > 0 && n > 0 && (!A || !w || !Z))) {
    tg_set_error("tg_sym_eigensystem: invalid argument (n <= 1024)");
    return TG_E_ARG;
  }
  if (count == 0 || n == 0) return TG_OK;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) {
    tg_set_error("no HIP device available");
    return TG_E_NODEVICE;
  }
  const size_t m = (size_t)lda * n, nn = (size_t)n * n;
  const size_t wb = sym_eigensystem_work_bytes(n, count);
  double *dA = nullptr, *dw = nullptr, *dZ = nullptr;
  void* work = nullptr;
  int* dinfo = nullptr;
  int r = TG_OK;
  if (hipMalloc(&dA, sizeof(double) * m * count) != hipSuccess ||
      hipMalloc(&dw, sizeof(double) * (size_t)n * count) != hipSuccess ||
      hipMalloc(&dZ, sizeof(double) * nn * count) != hipSuccess || hipMalloc(&work, wb) != hipSuccess ||
      hipMalloc(&dinfo, sizeof(int) * count) != hipSuccess) {
    tg_set_error("tg_sym_eigensystem: hipMalloc failed");
    r = TG_E_NOMEM;
  }
  if (r == TG_OK && hipMemcpy(dA, A, sizeof(double) * m * count, hipMemcpyHostToDevice) != hipSuccess)
    r = TG_E_HIP;
  if (r == TG_OK)
    r = batched_sym_eigensystem(0, n, dA, lda, (long long)m, count, work, dw, n, dZ, (long long)nn, dinfo);
  std::vector<int> hinfo(count, 0);
  if (r == TG_OK &&
      (hipDeviceSynchronize() != hipSuccess ||
       hipMemcpy(w, dw, sizeof(double) * (size_t)n * count, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(Z, dZ, sizeof(double) * nn * count, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(hinfo.data(), dinfo, sizeof(int) * count, hipMemcpyDeviceToHost) != hipSuccess)) {
    tg_set_error("tg_sym_eigensystem: device error");
    r = TG_E_HIP;
  }
  for (int k = 0; r == TG_OK && k < count; ++k)
    if (hinfo[k] != 0) {
      tg_set_error("tg_sym_eigensystem: secular equation did not converge (matrix " + std::to_string(k) + ")");
      r = TG_E_SINGULAR;
    }
  for (void* p : {(void*)dA, (void*)dw, (void*)dZ, work, (void*)dinfo})
    if (p) (void)hipFree(p);
  return r;
}
