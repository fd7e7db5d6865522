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
//      carrying 8 columns of the eigenvector matrix, in LDS, through every block.
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

// 1/x from the hardware reciprocal and two Newton steps (the secular sums only
// need the terms to a few ulps; the quotients that fix the eigenvectors divide)
__device__ inline double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// ---------------------------------------------------------------------------
// 1. Householder tridiagonalization (lower), one workgroup per matrix.
// A: n x n working copy (ld n), Vm: reflectors by column (v_i(i+1) = 1, zero
// above) for the back-transformation.  d, e (n - 1), tau (n - 2).
// The panel of PB columns keeps its reflectors and w vectors in LDS (Vp, Wp,
// n x PB each), so that only the product with the trailing matrix reads L2; that
// product runs by rows (A22 symmetric), each thread one row over a slice of the
// columns, with the slices summed through LDS.
// Dynamic LDS: (3 + 2 PB + TD_NS) n + 2 PB + TH / 64 doubles.
constexpr int TD_TH = 512;
constexpr int TD_NS = 4;        // column slices of the row products

template <int TH, int PB>
__global__ void __launch_bounds__(TH)
tridiag_kernel(int n, const double* __restrict__ Ain, int lda, long long sA, double* __restrict__ A,
               double* __restrict__ Vm, double* __restrict__ d, double* __restrict__ e,
               double* __restrict__ tau) {
  const int b = blockIdx.x;
  const long long nn = (long long)n * n;
  const double* src = Ain + (long long)b * sA;
  A += b * nn;
  Vm += b * nn;
  d += (long long)b * n;
  e += (long long)b * n;
  tau += (long long)b * n;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  constexpr int NW = TH / 64;
  extern __shared__ double sh[];
  double* col = sh;
  double* v = col + n;
  double* p = v + n;
  double* Vp = p + n;                       // Vp[r + jj n] = V(r, k0 + jj)
  double* Wp = Vp + (size_t)PB * n;         // Wp[r + jj n] = W(r, jj)
  double* part = Wp + (size_t)PB * n;       // part[sl n + r]
  double* cf = part + (size_t)TD_NS * n;    // 2 PB
  double* red = cf + 2 * PB;
  // symmetric working copy from the upper triangle
  for (int j = 0; j < n; ++j)
    for (int i = t; i < n; i += TH) {
      const long long o = i + (long long)j * n;
      A[o] = (i <= j) ? src[i + (long long)j * lda] : src[j + (long long)i * lda];
      Vm[o] = 0.0;
    }
  __syncthreads();
  const int nref = n - 2;
  for (int k0 = 0; k0 < nref; k0 += PB) {
    const int nbp = min(PB, nref - k0);
    // the panel arrays start at zero: the unrolled loops below read all PB columns
    for (int q = t; q < 2 * PB * n; q += TH) Vp[q] = 0.0;   // Vp and Wp are adjacent
    __syncthreads();
    for (int j = 0; j < nbp; ++j) {
      const int i = k0 + j;
      const int m = n - i;   // rows i..n-1
      // column i with the panel's earlier reflectors applied (columns jj >= j of
      // Vp, Wp are still zero); LDS operands loaded into registers first
      {
        double ci_w[PB], ci_v[PB];
#pragma unroll
        for (int jj = 0; jj < PB; ++jj) {
          ci_w[jj] = Wp[i + jj * n];
          ci_v[jj] = Vp[i + jj * n];
        }
        for (int r = t; r < m; r += TH) {
          const int row = i + r;
          double rv[PB], rw[PB];
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) {
            rv[jj] = Vp[row + jj * n];
            rw[jj] = Wp[row + jj * n];
          }
          double a = A[row + (long long)i * n];
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) a = fma(-rv[jj], ci_w[jj], fma(-rw[jj], ci_v[jj], a));
          col[r] = a;
        }
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
        Vp[(i + r) + j * n] = vr;
        Vm[(i + r) + (long long)i * n] = vr;
      }
      __syncthreads();
      // p = A22 v by rows over TD_NS column slices; t1 = W^T v, t2 = V^T v (LDS)
      const int mv = m - 1;
      const int ns = min(TD_NS, max(1, TH / mv));
      for (int tt = t; tt < ns * mv; tt += TH) {
        const int sl = tt / mv, r = tt - sl * mv;
        const double* ar = A + (i + 1 + r) + (long long)(i + 1) * n;
        // sixteen loads in flight per lane: the row product is bound by L2 latency
        constexpr int U = 16;
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = 0.0;
        int c = sl;
        const long long step = (long long)ns * n;
        for (; c + (U - 1) * ns < mv; c += U * ns) {
          double x[U], vv[U];
          const double* pc = ar + (long long)c * n;
#pragma unroll
          for (int u = 0; u < U; ++u) x[u] = pc[u * step];
#pragma unroll
          for (int u = 0; u < U; ++u) vv[u] = v[c + u * ns];
#pragma unroll
          for (int u = 0; u < U; ++u) acc[u] = fma(x[u], vv[u], acc[u]);
        }
        for (; c < mv; c += ns) acc[0] = fma(ar[(long long)c * n], v[c], acc[0]);
#pragma unroll
        for (int u = 1; u < U; ++u) acc[0] += acc[u];
        part[sl * n + r] = acc[0];
      }
      for (int q = wid; q < 2 * j; q += NW) {
        const double* x = (q < j) ? Wp + (i + 1) + q * n : Vp + (i + 1) + (q - j) * n;
        double acc = 0.0;
        for (int r = lane; r < mv; r += 64) acc = fma(x[r], v[r], acc);
        acc = wave_sum(acc);
        if (lane == 0) cf[q] = acc;   // cf[0, j) = t1, cf[j, 2j) = t2
      }
      __syncthreads();
      {
        double t1[PB], t2[PB];
#pragma unroll
        for (int jj = 0; jj < PB; ++jj) {
          t1[jj] = (jj < j) ? cf[jj] : 0.0;
          t2[jj] = (jj < j) ? cf[j + jj] : 0.0;
        }
        for (int r = t; r < mv; r += TH) {
          const int row = i + 1 + r;
          double rv[PB], rw[PB];
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) {
            rv[jj] = Vp[row + jj * n];
            rw[jj] = Wp[row + jj * n];
          }
          double a = part[r];
          for (int sl = 1; sl < ns; ++sl) a += part[sl * n + r];
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) a = fma(-rv[jj], t1[jj], fma(-rw[jj], t2[jj], a));
          p[r] = tv * a;
        }
      }
      __syncthreads();
      double dot = 0.0;
      for (int r = t; r < mv; r += TH) dot = fma(p[r], v[r], dot);
      dot = block_sum<TH>(dot, red);
      const double al = -0.5 * tv * dot;
      for (int r = t; r < mv; r += TH) Wp[(i + 1 + r) + j * n] = fma(al, v[r], p[r]);
      __syncthreads();
    }
    // trailing update A(s0:, s0:) -= V W^T + W V^T: each thread one row (its V, W
    // rows in registers) over a share of the columns
    const int s0 = k0 + nbp, ms = n - s0;
    if (ms > 0) {
      const int ng = max(1, TH / ms);
      for (int tt = t; tt < ng * ms; tt += TH) {
        const int r = s0 + tt % ms, cg = tt / ms;
        double vr[PB], wr[PB];
#pragma unroll
        for (int jj = 0; jj < PB; ++jj) {
          vr[jj] = Vp[r + jj * n];
          wr[jj] = Wp[r + jj * n];
        }
        for (int c = s0 + cg; c < n; c += ng) {
          double* ap = A + r + (long long)c * n;
          double wc[PB], vc[PB];
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) {
            wc[jj] = Wp[c + jj * n];
            vc[jj] = Vp[c + jj * n];
          }
          double a = *ap;
#pragma unroll
          for (int jj = 0; jj < PB; ++jj) a = fma(-vr[jj], wc[jj], fma(-wr[jj], vc[jj], a));
          *ap = a;
        }
      }
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
    #pragma unroll 8
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
        const double tt = sqrt(fma(zs[prev], zs[prev], zs[p] * zs[p]));
        const double it = 1.0 / tt;
        const double c = zs[p] * it, s = -zs[prev] * it;
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
    #pragma unroll 8
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
  for (int j = t; j < k; j += TH) {
    pos[nd[j]] = j;                   // position -> root index
    zh[j] = rho * zz[j] * zz[j];      // rho z^2 (zh holds z-hat after the roots)
  }
  __syncthreads();
  // secular roots: l_j = dd[og_j] + tz_j
  for (int j = t; j < k; j += TH) {
    int o;
    double a, bnd;
    if (j < k - 1) {
      const double mid = 0.5 * (dd[j + 1] - dd[j]);
      double fm = 1.0;
      #pragma unroll 8
      for (int i = 0; i < k; ++i) fm += zh[i] / ((dd[i] - dd[j]) - mid);
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
      const int isplit = (jh < 0) ? k : jh;
      const double dob = dd[o];
      #pragma unroll 8
      for (int i = 0; i < isplit; ++i) {
        const double rinv = recip((dd[i] - dob) - tau);
        const double term = zh[i] * rinv;
        psi += term;
        dpsi = fma(term, rinv, dpsi);
      }
      #pragma unroll 8
      for (int i = isplit; i < k; ++i) {
        const double rinv = recip((dd[i] - dob) - tau);
        const double term = zh[i] * rinv;
        phi += term;
        dphi = fma(term, rinv, dphi);
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
    #pragma unroll 8
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
    #pragma unroll 8
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
    #pragma unroll 8
    for (int q = 0; q < m; ++q) {
      const double y = ov[q];
      r += (y < x) || (y == x && q < p);
    }
    fp[p] = r;
  }
  for (int c = 0; c < m; ++c)
    for (int row = t; row < m; row += TH) Uf[(r0 + row) + (long long)(r0 + c) * n] = 0.0;
  __syncthreads();
  // the merge's eigenvector matrix: column fp[p] is e_p (deflated p) or the
  // secular eigenvector of root j (p = nd[j]) over the non-deflated rows
  for (int p = t; p < m; p += TH)
    if (flag[p]) Uf[(r0 + p) + (long long)(r0 + fp[p]) * n] = 1.0;
  for (int i = t; i < k; i += TH) {   // row nd[i]: consecutive threads, consecutive rows
    double* ur = Uf + (r0 + nd[i]);
    const double zi = zh[i], di = dd[i];
    #pragma unroll 8
    for (int j = 0; j < k; ++j)
      ur[(long long)(r0 + fp[nd[j]]) * n] = zi / ((di - dd[og[j]]) - tz[j]) * cn[j];
  }
  // the children's eigenvectors gathered in sorted order, then the deflation
  // rotations in their order, row by row
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
    for (int idx = t; idx < m * m; idx += TH) {
      const int c = idx / m, row = idx - c * m;
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
  const int t = threadIdx.x;
  constexpr int RC = 64, OPT = EG_NB * EG_NB / TH;
  __shared__ double G[EG_NB][EG_NB + 1], T[EG_NB][EG_NB + 1], Vc[RC][EG_NB + 1];
  // G = V^T V over rows k0+1 .. n-1, the rows passing through LDS in chunks;
  // thread t holds G(a, c) for c = t % 32, a = t / 32 + 8 u
  double acc[OPT];
#pragma unroll
  for (int u = 0; u < OPT; ++u) acc[u] = 0.0;
  const int cc = t % EG_NB, ab = t / EG_NB;
  for (int r0 = k0 + 1; r0 < n; r0 += RC) {
    const int rows = min(RC, n - r0);
    __syncthreads();
    for (int q = t; q < RC * EG_NB; q += TH) {
      const int rr = q % RC, a = q / RC;
      Vc[rr][a] = (rr < rows && a < nbp) ? Vm[(r0 + rr) + (long long)(k0 + a) * n] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < rows; ++rr) {
      const double vc = Vc[rr][cc];
#pragma unroll
      for (int u = 0; u < OPT; ++u) acc[u] = fma(Vc[rr][ab + (TH / EG_NB) * u], vc, acc[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < OPT; ++u) G[ab + (TH / EG_NB) * u][cc] = acc[u];
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
// V of the block passes through LDS in chunks of WY_RC rows, twice per block
// (for X = V^T Zb, then for Zb -= V (T X)).
// Dynamic LDS: n (CB + 1) + EG_NB CB + EG_NB (EG_NB + 1) + WY_RC (EG_NB + 1) doubles.
constexpr int WY_RC = 64;

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
  constexpr int LDZ = CB + 1, LDV = EG_NB + 1;
  constexpr int NG = TH / CB;               // thread groups per column
  constexpr int OPT = EG_NB / NG > 0 ? EG_NB / NG : 1;   // X entries per thread
  extern __shared__ double sh[];
  double* Zb = sh;                               // Zb[r * LDZ + c]
  double* X = Zb + (size_t)n * LDZ;              // X[a * CB + c]
  double* T = X + EG_NB * CB;                    // T[a * LDV + b]
  double* Vc = T + EG_NB * LDV;                  // Vc[rr * LDV + a]
  for (int cc = 0; cc < CB; ++cc)
    for (int r = t; r < n; r += TH) Zb[r * LDZ + cc] = (cc < nc) ? Z[r + (long long)(c0 + cc) * n] : 0.0;
  const int nref = n - 2;
  const int c = t % CB, g = t / CB;
  for (int pnl = npan - 1; pnl >= 0; --pnl) {
    const int k0 = pnl * EG_NB, nbp = min(EG_NB, nref - k0);
    for (int q = t; q < EG_NB * EG_NB; q += TH)
      T[(q % EG_NB) * LDV + q / EG_NB] = Tm[(long long)pnl * EG_NB * EG_NB + q];
    double acc[OPT];
#pragma unroll
    for (int u = 0; u < OPT; ++u) acc[u] = 0.0;
    // X = V^T Zb over rows k0+1 .. n-1 (v_a vanishes above row k0 + a + 1)
    for (int r0 = k0 + 1; r0 < n; r0 += WY_RC) {
      const int rows = min(WY_RC, n - r0);
      __syncthreads();
      for (int q = t; q < WY_RC * EG_NB; q += TH) {
        const int rr = q % WY_RC, a = q / WY_RC;
        Vc[rr * LDV + a] = (rr < rows && a < nbp) ? Vm[(r0 + rr) + (long long)(k0 + a) * n] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < OPT; ++u) {
        const int a = g + NG * u;
        if (a < EG_NB) {
          double x = acc[u];
#pragma unroll 8
          for (int rr = 0; rr < rows; ++rr) x = fma(Vc[rr * LDV + a], Zb[(r0 + rr) * LDZ + c], x);
          acc[u] = x;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < OPT; ++u) {
      const int a = g + NG * u;
      if (a < EG_NB) X[a * CB + c] = acc[u];
    }
    __syncthreads();
    // Y = T X (T upper triangular), kept in registers, then back into X
#pragma unroll
    for (int u = 0; u < OPT; ++u) {
      const int a = g + NG * u;
      double y = 0.0;
      if (a < nbp)
        for (int l = a; l < nbp; ++l) y = fma(T[a * LDV + l], X[l * CB + c], y);
      acc[u] = y;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < OPT; ++u) {
      const int a = g + NG * u;
      if (a < EG_NB) X[a * CB + c] = acc[u];
    }
    // Zb -= V Y
    for (int r0 = k0 + 1; r0 < n; r0 += WY_RC) {
      const int rows = min(WY_RC, n - r0);
      __syncthreads();
      for (int q = t; q < WY_RC * EG_NB; q += TH) {
        const int rr = q % WY_RC, a = q / WY_RC;
        Vc[rr * LDV + a] = (rr < rows && a < nbp) ? Vm[(r0 + rr) + (long long)(k0 + a) * n] : 0.0;
      }
      __syncthreads();
      for (int rr = g; rr < rows; rr += NG) {
        double z = Zb[(r0 + rr) * LDZ + c];
#pragma unroll 8
        for (int a = 0; a < nbp; ++a) z = fma(-Vc[rr * LDV + a], X[a * CB + c], z);
        Zb[(r0 + rr) * LDZ + c] = z;
      }
    }
    __syncthreads();
  }
  for (int cc = 0; cc < nc; ++cc)
    for (int r = t; r < n; r += TH) E[r + (long long)(c0 + cc) * n] = Zb[r * LDZ + cc];
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
  const size_t doubles = (size_t)count * (5 * nn + 4 * (size_t)n +
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
  double* d = Uf + count * nn;
  double* e = d + (size_t)count * n;
  double* tau = e + (size_t)count * n;
  double* wi = tau + (size_t)count * n;
  double* Tm = wi + (size_t)count * n;
  int* idesc = reinterpret_cast<int*>(Tm + (size_t)count * std::max(1, npan) * EG_NB * EG_NB);
  int* dinfo = idesc + 3 * n;
  const MergePlan& P = merge_plan(n);
  if (!P.desc.empty())
    TG_HIP(hipMemcpyAsync(idesc, P.desc.data(), sizeof(int) * P.desc.size(), hipMemcpyHostToDevice, s));
  {
    const int pb = n <= 384 ? 16 : (n <= 640 ? 8 : 4);
    const size_t lds1 = sizeof(double) * ((3 + 2 * (size_t)pb + TD_NS) * n + 2 * pb + TD_TH / 64);
    if (pb == 16)
      hipLaunchKernelGGL((tridiag_kernel<TD_TH, 16>), dim3(count), dim3(TD_TH), lds1, s, n, A, lda, sA,
                         Aw, Vm, d, e, tau);
    else if (pb == 8)
      hipLaunchKernelGGL((tridiag_kernel<TD_TH, 8>), dim3(count), dim3(TD_TH), lds1, s, n, A, lda, sA,
                         Aw, Vm, d, e, tau);
    else
      hipLaunchKernelGGL((tridiag_kernel<TD_TH, 4>), dim3(count), dim3(TD_TH), lds1, s, n, A, lda, sA,
                         Aw, Vm, d, e, tau);
  }
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
  auto apply_lds = [&](int cb) {
    return sizeof(double) * ((size_t)n * (cb + 1) + EG_NB * cb + EG_NB * (EG_NB + 1) +
                             WY_RC * (EG_NB + 1));
  };
  // eight columns per workgroup: n / 8 workgroups carry the blocks in parallel
  hipLaunchKernelGGL((wy_apply_kernel<EG_TH, 8>), dim3((n + 7) / 8, count), dim3(EG_TH), apply_lds(8), s,
                     n, npan, Vm, Tm, Z, E, sE);
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
