// Batched dense complex FP64 linear algebra for the gauge-angle loop:
//   * in-place Gauss-Jordan inverse with partial pivoting (plus determinant),
//     the role ZGETRF + ZGETRI play in calculate_thouless / calculate_overlap
//     (src/module_projection.f90:702-705, 860-883);
//   * Pfaffian of a complex skew-symmetric matrix by the Parlett-Reid LTL^T
//     reduction with the column pivoting of ZSKTF2, UPLO='U', MODE='P'
//     (src/subroutines_pfaffian.f:1694-2015 called by ZSKPFA :177-205).
// One 256-thread workgroup owns one matrix; the matrix stays in global memory
// (L2 / MALL resident), the pivot row and column of each step are staged in LDS.
#include "tg_internal.h"

#include <cstdlib>

namespace {

constexpr int NT = 256;

// argmax of |re|+|im| over v[i], i in [lo, hi), first index on ties (IZAMAX).
__device__ inline void block_argmax(double val, int idx, double* sv, int* si, double& best,
                                    int& besti) {
  const int t = threadIdx.x;
  sv[t] = val;
  si[t] = idx;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) {
      double v2 = sv[t + s];
      int i2 = si[t + s];
      if (v2 > sv[t] || (v2 == sv[t] && i2 < si[t])) {
        sv[t] = v2;
        si[t] = i2;
      }
    }
    __syncthreads();
  }
  best = sv[0];
  besti = si[0];
  __syncthreads();
}

__global__ void __launch_bounds__(NT)
gj_inverse_kernel(int n, cplx* __restrict__ Aall, int lda, long long sA, cplx* __restrict__ det,
                  int* __restrict__ info, int* __restrict__ perm_all) {
  extern __shared__ cplx sh[];
  cplx* rowk = sh;        // n
  cplx* colk = sh + n;    // n
  int* perm = reinterpret_cast<int*>(sh + 2 * n);   // n
  __shared__ double sv[NT];
  __shared__ int si[NT];
  cplx* a = Aall + (long long)blockIdx.x * sA;
  int* perm_out = perm_all + (long long)blockIdx.x * n;
  const int t = threadIdx.x;
  cplx d = cmk(1.0, 0.0);
  int inf = 0;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += NT) {
      double v = cabs1(a[i + (long long)k * lda]);
      if (v > best) { best = v; bi = i; }
    }
    double mx;
    int p;
    block_argmax(best, bi, sv, si, mx, p);
    // a vanishing (or non-finite) pivot column: singular, as zgetrf info = k
    const bool singular = !(mx > 0.0) || p < k || p >= n;
    if (singular) {
      if (inf == 0) inf = k + 1;
      p = k;
    }
    if (t == 0) perm[k] = p;
    if (p != k) {
      for (int j = t; j < n; j += NT) {
        cplx x = a[k + (long long)j * lda];
        a[k + (long long)j * lda] = a[p + (long long)j * lda];
        a[p + (long long)j * lda] = x;
      }
      d = cmk(-d.x, -d.y);
    }
    __syncthreads();
    if (singular) {
      d = cmk(0.0, 0.0);
      continue;
    }
    const cplx piv = a[k + (long long)k * lda];
    d = cmul(d, piv);
    const cplx pinv = cinv(piv);
    for (int j = t; j < n; j += NT) rowk[j] = (j == k) ? pinv : cmul(a[k + (long long)j * lda], pinv);
    for (int i = t; i < n; i += NT) colk[i] = (i == k) ? cmk(0.0, 0.0) : a[i + (long long)k * lda];
    __syncthreads();
    const long long nn = (long long)n * n;
    for (long long idx = t; idx < nn; idx += NT) {
      const int i = (int)(idx % n), j = (int)(idx / n);
      cplx* e = a + i + (long long)j * lda;
      if (i == k) {
        *e = rowk[j];
      } else {
        cplx v = (j == k) ? cmk(0.0, 0.0) : *e;
        const cplx f = colk[i], r = rowk[j];
        cfms(v, f, r);
        *e = v;
      }
    }
    __syncthreads();
  }
  // undo the row interchanges on the columns of the inverse
  for (int k = n - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k) {
      for (int i = t; i < n; i += NT) {
        cplx x = a[i + (long long)k * lda];
        a[i + (long long)k * lda] = a[i + (long long)p * lda];
        a[i + (long long)p * lda] = x;
      }
    }
    __syncthreads();
  }
  for (int k = t; k < n; k += NT) perm_out[k] = perm[k];
  if (t == 0) {
    det[blockIdx.x] = d;
    info[blockIdx.x] = inf;
  }
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inverse, one workgroup per matrix, thread r owns row r
// (n <= GJ_THREADS).  Columns are processed in panels of NB: the panel row of
// each thread lives in registers and is reduced with partial pivoting (pivot
// rows exchanged through LDS); the row interchanges and the accumulated
// elimination of the panel are then applied to every other column as one
// rank-NB update,  c <- P c + (T - E) (P c)[panel rows],  where T are the
// panel's transformed columns (the in-place inverse convention) and E the
// identity columns of the panel.  Traffic per matrix drops from n passes over
// the matrix to n/NB passes.  The panel rows of the other columns are staged in
// LDS a window at a time, so the update streams each thread's own row without
// further synchronisation.
// ---------------------------------------------------------------------------
constexpr int GJ_THREADS = 512;
constexpr int GJ_WIN = 320;   // columns staged per window (panel width x GJ_WIN complex in LDS)
constexpr int GJ_U = 8;       // columns per step of a thread's streaming update

__device__ inline void wave_argmax(double& v, int& i) {
  for (int off = 32; off > 0; off >>= 1) {
    double v2 = __shfl_xor(v, off, 64);
    int i2 = __shfl_xor(i, off, 64);
    if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
  }
}

// block-wide argmax (first index on ties) for TH threads
template <int TH>
__device__ inline void gj_argmax(double v, int i, double* wv, int* wi, double& bv, int& bi) {
  wave_argmax(v, i);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { wv[wid] = v; wi[wid] = i; }
  __syncthreads();
  bv = wv[0];
  bi = wi[0];
  for (int w = 1; w < TH / 64; ++w)
    if (wv[w] > bv || (wv[w] == bv && wi[w] < bi)) { bv = wv[w]; bi = wi[w]; }
  __syncthreads();
}

// TH threads per copy (n <= TH), RP copies, WIN columns per staged window
template <int GJ_NB, int RP, int TH, int WIN>
__global__ void __launch_bounds__(TH * RP)
gj_blocked_kernel(int n, cplx* __restrict__ Aall, int lda, long long sA, cplx* __restrict__ det,
                  int* __restrict__ info, int* __restrict__ perm_all) {
  __shared__ cplx sw[2][2][GJ_NB];
  __shared__ cplx Pw[GJ_NB][WIN + GJ_U];   // GJ_U columns of slack: no clamping below
  __shared__ int srcs[GJ_NB];
  __shared__ double wv[2][TH * RP / 64];
  __shared__ int wi[2][TH * RP / 64];
  __shared__ int perm[1024];
  // RP threads per row: thread tt handles row t; the copies (h = 1..RP-1) mirror
  // the panel reduction and share the column updates
  const int tt = threadIdx.x, t = tt % TH, h = tt / TH;
  const bool own = t < n;
  cplx* a = Aall + (long long)blockIdx.x * sA;
  cplx d = cmk(1.0, 0.0);
  int inf = 0;
  for (int k0 = 0; k0 < n; k0 += GJ_NB) {
    const int w = min(GJ_NB, n - k0);
    cplx pr[GJ_NB];
#pragma unroll
    for (int m = 0; m < GJ_NB; ++m)
      pr[m] = (own && m < w) ? a[t + (long long)(k0 + m) * lda] : cmk(0.0, 0.0);
    // ---- panel reduction in registers: two barriers per pivot (argmax, pivot
    // rows); the LDS areas alternate between consecutive pivots
    for (int mm = 0; mm < w; ++mm) {
      const int k = k0 + mm;
      const int buf = mm & 1;
      cplx pv = cmk(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < GJ_NB; ++m)
        if (m == mm) pv = pr[m];
      double v = (own && h == 0 && t >= k) ? cabs1(pv) : -1.0;
      int vi = (own && h == 0 && t >= k) ? t : n;
      wave_argmax(v, vi);
      if ((tt & 63) == 0) {
        wv[buf][tt >> 6] = v;
        wi[buf][tt >> 6] = vi;
      }
      __syncthreads();
      double mx = wv[buf][0];
      int p = wi[buf][0];
      for (int q = 1; q < TH * RP / 64; ++q)
        if (wv[buf][q] > mx || (wv[buf][q] == mx && wi[buf][q] < p)) {
          mx = wv[buf][q];
          p = wi[buf][q];
        }
      const bool singular = !(mx > 0.0) || p < k || p >= n;
      if (singular) {
        if (inf == 0) inf = k + 1;
        if (tt == 0) perm[k] = k;
        d = cmk(0.0, 0.0);
        continue;
      }
      if (tt == 0) perm[k] = p;
      if (h == 0 && t == p)
#pragma unroll
        for (int m = 0; m < GJ_NB; ++m) sw[buf][0][m] = pr[m];
      if (h == 0 && t == k && p != k)
#pragma unroll
        for (int m = 0; m < GJ_NB; ++m) sw[buf][1][m] = pr[m];
      __syncthreads();
      const cplx piv = sw[buf][0][mm];
      const cplx pinv = cinv(piv);
      if (p != k) d = cmk(-d.x, -d.y);
      d = cmul(d, piv);
      if (t == p && p != k)
#pragma unroll
        for (int m = 0; m < GJ_NB; ++m) pr[m] = sw[buf][1][m];
      if (own) {
        // pivot row k: r / pivot (pivot column 1 / pivot); other rows:
        // row - (f / pivot) r (pivot column -f / pivot).  One code path:
        // g = -1/pivot and a zero base for row k, g = f/pivot otherwise.
        cplx f = cmk(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < GJ_NB; ++m)
          if (m == mm) f = pr[m];
        const bool pivrow = (t == k);
        const cplx g = pivrow ? cmk(-pinv.x, -pinv.y) : cmul(f, pinv);
#pragma unroll
        for (int m = 0; m < GJ_NB; ++m) {
          if (m == mm) {
            pr[m] = cmk(-g.x, -g.y);
          } else {
            const cplx r = sw[buf][0][m];
            const cplx b = pivrow ? cmk(0.0, 0.0) : pr[m];
            cplx v = b;
            cfms(v, g, r);
            pr[m] = v;
          }
        }
      }
    }
    __syncthreads();   // perm of the panel visible to every thread
#pragma unroll
    for (int m = 0; m < GJ_NB; ++m)
      if (own && h == 0 && m < w) a[t + (long long)(k0 + m) * lda] = pr[m];
    // ---- source row of this thread after the panel's interchanges
    int src = t;
    for (int m = w - 1; m >= 0; --m) {
      const int kk = k0 + m, pp = perm[kk];
      if (src == kk) src = pp;
      else if (src == pp) src = kk;
    }
    const int blk = t - k0;   // position of this row inside the panel, if any
    const bool inblk = own && blk >= 0 && blk < w;
    if (inblk && h == 0) srcs[blk] = src;
    __syncthreads();
    // ---- interchanges and rank-w update of the other columns, GJ_WIN at a time.
    // Rows moved out of the panel come from panel positions (each position is
    // swapped once, with a row at or below it), so after the panel source rows
    // are staged in LDS every thread only reads and writes its own row.
    const int nrest = n - w;
    for (int c0 = 0; c0 < nrest; c0 += WIN) {
      const int cw = min(WIN, nrest - c0);
      for (int e = tt; e < w * cw; e += TH * RP) {
        const int m = e % w, c = e / w;
        const int jj = c0 + c;
        Pw[m][c] = a[srcs[m] + (long long)((jj < k0) ? jj : jj + w) * lda];
      }
      __syncthreads();
      if (own && !inblk && src != t)
        for (int c = h; c < cw; c += RP) {
          const int jj = c0 + c;
          const long long j = (jj < k0) ? jj : jj + w;
          a[t + j * lda] = a[src + j * lda];
        }
      __syncthreads();
      if (own) {
        for (int c = h * GJ_U; c < cw; c += GJ_U * RP) {
          cplx acc[GJ_U];
#pragma unroll
          for (int u = 0; u < GJ_U; ++u) {
            const int jj = c0 + c + u;
            acc[u] = (!inblk && c + u < cw) ? a[t + (long long)((jj < k0) ? jj : jj + w) * lda]
                                            : cmk(0.0, 0.0);
          }
#pragma unroll
          for (int m = 0; m < GJ_NB; ++m) {
            if (m < w) {
              const cplx pm = pr[m];
              const cplx* prow = &Pw[m][c];     // columns past cw are discarded
#pragma unroll
              for (int u = 0; u < GJ_U; ++u) {
                const cplx b = prow[u];
                cfma(acc[u], pm, b);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < GJ_U; ++u) {
            const int jj = c0 + c + u;
            if (c + u < cw) a[t + (long long)((jj < k0) ? jj : jj + w) * lda] = acc[u];
          }
        }
      }
      __syncthreads();
    }
  }
  // undo the row interchanges on the columns of the inverse (own row only)
  if (own && h == 0)
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        cplx x = a[t + (long long)k * lda];
        a[t + (long long)k * lda] = a[t + (long long)p * lda];
        a[t + (long long)p * lda] = x;
      }
    }
  for (int k = tt; k < n; k += TH * RP) perm_all[(long long)blockIdx.x * n + k] = perm[k];
  if (tt == 0) {
    det[blockIdx.x] = d;
    info[blockIdx.x] = inf;
  }
}

__global__ void __launch_bounds__(NT)
pfaffian_kernel(int n, const cplx* __restrict__ Aall, int lda, long long sA,
                cplx* __restrict__ work, double* __restrict__ logabs, cplx* __restrict__ phase,
                int* __restrict__ zero) {
  extern __shared__ cplx sh[];
  cplx* xs = sh;       // column k
  cplx* ys = sh + n;   // column k-1
  __shared__ double sv[NT];
  __shared__ int si[NT];
  const int t = threadIdx.x;
  const cplx* A = Aall + (long long)blockIdx.x * sA;
  cplx* W = work + (long long)blockIdx.x * n * n;
  const long long nn = (long long)n * n;
  // full skew copy of the upper triangle
  for (long long idx = t; idx < nn; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    cplx v = cmk(0.0, 0.0);
    if (i < j) v = A[i + (long long)j * lda];
    else if (i > j) { v = A[j + (long long)i * lda]; v = cmk(-v.x, -v.y); }
    W[idx] = v;
  }
  __syncthreads();
  double la = 0.0;
  cplx ph = cmk(1.0, 0.0);
  int z = 0;
  for (int k = n - 1; k >= 1; k -= 2) {
    const int kk = k - 1;
    double best = -1.0;
    int bi = n;
    for (int i = t; i < k; i += NT) {
      double v = cabs1(W[i + (long long)k * n]);
      if (v > best) { best = v; bi = i; }
    }
    double mx;
    int kp;
    block_argmax(best, bi, sv, si, mx, kp);
    // zero (or non-finite) pivot column: ZSKTF2 INFO > 0, ZSKPFA returns 0
    if (!(mx > 0.0) || kp < 0 || kp >= k) { z = 1; break; }
    if (kp != kk) {
      // symmetric interchange kp <-> kk on rows/columns 0..k
      for (int j = t; j <= k; j += NT) {
        cplx x = W[kk + (long long)j * n];
        W[kk + (long long)j * n] = W[kp + (long long)j * n];
        W[kp + (long long)j * n] = x;
      }
      __syncthreads();
      for (int i = t; i <= k; i += NT) {
        cplx x = W[i + (long long)kk * n];
        W[i + (long long)kk * n] = W[i + (long long)kp * n];
        W[i + (long long)kp * n] = x;
      }
      ph = cmk(-ph.x, -ph.y);
      __syncthreads();
    }
    const cplx p = W[kk + (long long)k * n];
    const double ap = sqrt(p.x * p.x + p.y * p.y);
    la += log(ap);
    ph = cmul(ph, cmk(p.x / ap, p.y / ap));
    const cplx pinv = cinv(p);
    for (int i = t; i < kk; i += NT) {
      xs[i] = cmul(W[i + (long long)k * n], pinv);
      ys[i] = W[i + (long long)kk * n];
    }
    __syncthreads();
    // W(0:kk,0:kk) += (x y^T - y x^T) / p
    const long long m2 = (long long)kk * kk;
    for (long long idx = t; idx < m2; idx += NT) {
      const int i = (int)(idx % kk), j = (int)(idx / kk);
      const cplx xi = xs[i], yj = ys[j], yi = ys[i], xj = xs[j];
      cplx* e = W + i + (long long)j * n;
      cplx v = *e;
      cfma(v, xi, yj);
      cfms(v, yi, xj);
      *e = v;
    }
    __syncthreads();
  }
  if (t == 0) {
    logabs[blockIdx.x] = z ? -INFINITY : la;
    phase[blockIdx.x] = z ? cmk(0.0, 0.0) : ph;
    zero[blockIdx.x] = z;
  }
}

// ---------------------------------------------------------------------------
// Blocked Parlett-Reid Pfaffian (n <= PF_THREADS), one workgroup per matrix,
// thread i owns row i of the full skew copy W.  Pair steps are grouped in
// panels of PF_PS: within a panel the rank-2 updates
//   W(0:kk,0:kk) += x y^T - y x^T,  x = W(:,k)/W(kk,k), y = W(:,kk)
// are kept as pending vectors X(i,s), Y(i,s) in the registers of row i; the
// columns a step needs are formed on the fly (global column + pending terms)
// and the symmetric interchange moves the pending rows with the matrix rows.
// After the panel the leading block receives one rank-2*PF_PS update.  Pivoting
// is that of ZSKTF2: IZAMAX over W(0:k-1, k), first index on ties.
// ---------------------------------------------------------------------------
constexpr int PF_THREADS = 512;
constexpr int PF_PS = 8;
constexpr int PF_WIN = 256;   // columns whose pending vectors are staged in LDS at once

// TH threads (n <= TH), WIN columns per staged window.  (A cap at 168 VGPRs for
// TH = 384, so that GEMM tiles could share its CUs, paid while the Pfaffians ran
// beside the particle-hole contraction; since they start beside the inverse the
// uncapped kernel makes the n = 336 step 0.5 ms shorter.)
template <int TH, int WIN>
__global__ void __launch_bounds__(TH)
pfaffian_blocked_kernel(int n, const cplx* __restrict__ Aall, int lda, long long sA,
                        cplx* __restrict__ work, double* __restrict__ logabs,
                        cplx* __restrict__ phase, int* __restrict__ zero) {
  __shared__ cplx xk[PF_PS], yk[PF_PS];
  __shared__ cplx swx[2][PF_PS], swy[2][PF_PS];
  __shared__ cplx swc[2];
  __shared__ cplx cx[WIN][PF_PS], cy[WIN][PF_PS];
  __shared__ double wv[TH / 64];
  __shared__ int wi[TH / 64];
  __shared__ cplx spiv;
  const int t = threadIdx.x;
  const bool own = t < n;
  const cplx* A = Aall + (long long)blockIdx.x * sA;
  cplx* W = work + (long long)blockIdx.x * n * n;
  const long long nn = (long long)n * n;
  for (long long idx = t; idx < nn; idx += TH) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    cplx v = cmk(0.0, 0.0);
    if (i < j) v = A[i + (long long)j * lda];
    else if (i > j) { v = A[j + (long long)i * lda]; v = cmk(-v.x, -v.y); }
    W[idx] = v;
  }
  __syncthreads();
  double la = 0.0;
  cplx ph = cmk(1.0, 0.0);
  int z = 0;
  cplx X[PF_PS], Y[PF_PS];
  int k_top = n - 1;
  while (k_top >= 1) {
#pragma unroll
    for (int s2 = 0; s2 < PF_PS; ++s2) { X[s2] = cmk(0.0, 0.0); Y[s2] = cmk(0.0, 0.0); }
    int ns = 0, m = k_top + 1;
    for (int s = 0; s < PF_PS; ++s) {
      const int k = k_top - 2 * s;
      if (k < 1) break;
      const int kk = k - 1;
      // pending vectors of row k
      if (t == k)
#pragma unroll
        for (int s2 = 0; s2 < PF_PS; ++s2) { xk[s2] = X[s2]; yk[s2] = Y[s2]; }
      __syncthreads();
      cplx cik = cmk(0.0, 0.0);
      if (own && t < k) {
        cik = W[t + (long long)k * n];
#pragma unroll
        for (int s2 = 0; s2 < PF_PS; ++s2)
          if (s2 < s) {
            cfma(cik, X[s2], yk[s2]);
            cfms(cik, Y[s2], xk[s2]);
          }
      }
      double v = (own && t < k) ? cabs1(cik) : -1.0;
      int vi = (own && t < k) ? t : n;
      double mx;
      int kp;
      gj_argmax<TH>(v, vi, wv, wi, mx, kp);
      if (!(mx > 0.0) || kp < 0 || kp >= k) { z = 1; break; }
      if (kp != kk) {
        if (t == kp) {
#pragma unroll
          for (int s2 = 0; s2 < PF_PS; ++s2) { swx[0][s2] = X[s2]; swy[0][s2] = Y[s2]; }
          swc[0] = cik;
        }
        if (t == kk) {
#pragma unroll
          for (int s2 = 0; s2 < PF_PS; ++s2) { swx[1][s2] = X[s2]; swy[1][s2] = Y[s2]; }
          swc[1] = cik;
        }
        // rows kp <-> kk of the leading columns
        for (int j = t; j < k; j += TH) {
          cplx* r1 = W + kp + (long long)j * n;
          cplx* r2 = W + kk + (long long)j * n;
          const cplx x1 = *r1;
          *r1 = *r2;
          *r2 = x1;
        }
        __syncthreads();
        if (t == kp) {
#pragma unroll
          for (int s2 = 0; s2 < PF_PS; ++s2) { X[s2] = swx[1][s2]; Y[s2] = swy[1][s2]; }
          cik = swc[1];
        }
        if (t == kk) {
#pragma unroll
          for (int s2 = 0; s2 < PF_PS; ++s2) { X[s2] = swx[0][s2]; Y[s2] = swy[0][s2]; }
          cik = swc[0];
        }
        // columns kp <-> kk of the leading rows
        if (own && t < k) {
          cplx* c1 = W + t + (long long)kp * n;
          cplx* c2 = W + t + (long long)kk * n;
          const cplx x1 = *c1;
          *c1 = *c2;
          *c2 = x1;
        }
        ph = cmk(-ph.x, -ph.y);
        __syncthreads();
      }
      if (t == kk) {
        spiv = cik;
#pragma unroll
        for (int s2 = 0; s2 < PF_PS; ++s2) { xk[s2] = X[s2]; yk[s2] = Y[s2]; }
      }
      __syncthreads();
      const cplx p = spiv;
      const double ap = sqrt(p.x * p.x + p.y * p.y);
      la += log(ap);
      ph = cmul(ph, cmk(p.x / ap, p.y / ap));
      const cplx pinv = cinv(p);
      cplx xn = cmk(0.0, 0.0), yn = cmk(0.0, 0.0);
      if (own && t < kk) {
        cplx ci = W[t + (long long)kk * n];
#pragma unroll
        for (int s2 = 0; s2 < PF_PS; ++s2)
          if (s2 < s) {
            cfma(ci, X[s2], yk[s2]);
            cfms(ci, Y[s2], xk[s2]);
          }
        xn = cmul(cik, pinv);
        yn = ci;
      }
#pragma unroll
      for (int s2 = 0; s2 < PF_PS; ++s2)
        if (s2 == s) { X[s2] = xn; Y[s2] = yn; }
      ns = s + 1;
      m = kk;
      __syncthreads();
    }
    if (z) break;
    // rank-2*ns update of the leading block [0, m): the pending vectors of
    // WIN columns are staged in LDS, then every thread streams its own row
    for (int j0 = 0; j0 < m; j0 += WIN) {
      const int cw = min(WIN, m - j0);
      if (t >= j0 && t < j0 + cw)
#pragma unroll
        for (int s2 = 0; s2 < PF_PS; ++s2) { cx[t - j0][s2] = X[s2]; cy[t - j0][s2] = Y[s2]; }
      __syncthreads();
      if (own && t < m) {
        for (int c = 0; c < cw; ++c) {
          const long long j = j0 + c;
          cplx acc = W[t + j * n];
#pragma unroll
          for (int s2 = 0; s2 < PF_PS; ++s2)
            if (s2 < ns) {
              cfma(acc, X[s2], cy[c][s2]);
              cfms(acc, Y[s2], cx[c][s2]);
            }
          W[t + j * n] = acc;
        }
      }
      __syncthreads();
    }
    k_top = m - 1;
  }
  if (t == 0) {
    logabs[blockIdx.x] = z ? -INFINITY : la;
    phase[blockIdx.x] = z ? cmk(0.0, 0.0) : ph;
    zero[blockIdx.x] = z;
  }
}

}  // namespace

int batched_inverse(hipStream_t s, int n, cplx* A, int lda, long long sA, int batch, cplx* det,
                    int* info, int* perm) {
  if (n <= 0 || batch <= 0) return TG_OK;
  if (n <= GJ_THREADS) {
    // one thread per row in the smallest workgroup that holds the rows (the
    // LDS window shrinks with it, so small matrices share a CU); TG_GJ=2: two
    // threads per row (the second mirrors the panel reduction and takes half
    // of the column updates), n <= 384
    static const int rp = [] {
      const char* e = std::getenv("TG_GJ");
      return (e && std::atoi(e) == 2) ? 2 : 1;
    }();
#define TG_GJ_LAUNCH(RP, TH, WIN)                                                              \
  hipLaunchKernelGGL((gj_blocked_kernel<16, RP, TH, WIN>), dim3(batch), dim3(TH * RP), 0, s, n, A, \
                     lda, sA, det, info, perm)
    if (rp == 2 && n <= 384)
      TG_GJ_LAUNCH(2, 384, GJ_WIN);
    else if (n <= 64)
      TG_GJ_LAUNCH(1, 64, 64);
    else if (n <= 128)
      TG_GJ_LAUNCH(1, 128, 128);
    else if (n <= 256)
      TG_GJ_LAUNCH(1, 256, 256);
    else if (n <= 384)
      TG_GJ_LAUNCH(1, 384, GJ_WIN);
    else
      TG_GJ_LAUNCH(1, GJ_THREADS, GJ_WIN);
#undef TG_GJ_LAUNCH
    TG_HIP(hipGetLastError());
    return TG_OK;
  }
  size_t shm = 2 * (size_t)n * sizeof(cplx) + (size_t)n * sizeof(int);
  hipLaunchKernelGGL(gj_inverse_kernel, dim3(batch), dim3(NT), shm, s, n, A, lda, sA, det, info,
                     perm);
  TG_HIP(hipGetLastError());
  return TG_OK;
}

int batched_pfaffian(hipStream_t s, int n, const cplx* A, int lda, long long sA, int batch,
                     cplx* work, double* logabs, cplx* phase, int* zero) {
  if (batch <= 0) return TG_OK;
  if (n % 2 != 0 || n == 0) {
    tg_set_error("batched_pfaffian: n must be even and positive here");
    return TG_E_ARG;
  }
  if (n <= PF_THREADS) {
    // one thread per row in the smallest workgroup that holds the rows
#define TG_PF_LAUNCH(TH, WIN)                                                                 \
  hipLaunchKernelGGL((pfaffian_blocked_kernel<TH, WIN>), dim3(batch), dim3(TH), 0, s, n, A, lda, \
                     sA, work, logabs, phase, zero)
    if (n <= 64)
      TG_PF_LAUNCH(64, 64);
    else if (n <= 128)
      TG_PF_LAUNCH(128, 128);
    else if (n <= 256)
      TG_PF_LAUNCH(256, PF_WIN);
    else if (n <= 384)
      TG_PF_LAUNCH(384, PF_WIN);
    else
      TG_PF_LAUNCH(PF_THREADS, PF_WIN);
#undef TG_PF_LAUNCH
    TG_HIP(hipGetLastError());
    return TG_OK;
  }
  size_t shm = 2 * (size_t)n * sizeof(cplx);
  hipLaunchKernelGGL(pfaffian_kernel, dim3(batch), dim3(NT), shm, s, n, A, lda, sA, work, logabs,
                     phase, zero);
  TG_HIP(hipGetLastError());
  return TG_OK;
}
