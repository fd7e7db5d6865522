// Batched dense complex FP64 linear algebra for the gauge-angle loop:
//   * in-place Gauss-Jordan inverse with partial pivoting (plus determinant),
//     the role ZGETRF + ZGETRI play in calculate_thouless / calculate_overlap
//     (src/module_projection.f90:702-705, 860-883);
//   * Pfaffian of a complex skew-symmetric matrix by the Parlett-Reid LTL^T
//     reduction with the column pivoting of ZSKTF2, UPLO='U', MODE='P'
//     (src/subroutines_pfaffian.f:1694-2015 called by ZSKPFA :177-205).
// One workgroup owns one matrix, which stays in global memory; pivot rows,
// panel source rows and matrix-core operands are staged in LDS.
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

// One DPP step of the argmax below: (v, i) of the lane the row-level move CTRL
// brings in (lanes it brings nothing to keep their own), the larger v and then
// the smaller i kept.
template <int CTRL, int RM>
__device__ __forceinline__ void gj_dpp_argmax_step(double& v, int& i) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(unsigned long long)b, hi = (int)(b >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, 0xf, false);
  const int i2 = __builtin_amdgcn_update_dpp(i, i, CTRL, RM, 0xf, false);
  const double v2 =
      __longlong_as_double((long long)(((unsigned long long)(unsigned)hi2 << 32) | (unsigned)lo2));
  const bool take = (v2 > v) | ((v2 == v) & (i2 < i));   // no short circuit: no branch
  v = take ? v2 : v;
  i = take ? i2 : i;
}

// argmax (larger v, then smaller i: the first maximum) over lanes 0 .. W-1 of a
// wave, every lane receiving the result.  Row shifts by 1, 2, 4, 8 leave the
// maximum of lanes [16 r, l] in lane l of row r; the row broadcasts of lanes 15
// and 31 then carry rows 0-1 into lane 31 and all four rows into lane 63.
template <int W>
__device__ __forceinline__ void gj_wave_argmax_dpp(double& v, int& i) {
  static_assert(W == 64 || W <= 16, "a whole wave or lanes of its first row");
  gj_dpp_argmax_step<0x111, 0xf>(v, i);   // row_shr:1
  if constexpr (W > 2) gj_dpp_argmax_step<0x112, 0xf>(v, i);   // row_shr:2
  if constexpr (W > 4) gj_dpp_argmax_step<0x114, 0xf>(v, i);   // row_shr:4
  if constexpr (W > 8) gj_dpp_argmax_step<0x118, 0xf>(v, i);   // row_shr:8
  constexpr int last = (W == 64) ? 63 : W - 1;
  if constexpr (W == 64) {
    gj_dpp_argmax_step<0x142, 0xa>(v, i);   // row_bcast:15 into rows 1 and 3
    gj_dpp_argmax_step<0x143, 0xc>(v, i);   // row_bcast:31 into rows 2 and 3
  }
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)b, last);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), last);
  v = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  i = __builtin_amdgcn_readlane(i, last);
}

// argmax over the wave (first index on ties), every lane receiving it
__device__ inline void wave_argmax(double& v, int& i) { gj_wave_argmax_dpp<64>(v, i); }

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
      // lane q of every wave takes wave q's maximum, one more reduction over them
      double mx = -1.0;
      int p = n;
      if ((tt & 63) < TH * RP / 64) {
        mx = wv[buf][tt & 63];
        p = wi[buf][tt & 63];
      }
      gj_wave_argmax_dpp<TH * RP / 64>(mx, p);
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

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse with 32-column panels, one thread per row (n <= TH).
// A thread holds its row of the panel as two 16-column register halves A and
// B.  Half A is reduced as in gj_blocked_kernel; B's columns then receive A's
// interchanges and rank-16 update (16 x 16 source values through LDS) and are
// reduced in turn; B's interchanges reach the A half through two LDS tables
// (the A half of the block-B rows, and of the rows moved into block B).  The
// other columns then take one rank-32 update, c <- P c + (T - E)(P c)[panel],
// so a matrix is streamed n/32 times instead of n/16.  While a group of U
// columns is updated, the next group's loads are in flight.  The pivots are
// those of gj_blocked_kernel (first maximum of |Re| + |Im|, as IZAMAX in
// ZGETRF).
// ---------------------------------------------------------------------------
template <int TH>
__device__ __forceinline__ void gj_reduce16(cplx (&pr)[16], int k0, int w, int n, int t,
                                            cplx (*sw)[2][16], double (*wv)[TH / 64],
                                            int (*wi)[TH / 64], int* perm, cplx& d, int& inf) {
  const bool own = t < n;
  for (int mm = 0; mm < w; ++mm) {
    const int k = k0 + mm;
    const int buf = mm & 1;
    cplx pv = cmk(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < 16; ++m)
      if (m == mm) pv = pr[m];
    double v = (own && t >= k) ? cabs1(pv) : -1.0;
    int vi = (own && t >= k) ? t : n;
    gj_wave_argmax_dpp<64>(v, vi);
    if ((t & 63) == 0) {
      wv[buf][t >> 6] = v;
      wi[buf][t >> 6] = vi;
    }
    __syncthreads();
    // lane q of every wave takes wave q's maximum, one more reduction over them
    double mx = -1.0;
    int p = n;
    if ((t & 63) < TH / 64) {
      mx = wv[buf][t & 63];
      p = wi[buf][t & 63];
    }
    gj_wave_argmax_dpp<TH / 64>(mx, p);
    if (!(mx > 0.0) || p < k || p >= n) {
      if (inf == 0) inf = k + 1;
      if (t == 0) perm[k] = k;
      d = cmk(0.0, 0.0);
      continue;
    }
    if (t == 0) perm[k] = p;
    if (t == p)
#pragma unroll
      for (int m = 0; m < 16; ++m) sw[buf][0][m] = pr[m];
    if (t == k && p != k)
#pragma unroll
      for (int m = 0; m < 16; ++m) sw[buf][1][m] = pr[m];
    __syncthreads();
    const cplx piv = sw[buf][0][mm];
    const cplx pinv = cinv(piv);
    if (p != k) d = cmk(-d.x, -d.y);
    d = cmul(d, piv);
    if (t == p && p != k)
#pragma unroll
      for (int m = 0; m < 16; ++m) pr[m] = sw[buf][1][m];
    if (own) {
      cplx f = cmk(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < 16; ++m)
        if (m == mm) f = pr[m];
      const bool pivrow = (t == k);
      const cplx g = pivrow ? cmk(-pinv.x, -pinv.y) : cmul(f, pinv);
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if (m == mm) {
          pr[m] = cmk(-g.x, -g.y);
        } else {
          const cplx r = sw[buf][0][m];
          cplx v2 = pivrow ? cmk(0.0, 0.0) : pr[m];
          cfms(v2, g, r);
          pr[m] = v2;
        }
      }
    }
  }
}

// source row of row t under the interchanges perm[k0 .. k0 + w)
__device__ __forceinline__ int gj_source(const int* perm, int k0, int w, int t) {
  int src = t;
  for (int m = w - 1; m >= 0; --m) {
    const int kk = k0 + m, pp = perm[kk];
    if (src == kk) src = pp;
    else if (src == pp) src = kk;
  }
  return src;
}

typedef double gj_d4 __attribute__((ext_vector_type(4)));
constexpr int GJ_MFQ = 3;   // row blocks per wave in the matrix-core update (TH = 512)

// MF (TH = 512, n <= 384): the rank-32 update of the other columns as 16 x 16
// tiles of v_mfma_f64_16x16x4f64 (U = 16 then: one column tile of slack).  The
// tile computes C^T = Pw^T W^T, so a lane holds 16 consecutive rows of four
// columns (coalesced loads and stores) and each LDS operand serves 48 rows
// instead of one; with the next tile's loads in flight during the products,
// eight waves keep about four times the bytes of the scalar update in flight.
template <int TH, int WIN, int U, bool MF>
__global__ void __launch_bounds__(TH)
gj_wide_kernel(int n, cplx* __restrict__ Aall, int lda, long long sA, cplx* __restrict__ det,
               int* __restrict__ info, int* __restrict__ perm_all) {
  constexpr int H = 16;
  __shared__ cplx sw[2][2][H];
  __shared__ cplx Pw[2 * H][WIN + U];   // U slack columns: copies of the window's last
  __shared__ cplx tab[2][H][H];
  __shared__ int srcs[2 * H];
  __shared__ double wv[2][TH / 64];
  __shared__ int wi[2][TH / 64];
  __shared__ int perm[TH];
  const int t = threadIdx.x;
  const bool own = t < n;
  cplx* a = Aall + (long long)blockIdx.x * sA;
  cplx d = cmk(1.0, 0.0);
  int inf = 0;
  for (int k0 = 0; k0 < n; k0 += 2 * H) {
    const int w = min(2 * H, n - k0), wA = min(H, w), wB = w - wA, kB = k0 + H;
    cplx pA[H], pB[H];
#pragma unroll
    for (int m = 0; m < H; ++m)
      pA[m] = (own && m < wA) ? a[t + (long long)(k0 + m) * lda] : cmk(0.0, 0.0);
    gj_reduce16<TH>(pA, k0, wA, n, t, sw, wv, wi, perm, d, inf);
#pragma unroll
    for (int m = 0; m < H; ++m) pB[m] = cmk(0.0, 0.0);
    if (wB > 0) {
      __syncthreads();   // perm of half A
      const int srcA = gj_source(perm, k0, wA, t);
      const bool inA = own && t >= k0 && t < k0 + wA;
      if (inA) srcs[t - k0] = srcA;
      __syncthreads();
      // (P_A c)[A rows] of B's columns (zero past wB), and this row's P_A c
      // (zero in block A); wA = H here
      for (int e = t; e < H * H; e += TH) {
        const int m = e % H, c = e / H;
        Pw[m][c] = (c < wB) ? a[srcs[m] + (long long)(kB + c) * lda] : cmk(0.0, 0.0);
      }
#pragma unroll
      for (int m = 0; m < H; ++m)
        if (own && !inA && m < wB) pB[m] = a[srcA + (long long)(kB + m) * lda];
      __syncthreads();
      if (own) {
#pragma unroll
        for (int m = 0; m < H; ++m) {
          const cplx am = pA[m];
#pragma unroll
          for (int c = 0; c < H; ++c) cfma(pB[c], am, Pw[m][c]);   // rows m >= wA are zero
        }
      }
      gj_reduce16<TH>(pB, kB, wB, n, t, sw, wv, wi, perm, d, inf);
      __syncthreads();   // perm of half B
      // B's interchanges and rank-16 update on the A half
      const int srcB = gj_source(perm, kB, wB, t);
      const bool inB = own && t >= kB && t < kB + wB;
      if (inB) {
        srcs[H + t - kB] = srcB;
#pragma unroll
        for (int m = 0; m < H; ++m) tab[0][t - kB][m] = pA[m];
      }
      __syncthreads();
      if (own && t >= kB)
        for (int q = 0; q < wB; ++q)
          if (srcs[H + q] == t) {
#pragma unroll
            for (int m = 0; m < H; ++m) tab[1][q][m] = pA[m];
          }
      for (int e = t; e < (H - wB) * H; e += TH) tab[1][wB + e / H][e % H] = cmk(0.0, 0.0);
      __syncthreads();
      if (own) {
        if (inB) {
#pragma unroll
          for (int m = 0; m < H; ++m) pA[m] = cmk(0.0, 0.0);
        } else if (srcB != t) {
#pragma unroll
          for (int m = 0; m < H; ++m) pA[m] = tab[0][srcB - kB][m];
        }
#pragma unroll
        for (int q = 0; q < H; ++q) {   // pB[q] = 0 for q >= wB
          const cplx bq = pB[q];
#pragma unroll
          for (int m = 0; m < H; ++m) cfma(pA[m], bq, tab[1][q][m]);
        }
      }
    }
    __syncthreads();   // perm of the panel; LDS tables read
#pragma unroll
    for (int m = 0; m < H; ++m) {
      if (own && m < wA) a[t + (long long)(k0 + m) * lda] = pA[m];
      if (own && m < wB) a[t + (long long)(kB + m) * lda] = pB[m];
    }
    const int src = gj_source(perm, k0, w, t);
    const int blk = t - k0;
    const bool inblk = own && blk >= 0 && blk < w;
    if (inblk) srcs[blk] = src;
    // MF: matrix-core operands of W.  Wave v updates row blocks v + 8 q (q < 3,
    // rows 16 (v + 8 q) + li), so every SIMD holds six of the 24 blocks of
    // n <= 384; fr/fi[q][kk] = W(16 (v + 8 q) + li, 4 kk + lq), moved between
    // waves through a TH x 16 slab of the (idle) Pw area per half
    const int lane = t & 63, li = lane & 15, lq = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    double fr[GJ_MFQ][8], fi[GJ_MFQ][8];
    if constexpr (MF) {
      cplx* slab = &Pw[0][0];
#pragma unroll
      for (int m = 0; m < H; ++m) slab[m * TH + t] = pA[m];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < GJ_MFQ; ++q)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const cplx v = slab[(4 * kk + lq) * TH + 16 * (wv + 8 * q) + li];
          fr[q][kk] = v.x;
          fi[q][kk] = v.y;
        }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < H; ++m) slab[m * TH + t] = pB[m];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < GJ_MFQ; ++q)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const cplx v = slab[(4 * kk + lq) * TH + 16 * (wv + 8 * q) + li];
          fr[q][4 + kk] = v.x;
          fi[q][4 + kk] = v.y;
        }
    }
    __syncthreads();
    // interchanges and rank-w update of the other columns, WIN at a time; Pw row
    // m < wA is panel position k0 + m, row H + q position kB + q
    const int nrest = n - w;
    for (int c0 = 0; c0 < nrest; c0 += WIN) {
      const int cw = min(WIN, nrest - c0);
      // all 2H rows are written (rows past the panel width as zeros), so the
      // update below needs no bounds on m; the U slack columns repeat the last
      // column, so a group that runs past the window computes that column again
      for (int e = t; e < 2 * H * (cw + U); e += TH) {
        const int m2 = e % (2 * H), c = e / (2 * H);
        const int pos = (m2 < H) ? (m2 < wA ? m2 : -1) : (m2 - H < wB ? m2 : -1);
        const int jj = c0 + min(c, cw - 1);
        Pw[m2][c] = (pos >= 0) ? a[srcs[pos] + (long long)((jj < k0) ? jj : jj + w) * lda]
                               : cmk(0.0, 0.0);
      }
      __syncthreads();
      if (own && !inblk && src != t)
        for (int c = 0; c < cw; ++c) {
          const int jj = c0 + c;
          const long long j = (jj < k0) ? jj : jj + w;
          a[t + j * lda] = a[src + j * lda];
        }
      __syncthreads();
      if constexpr (MF) {
        // this lane's elements of tile jt: rows 16 (wv + 8 q) + li, columns
        // jt + lq + 4 r (loads clamped into the matrix and the window; stored
        // only inside both); block rows start from zero.  The next tile's
        // values are loaded while the current tile's products run.
        int rowc[GJ_MFQ];
        bool ld[GJ_MFQ], st[GJ_MFQ];
#pragma unroll
        for (int q = 0; q < GJ_MFQ; ++q) {
          const int row = 16 * (wv + 8 * q) + li;
          rowc[q] = min(row, n - 1);
          st[q] = row < n;
          ld[q] = row < n && !(row >= k0 && row < k0 + w);
        }
        auto colo = [&](int c) -> long long {
          const int jj = c0 + min(c, cw - 1);
          return (long long)((jj < k0) ? jj : jj + w) * lda;
        };
        cplx nx[GJ_MFQ][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long long o = colo(lq + 4 * r);
#pragma unroll
          for (int q = 0; q < GJ_MFQ; ++q) nx[q][r] = a[rowc[q] + o];
        }
#pragma unroll 1
        for (int jt = 0; jt < cw; jt += 16) {
          // three real products per complex one (Gauss): t1 = C_re + sum Pr Wr,
          // t2 = sum Pi Wi, t3 = C_re + C_im + sum (Pr + Pi)(Wr + Wi); then
          // C = (t1 - t2, t3 - t1 - t2)
          gj_d4 t1[GJ_MFQ], t2[GJ_MFQ], t3[GJ_MFQ];
#pragma unroll
          for (int q = 0; q < GJ_MFQ; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double cr = ld[q] ? nx[q][r].x : 0.0, ci = ld[q] ? nx[q][r].y : 0.0;
              t1[q][r] = cr;
              t2[q][r] = 0.0;
              t3[q][r] = cr + ci;
            }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const long long o = colo(jt + 16 + lq + 4 * r);
#pragma unroll
            for (int q = 0; q < GJ_MFQ; ++q) nx[q][r] = a[rowc[q] + o];
          }
#pragma unroll
          for (int kk = 0; kk < 8; ++kk) {
            const cplx pv = Pw[4 * kk + lq][jt + li];
            const double ps = pv.x + pv.y;
#pragma unroll
            for (int q = 0; q < GJ_MFQ; ++q) {
              const double ws = fr[q][kk] + fi[q][kk];
              t1[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv.x, fr[q][kk], t1[q], 0, 0, 0);
              t2[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv.y, fi[q][kk], t2[q], 0, 0, 0);
              t3[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ps, ws, t3[q], 0, 0, 0);
            }
          }
          gj_d4 dr[GJ_MFQ], di[GJ_MFQ];
#pragma unroll
          for (int q = 0; q < GJ_MFQ; ++q) {
            dr[q] = t1[q] - t2[q];
            di[q] = t3[q] - t1[q] - t2[q];
          }
          // the empty asm statements keep the products ahead of the guarded stores
#pragma unroll
          for (int q = 0; q < GJ_MFQ; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              double x = dr[q][r], y = di[q][r];
              asm volatile("" : "+v"(x), "+v"(y));
              dr[q][r] = x;
              di[q][r] = y;
            }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const long long o = colo(jt + lq + 4 * r);
            const bool cok = jt + lq + 4 * r < cw;
#pragma unroll
            for (int q = 0; q < GJ_MFQ; ++q)
              if (st[q] && cok) a[rowc[q] + o] = cmk(dr[q][r], di[q][r]);
          }
        }
      } else if (own) {
        // column c of the window, clamped into it: a group that runs past the
        // window loads, computes and stores its last column again (same value),
        // so loads and stores are unconditional; block rows start from zero
        auto col = [&](int c) -> cplx* {
          const int jj = c0 + min(c, cw - 1);
          return a + t + (long long)((jj < k0) ? jj : jj + w) * lda;
        };
        const cplx zero = cmk(0.0, 0.0);
        cplx nx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) nx[u] = *col(u);
        // one basic block per group, so only the LDS operands of a few panel
        // columns are live at a time
#pragma unroll 1
        for (int c = 0; c <= cw - 1; c += U) {
          cplx acc[U];
#pragma unroll
          for (int u = 0; u < U; ++u) acc[u] = inblk ? zero : nx[u];
#pragma unroll
          for (int u = 0; u < U; ++u) nx[u] = *col(c + U + u);
#pragma unroll
          for (int m = 0; m < H; ++m) {
            const cplx am = pA[m], bm = pB[m];
            const cplx* ra = &Pw[m][c];
            const cplx* rb = &Pw[H + m][c];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              cfma(acc[u], am, ra[u]);
              cfma(acc[u], bm, rb[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) *col(c + u) = acc[u];
        }
      }
      __syncthreads();
    }
  }
  if (own)
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        cplx x = a[t + (long long)k * lda];
        a[t + (long long)k * lda] = a[t + (long long)p * lda];
        a[t + (long long)p * lda] = x;
      }
    }
  for (int k = t; k < n; k += TH) perm_all[(long long)blockIdx.x * n + k] = perm[k];
  if (t == 0) {
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

// ---------------------------------------------------------------------------
// Pfaffian on the matrix cores (n <= NMAX <= TH), one workgroup per matrix.
//
// Parlett-Reid with the pivoting of ZSKTF2 (UPLO = 'U', MODE = 'P': pivot =
// first maximum of |Re|+|Im| over rows 0..k-1 of column k, interchanged with
// row/column k-1), restated for a matrix that never fits a CU:
//   * only the upper triangle of the leading block is stored and updated,
//     (4/3) n^3 flops in all;
//   * pair steps run in panels of PF_PS.  Within a panel nothing is written
//     to the matrix: thread t stands for stored row t and carries its current
//     row label (a symmetric interchange kp <-> k-1 swaps two labels) and its
//     pending rank-2 vectors X_t[s] = c_t / p, Y_t[s] (column k / pivot and
//     column k-1 of step s), in registers and in LDS.  A column a step needs is
//     the stored column of the row that carries the label, plus the pending
//     terms X_t Y_k - Y_t X_k;
//   * after the panel the leading block [0, m) receives the whole rank-2*PF_PS
//     update in 16 x 16 tiles of v_mfma_f64_16x16x4f64 (complex as three real
//     products), read through the label permutation from one buffer and
//     written in label order to the other (ping-pong between the input, which
//     is destroyed, and the work matrix), so the interchanges cost no pass of
//     their own.
// Pf = prod of the pivots W(k-1,k) x (-1)^(number of interchanges), kept as
// log|pf| and a phase.
// ---------------------------------------------------------------------------
constexpr int PFM_PS = 8;   // pair steps per panel: rank-16 tile updates

typedef double __attribute__((ext_vector_type(4))) pf_d4;

template <int TH, int NMAX>
__global__ void __launch_bounds__(TH)
pfaffian_mfma_kernel(int n, cplx* __restrict__ A, int lda, long long sA, cplx* __restrict__ work,
                     double* __restrict__ logabs, cplx* __restrict__ phase, int* __restrict__ zero) {
  constexpr int NW = TH / 64;
  constexpr int LDP = NMAX + 16;        // plane stride (spreads the four k rows over the banks)
  __shared__ double Xr[PFM_PS * LDP], Xi[PFM_PS * LDP], Yr[PFM_PS * LDP], Yi[PFM_PS * LDP];
  __shared__ int rowslot[NMAX];         // stored row carrying each label
  __shared__ double wabs[NW], wcr[NW], wci[NW];
  __shared__ int wlab[NW], wslot[NW];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const bool own = t < n;
  cplx* bufA = A + (long long)blockIdx.x * sA;
  cplx* bufW = work + (long long)blockIdx.x * n * n;
  cplx* R = bufA;          // stored leading block (read)
  int ldr = lda;
  cplx* Wt = bufW;         // updated leading block (written)
  int ldw = n;

  double la = 0.0;
  cplx ph = cmk(1.0, 0.0);
  int z = 0;
  int k_top = n - 1;
  cplx X[PFM_PS], Y[PFM_PS];
  while (k_top >= 1) {
    int lab = t;
    if (own) rowslot[t] = t;
#pragma unroll
    for (int s2 = 0; s2 < PFM_PS; ++s2) { X[s2] = cmk(0.0, 0.0); Y[s2] = cmk(0.0, 0.0); }
    __syncthreads();
    int ns = 0, m = k_top + 1;
    for (int s = 0; s < PFM_PS; ++s) {
      const int k = k_top - 2 * s;
      if (k < 1) break;
      const int kk = k - 1;
      // ---- column k (rows labelled < k): stored column of row rowslot[k] + pending
      const int sk = rowslot[k];
      const bool act = own && lab < k;
      cplx c = cmk(0.0, 0.0);
      if (act) {
        if (t < sk) {
          c = R[t + (long long)sk * ldr];
        } else {
          const cplx v = R[sk + (long long)t * ldr];
          c = cmk(-v.x, -v.y);
        }
#pragma unroll
        for (int s2 = 0; s2 < PFM_PS; ++s2)
          if (s2 < s) {
            cfma(c, X[s2], cmk(Yr[s2 * LDP + sk], Yi[s2 * LDP + sk]));
            cfms(c, Y[s2], cmk(Xr[s2 * LDP + sk], Xi[s2 * LDP + sk]));
          }
      }
      // ---- pivot: first maximum of |Re|+|Im| in label order (IZAMAX)
      double v = act ? cabs1(c) : -1.0;
      int vl = act ? lab : n;
      wave_argmax(v, vl);
      if ((vl == n && lane == 0) || (act && lab == vl)) {
        wabs[wid] = v;
        wlab[wid] = vl;
        wslot[wid] = t;
        wcr[wid] = c.x;
        wci[wid] = c.y;
      }
      __syncthreads();
      double mx = wabs[0];
      int kp = wlab[0], w = wslot[0];
      double pr = wcr[0], pi = wci[0];
#pragma unroll
      for (int q = 1; q < NW; ++q)
        if (wabs[q] > mx || (wabs[q] == mx && wlab[q] < kp)) {
          mx = wabs[q]; kp = wlab[q]; w = wslot[q]; pr = wcr[q]; pi = wci[q];
        }
      // zero (or non-finite) pivot column: ZSKTF2 INFO > 0, ZSKPFA returns 0
      if (!(mx > 0.0) || kp >= k) { z = 1; break; }
      // ---- symmetric interchange kp <-> kk: two labels change hands
      if (kp != kk) {
        if (lab == kk) { lab = kp; rowslot[kp] = t; }
        else if (t == w) { lab = kk; rowslot[kk] = t; }
        ph = cmk(-ph.x, -ph.y);
      }
      const cplx p = cmk(pr, pi);
      const double ap = sqrt(p.x * p.x + p.y * p.y);
      la += log(ap);
      ph = cmul(ph, cmk(p.x / ap, p.y / ap));
      const cplx pinv = cinv(p);
      // ---- column kk after the interchange = column of stored row w
      cplx x = cmk(0.0, 0.0), y = cmk(0.0, 0.0);
      if (own && lab < kk) {
        cplx c2;
        if (t < w) {
          c2 = R[t + (long long)w * ldr];
        } else {
          const cplx u = R[w + (long long)t * ldr];
          c2 = cmk(-u.x, -u.y);
        }
#pragma unroll
        for (int s2 = 0; s2 < PFM_PS; ++s2)
          if (s2 < s) {
            cfma(c2, X[s2], cmk(Yr[s2 * LDP + w], Yi[s2 * LDP + w]));
            cfms(c2, Y[s2], cmk(Xr[s2 * LDP + w], Xi[s2 * LDP + w]));
          }
        x = cmul(c, pinv);
        y = c2;
      }
#pragma unroll
      for (int s2 = 0; s2 < PFM_PS; ++s2)
        if (s2 == s) { X[s2] = x; Y[s2] = y; }
      if (own) {
        Xr[s * LDP + t] = x.x;
        Xi[s * LDP + t] = x.y;
        Yr[s * LDP + t] = y.x;
        Yi[s * LDP + t] = y.y;
      }
      ns = s + 1;
      m = kk;
      __syncthreads();
    }
    if (z) break;
    // ---- rank-2*ns update of the upper triangle of the leading block [0, m):
    //   W(i,j) = stored(slot i, slot j) + sum_s X_i[s] Y_j[s] - Y_i[s] X_j[s]
    // as D = B^T A^T per 16 x 16 tile, A(i,q) = [X_i | Y_i], B(q,j) = [Y_j ; -X_j],
    // so that lane l holds rows I0 + l%16 of columns J0 + l/16 + 4r (coalesced).
    if (m >= 2) {
      const int nt = (m + 15) >> 4;
      const int ntiles = nt * (nt + 1) / 2;
      const int li = lane & 15, lq = lane >> 4;
      for (int tau = wid; tau < ntiles; tau += NW) {
        int bj = (int)((sqrt(8.0 * tau + 1.0) - 1.0) * 0.5);
        while ((bj + 1) * (bj + 2) / 2 <= tau) ++bj;
        while (bj * (bj + 1) / 2 > tau) --bj;
        const int bi = tau - bj * (bj + 1) / 2;
        const int I0 = bi * 16, J0 = bj * 16;
        const int i = I0 + li, ja = J0 + li;
        const bool iok = i < m, jok = ja < m;
        const int ri = iok ? rowslot[i] : 0;
        const int rja = jok ? rowslot[ja] : 0;
        // stored values of this lane's four elements (issued before the products)
        cplx cv[4];
        bool wr[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = J0 + lq + 4 * r;
          wr[r] = iok && j < m && i < j;
          cv[r] = cmk(0.0, 0.0);
          if (wr[r]) {
            const int rj = rowslot[j];
            if (ri < rj) {
              cv[r] = R[ri + (long long)rj * ldr];
            } else {
              const cplx u = R[rj + (long long)ri * ldr];
              cv[r] = cmk(-u.x, -u.y);
            }
          }
        }
        // three real products per complex one (Gauss): t1 = sum ar br, t2 = sum ai bi,
        // t3 = sum (ar + ai)(br + bi); the product is (t1 - t2, t3 - t1 - t2)
        pf_d4 t1 = {0.0, 0.0, 0.0, 0.0}, t2 = {0.0, 0.0, 0.0, 0.0}, t3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int cstep = 0; cstep < PFM_PS / 2; ++cstep) {
          const int q = lq + 4 * cstep;          // 0 .. 2*PFM_PS-1
          const bool xq = q < PFM_PS;
          const int qs = xq ? q : q - PFM_PS;
          const bool live = qs < ns;
          // A operand of the product: B(q, ja) = Y_ja[q] or -X_ja[q - PS]
          double ar = 0.0, ai = 0.0, br = 0.0, bi2 = 0.0;
          if (live && jok) {
            if (xq) { ar = Yr[qs * LDP + rja]; ai = Yi[qs * LDP + rja]; }
            else    { ar = -Xr[qs * LDP + rja]; ai = -Xi[qs * LDP + rja]; }
          }
          // B operand: A(i, q) = X_i[q] or Y_i[q - PS]
          if (live && iok) {
            if (xq) { br = Xr[qs * LDP + ri]; bi2 = Xi[qs * LDP + ri]; }
            else    { br = Yr[qs * LDP + ri]; bi2 = Yi[qs * LDP + ri]; }
          }
          t1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, t1, 0, 0, 0);
          t2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi2, t2, 0, 0, 0);
          t3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi2, t3, 0, 0, 0);
        }
        const pf_d4 dr = t1 - t2, di = t3 - t1 - t2;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (wr[r]) {
            const int j = J0 + lq + 4 * r;
            Wt[i + (long long)j * ldw] = cmk(cv[r].x + dr[r], cv[r].y + di[r]);
          }
      }
    }
    __syncthreads();
    cplx* tp = R; R = Wt; Wt = tp;
    const int tl = ldr; ldr = ldw; ldw = tl;
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
    // LDS window shrinks with it, so small matrices share a CU); above 256 rows
    // the 32-column panels of gj_wide_kernel.  TG_GJ=1: 16-column panels at
    // every order; TG_GJ=2: 16-column panels with two threads per row (the
    // second mirrors the panel reduction and takes half of the column updates),
    // n <= 384
    static const int rp = [] {
      const char* e = std::getenv("TG_GJ");
      return e ? std::atoi(e) : 0;
    }();
#define TG_GJ_LAUNCH(RP, TH, WIN)                                                              \
  hipLaunchKernelGGL((gj_blocked_kernel<16, RP, TH, WIN>), dim3(batch), dim3(TH * RP), 0, s, n, A, \
                     lda, sA, det, info, perm)
    if (rp == 2 && n <= 384)
      TG_GJ_LAUNCH(2, 384, GJ_WIN);
    else if (rp == 4 && n > 256 && n <= 384)
      hipLaunchKernelGGL((gj_wide_kernel<384, 256, 4, false>), dim3(batch), dim3(384), 0, s, n, A,
                         lda, sA, det, info, perm);
    else if (rp != 1 && n > 256 && n <= 384)
      hipLaunchKernelGGL((gj_wide_kernel<512, 256, 16, true>), dim3(batch), dim3(512), 0, s, n, A,
                         lda, sA, det, info, perm);
    else if (rp != 1 && n > 384)
      hipLaunchKernelGGL((gj_wide_kernel<512, 256, 4, false>), dim3(batch), dim3(512), 0, s, n, A,
                         lda, sA, det, info, perm);
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

int batched_pfaffian(hipStream_t s, int n, cplx* A, int lda, long long sA, int batch,
                     cplx* work, double* logabs, cplx* phase, int* zero) {
  if (batch <= 0) return TG_OK;
  if (n % 2 != 0 || n == 0) {
    tg_set_error("batched_pfaffian: n must be even and positive here");
    return TG_E_ARG;
  }
  static const bool old_pf = [] {
    const char* e = std::getenv("TG_PF");
    return e && std::atoi(e) == 1;
  }();
  if (!old_pf && n <= 512) {
    // matrix-core Pfaffian, one thread per row (at least four waves)
#define TG_PF2_LAUNCH(TH, NMAX)                                                                \
  hipLaunchKernelGGL((pfaffian_mfma_kernel<TH, NMAX>), dim3(batch), dim3(TH), 0, s, n, A, lda, sA, \
                     work, logabs, phase, zero)
    if (n <= 64)
      TG_PF2_LAUNCH(256, 64);
    else if (n <= 128)
      TG_PF2_LAUNCH(256, 128);
    else if (n <= 256)
      TG_PF2_LAUNCH(256, 256);
    else if (n <= 384)
      TG_PF2_LAUNCH(384, 384);
    else
      TG_PF2_LAUNCH(512, 512);
#undef TG_PF2_LAUNCH
    TG_HIP(hipGetLastError());
    return TG_OK;
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
