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
        v.x -= f.x * r.x - f.y * r.y;
        v.y -= f.x * r.y + f.y * r.x;
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
      v.x += (xi.x * yj.x - xi.y * yj.y) - (yi.x * xj.x - yi.y * xj.y);
      v.y += (xi.x * yj.y + xi.y * yj.x) - (yi.x * xj.y + yi.y * xj.x);
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

}  // namespace

int batched_inverse(hipStream_t s, int n, cplx* A, int lda, long long sA, int batch, cplx* det,
                    int* info, int* perm) {
  if (n <= 0 || batch <= 0) return TG_OK;
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
  size_t shm = 2 * (size_t)n * sizeof(cplx);
  hipLaunchKernelGGL(pfaffian_kernel, dim3(batch), dim3(NT), shm, s, n, A, lda, sA, work, logabs,
                     phase, zero);
  TG_HIP(hipGetLastError());
  return TG_OK;
}
