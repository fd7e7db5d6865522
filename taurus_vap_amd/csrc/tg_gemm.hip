// Tiled FP64 GEMM kernels (real / complex operands) for gfx950.
//
// One 256-thread workgroup (4 wave64) computes a 64x64 tile of C; each thread
// owns a 4x4 sub-tile with rows tx+16*i and columns ty+16*j so that the column-
// major stores of a wave are 16-element contiguous runs.  K advances in chunks
// of 16 through LDS.  Operands may be real (double), complex (double2) or FP32
// (float, converted to double when staged): the contraction blocks keep their
// matrix elements in FP32 while every product and sum is FP64.
#include "tg_internal.h"
#include "tg_mfma.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace {

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;

template <typename T> struct Stage { typedef T type; };
template <> struct Stage<float> { typedef double type; };

template <typename T> __device__ inline typename Stage<T>::type stage(T v) { return v; }
template <> __device__ inline double stage<float>(float v) { return (double)v; }

template <typename T> __device__ inline T zero_of();
template <> __device__ inline double zero_of<double>() { return 0.0; }
template <> __device__ inline float zero_of<float>() { return 0.0f; }
template <> __device__ inline cplx zero_of<cplx>() { return cmk(0.0, 0.0); }

template <typename T> __device__ inline T conj_if(T v, bool c) { return v; }
template <> __device__ inline cplx conj_if<cplx>(cplx v, bool c) { return c ? cconj(v) : v; }

// acc += a * b for the four type combinations
__device__ inline void mac(cplx& acc, cplx a, cplx b) { cfma(acc, a, b); }
__device__ inline void mac(cplx& acc, double a, cplx b) { acc.x = fma(a, b.x, acc.x); acc.y = fma(a, b.y, acc.y); }
__device__ inline void mac(cplx& acc, cplx a, double b) { acc.x = fma(a.x, b, acc.x); acc.y = fma(a.y, b, acc.y); }
__device__ inline void mac(double& acc, double a, double b) { acc = fma(a, b, acc); }

template <typename TC> struct Scal;
template <> struct Scal<cplx> { typedef cplx type; };
template <> struct Scal<double> { typedef double type; };

__device__ inline cplx epi(cplx alpha, cplx acc, cplx beta, const cplx* cptr, bool use_beta) {
  cplx r = cmul(alpha, acc);
  if (use_beta) r = cadd(r, cmul(beta, *cptr));
  return r;
}
__device__ inline double epi(double alpha, double acc, double beta, const double* cptr, bool use_beta) {
  double r = alpha * acc;
  if (use_beta) r = fma(beta, *cptr, r);
  return r;
}

// One 64x64 tile: C(m0:m0+64, n0:n0+64) of a column-major M x N result.
template <typename TA, typename TB, typename TC, int OPA, int OPB>
__device__ inline void gemm_tile(int M, int N, int K, typename Scal<TC>::type alpha,
                                 const TA* __restrict__ A, int lda,
                                 const TB* __restrict__ B, int ldb,
                                 typename Scal<TC>::type beta, bool use_beta,
                                 TC* __restrict__ C, int ldc, int m0, int n0,
                                 typename Stage<TA>::type (*As)[BM + 1],
                                 typename Stage<TB>::type (*Bs)[BN + 1]) {
  typedef typename Stage<TA>::type SA;
  typedef typename Stage<TB>::type SB;
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  TC acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zero_of<TC>();

  for (int k0 = 0; k0 < K; k0 += BK) {
    // ---- stage A(m0:+64, k0:+16) into As[k][i]
    if (OPA == OP_N) {
      const int i = t & 63;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = (t >> 6) + 4 * r;
        const int gi = m0 + i, gk = k0 + k;
        SA v = stage(zero_of<TA>());
        if (gi < M && gk < K) v = stage(A[gi + (long long)gk * lda]);
        As[k][i] = v;
      }
    } else {
      const int k = t & 15;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = (t >> 4) + 16 * r;
        const int gi = m0 + i, gk = k0 + k;
        SA v = stage(zero_of<TA>());
        if (gi < M && gk < K) v = conj_if(stage(A[gk + (long long)gi * lda]), OPA == OP_C);
        As[k][i] = v;
      }
    }
    // ---- stage B(k0:+16, n0:+64) into Bs[k][j]
    if (OPB == OP_N) {
      const int k = t & 15;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = (t >> 4) + 16 * r;
        const int gk = k0 + k, gj = n0 + j;
        SB v = stage(zero_of<TB>());
        if (gk < K && gj < N) v = stage(B[gk + (long long)gj * ldb]);
        Bs[k][j] = v;
      }
    } else {
      const int j = t & 63;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = (t >> 6) + 4 * r;
        const int gk = k0 + k, gj = n0 + j;
        SB v = stage(zero_of<TB>());
        if (gk < K && gj < N) v = conj_if(stage(B[gj + (long long)gk * ldb]), OPB == OP_C);
        Bs[k][j] = v;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < BK; ++k) {
      SA a[4];
      SB b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) mac(acc[i][j], a[i], b[j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gj = n0 + ty + 16 * j;
    if (gj >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int gi = m0 + tx + 16 * i;
      if (gi >= M) continue;
      TC* cp = C + gi + (long long)gj * ldc;
      *cp = epi(alpha, acc[i][j], beta, cp, use_beta);
    }
  }
}

template <typename TA, typename TB, typename TC, int OPA, int OPB>
__global__ void __launch_bounds__(NT)
gemm_batched_kernel(int M, int N, int K, typename Scal<TC>::type alpha,
                    const TA* __restrict__ A, int lda, long long sA,
                    const TB* __restrict__ B, int ldb, long long sB,
                    typename Scal<TC>::type beta, int use_beta,
                    TC* __restrict__ C, int ldc, long long sC, int divA, int divB) {
  __shared__ typename Stage<TA>::type As[BK][BM + 1];
  __shared__ typename Stage<TB>::type Bs[BK][BN + 1];
  const long long b = blockIdx.z;
  gemm_tile<TA, TB, TC, OPA, OPB>(M, N, K, alpha, A + (b / divA) * sA, lda, B + (b / divB) * sB,
                                  ldb, beta,
                                  use_beta != 0, C + b * sC, ldc, blockIdx.x * BM,
                                  blockIdx.y * BN, As, Bs);
}

template <typename TW>
__global__ void __launch_bounds__(NT)
gemm_grouped_kernel(int M, const cplx* __restrict__ X, cplx* __restrict__ C,
                    const TW* __restrict__ W, const GroupDesc* __restrict__ groups,
                    const int* __restrict__ tile_group, const int* __restrict__ tile_n0) {
  __shared__ cplx As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tile = blockIdx.x;
  const GroupDesc g = groups[tile_group[tile]];
  const int n0 = tile_n0[tile];
  const cplx* Xg = X + (long long)g.off * M;
  cplx* Cg = C + (long long)g.off * M;
  gemm_tile<cplx, TW, cplx, OP_N, OP_N>(M, g.size, g.size, cmk(1.0, 0.0), Xg, M, W + g.woff,
                                        g.size, cmk(0.0, 0.0), false, Cg, M, blockIdx.y * BM, n0,
                                        As, Bs);
}

template <typename TA, typename TB, typename TC, int OPA, int OPB>
int launch(hipStream_t s, int M, int N, int K, typename Scal<TC>::type alpha, const TA* A,
           int lda, long long sA, const TB* B, int ldb, long long sB,
           typename Scal<TC>::type beta, int use_beta, TC* C, int ldc, long long sC, int batch,
           int divA, int divB) {
  if (M <= 0 || N <= 0 || batch <= 0) return TG_OK;
  if (batch > 65535) {
    tg_set_error("gemm: batch too large");
    return TG_E_ARG;
  }
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  hipLaunchKernelGGL((gemm_batched_kernel<TA, TB, TC, OPA, OPB>), grid, dim3(NT), 0, s, M, N,
                     K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, divA, divB);
  TG_HIP(hipGetLastError());
  return TG_OK;
}

template <typename TA, typename TB, typename TC>
int dispatch(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K,
             typename Scal<TC>::type alpha, const TA* A, int lda, long long sA, const TB* B,
             int ldb, long long sB, typename Scal<TC>::type beta, int use_beta, TC* C, int ldc,
             long long sC, int batch, int divA, int divB) {
#define TG_CASE(OA, OB)                                                                   \
  if (opa == OA && opb == OB)                                                             \
    return launch<TA, TB, TC, OA, OB>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,    \
                                      use_beta, C, ldc, sC, batch, divA, divB);
  TG_CASE(OP_N, OP_N) TG_CASE(OP_N, OP_T) TG_CASE(OP_N, OP_C)
  TG_CASE(OP_T, OP_N) TG_CASE(OP_T, OP_T) TG_CASE(OP_T, OP_C)
  TG_CASE(OP_C, OP_N) TG_CASE(OP_C, OP_T) TG_CASE(OP_C, OP_C)
#undef TG_CASE
  tg_set_error("gemm: bad op");
  return TG_E_ARG;
}

inline bool is_zero(cplx z) { return z.x == 0.0 && z.y == 0.0; }

// Kernel family of the complex GEMMs: matrix-core (MFMA) tiles unless
// TG_GEMM=vector; TG_GEMM_VARIANT / TG_GROUPED_VARIANT force one tile shape
// (and switch the autotuner off for that family).
struct GemmSel {
  int mfma = 1, variant = -1, grouped = -1;
  GemmSel() {
    const char* e = std::getenv("TG_GEMM");
    if (e && std::strcmp(e, "vector") == 0) mfma = 0;
    const char* v = std::getenv("TG_GEMM_VARIANT");
    if (v) variant = std::atoi(v);
    const char* g = std::getenv("TG_GROUPED_VARIANT");
    if (g) grouped = std::atoi(g);
  }
};
const GemmSel& sel() {
  static GemmSel s;
  return s;
}

// Autotuned tile shape per GEMM signature (kind, ops, sizes, batch).  Filled by
// gemm_tune_* (timed runs of every variant); signatures never tuned use tile 0.
struct GemmKey {
  int kind, opa, opb, M, N, K, batch;
  bool operator<(const GemmKey& o) const {
    const int a[7] = {kind, opa, opb, M, N, K, batch};
    const int b[7] = {o.kind, o.opa, o.opb, o.M, o.N, o.K, o.batch};
    for (int i = 0; i < 7; ++i)
      if (a[i] != b[i]) return a[i] < b[i];
    return false;
  }
};
enum { K_ZZ = 0, K_DZ = 1, K_ZD = 2, K_GROUPED = 3 };
std::map<GemmKey, int>& choices() {
  static std::map<GemmKey, int> m;
  return m;
}
int variant_for(const GemmKey& k) {
  const int forced = (k.kind == K_GROUPED) ? sel().grouped : sel().variant;
  if (forced >= 0) return forced;
  auto it = choices().find(k);
  return it == choices().end() ? 0 : it->second;
}

// Times `reps` runs of each variant on stream s and records the fastest.
template <typename F>
int tune(hipStream_t s, const GemmKey& key, int nvar, F run) {
  if (!sel().mfma) return TG_OK;
  const int forced = (key.kind == K_GROUPED) ? sel().grouped : sel().variant;
  if (forced >= 0 || choices().count(key)) return TG_OK;
  hipEvent_t e0, e1;
  TG_HIP(hipEventCreate(&e0));
  TG_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  int rc = TG_OK;
  for (int v = 0; v < nvar && rc == TG_OK; ++v) {
    rc = run(v);                                  // warm-up (first launch of the kernel)
    if (rc != TG_OK) break;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < 2 && rc == TG_OK; ++r) rc = run(v);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) rc = TG_E_HIP;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (rc == TG_OK && std::getenv("TG_GEMM_VERBOSE"))
      fprintf(stderr, "[taurus_gpu]   kind %d ops %d%d M %d N %d K %d batch %d tile %d: %.3f ms\n",
              key.kind, key.opa, key.opb, key.M, key.N, key.K, key.batch, v, ms / 2);
    if (rc == TG_OK && ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TG_OK) return rc;
  choices()[key] = best;
  if (std::getenv("TG_GEMM_VERBOSE"))
    fprintf(stderr, "[taurus_gpu] gemm kind %d ops %d%d M %d N %d K %d batch %d -> tile %d (%.3f ms)\n",
            key.kind, key.opa, key.opb, key.M, key.N, key.K, key.batch, best, best_ms / 2);
  return TG_OK;
}

}  // namespace

int gemm_tuned_variant(int kind, TgOp opa, TgOp opb, int M, int N, int K, int batch) {
  return variant_for(GemmKey{kind, (int)opa, (int)opb, M, N, K, batch});
}

int gemm_tune_zz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const cplx* A, int lda,
                 long long sA, const cplx* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch) {
  const GemmKey key{K_ZZ, (int)opa, (int)opb, M, N, K, batch};
  return tune(s, key, TG_MFMA_VARIANTS, [&](int v) {
    return mfma_gemm_zz(v, s, opa, opb, M, N, K, cmk(1, 0), A, lda, sA, B, ldb, sB, cmk(0, 0), 0,
                        C, ldc, sC, batch);
  });
}

int gemm_tune_dz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const double* A,
                 int lda, long long sA, const cplx* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch) {
  if (opa == OP_C) opa = OP_T;
  const GemmKey key{K_DZ, (int)opa, (int)opb, M, N, K, batch};
  return tune(s, key, TG_MFMA_VARIANTS, [&](int v) {
    return mfma_gemm_dz(v, s, opa, opb, M, N, K, cmk(1, 0), A, lda, sA, B, ldb, sB, cmk(0, 0), 0,
                        C, ldc, sC, batch);
  });
}

int gemm_tune_zd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, const cplx* A, int lda,
                 long long sA, const double* B, int ldb, long long sB, cplx* C, int ldc,
                 long long sC, int batch) {
  if (opb == OP_C) opb = OP_T;
  const GemmKey key{K_ZD, (int)opa, (int)opb, M, N, K, batch};
  return tune(s, key, TG_MFMA_VARIANTS, [&](int v) {
    return mfma_gemm_zd(v, s, opa, opb, M, N, K, cmk(1, 0), A, lda, sA, B, ldb, sB, cmk(0, 0), 0,
                        C, ldc, sC, batch);
  });
}

bool gemm_grouped_tuned(int M, int ntiles_total, int tag) {
  if (!sel().mfma || sel().grouped >= 0) return true;
  return choices().count(GemmKey{K_GROUPED, tag, 0, M, ntiles_total, 0, 1}) != 0;
}

int gemm_tune_grouped(hipStream_t s, int M, const cplx* X, cplx* C, const void* W, int w_is_fp32,
                      const GroupDesc* groups_dev, const int* tile_group_dev,
                      const int* tile_n0_dev, int ntiles_total, int tag) {
  const GemmKey key{K_GROUPED, tag, 0, M, ntiles_total, 0, 1};
  return tune(s, key, TG_MFMA_GROUPED_VARIANTS, [&](int v) {
    return mfma_gemm_grouped(v, s, M, X, C, W, w_is_fp32, groups_dev, tile_group_dev, tile_n0_dev,
                             ntiles_total);
  });
}

int gemm_zz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const cplx* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA, int divB,
            int variant) {
  if (M <= 0 || N <= 0 || batch <= 0) return TG_OK;
  if (sel().mfma)
    return mfma_gemm_zz(variant >= 0 ? variant
                                     : variant_for(GemmKey{K_ZZ, (int)opa, (int)opb, M, N, K, batch}),
                        s, opa,
                        opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                        !is_zero(beta), C, ldc, sC, batch, divA, divB);
  return dispatch<cplx, cplx, cplx>(s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                                    !is_zero(beta), C, ldc, sC, batch, divA, divB);
}

int gemm_dz(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const double* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA, int divB) {
  if (M <= 0 || N <= 0 || batch <= 0) return TG_OK;
  if (opa == OP_C) opa = OP_T;
  if (sel().mfma)
    return mfma_gemm_dz(variant_for(GemmKey{K_DZ, (int)opa, (int)opb, M, N, K, batch}), s, opa,
                        opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                        !is_zero(beta), C, ldc, sC, batch, divA, divB);
  return dispatch<double, cplx, cplx>(s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                                      !is_zero(beta), C, ldc, sC, batch, divA, divB);
}

int gemm_zd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
            const cplx* A, int lda, long long sA, const double* B, int ldb, long long sB,
            cplx beta, cplx* C, int ldc, long long sC, int batch, int divA, int divB) {
  if (M <= 0 || N <= 0 || batch <= 0) return TG_OK;
  if (opb == OP_C) opb = OP_T;
  if (sel().mfma)
    return mfma_gemm_zd(variant_for(GemmKey{K_ZD, (int)opa, (int)opb, M, N, K, batch}), s, opa,
                        opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                        !is_zero(beta), C, ldc, sC, batch, divA, divB);
  return dispatch<cplx, double, cplx>(s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                                      !is_zero(beta), C, ldc, sC, batch, divA, divB);
}

int gemm_dd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, double alpha,
            const double* A, int lda, long long sA, const double* B, int ldb, long long sB,
            double beta, double* C, int ldc, long long sC, int batch, int divA, int divB) {
  if (opa == OP_C) opa = OP_T;
  if (opb == OP_C) opb = OP_T;
  if (sel().mfma)
    return mfma_gemm_dd(s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, beta != 0.0,
                        C, ldc, sC, batch, divA, divB);
  return dispatch<double, double, double>(s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB,
                                          beta, beta != 0.0, C, ldc, sC, batch, divA, divB);
}

int gemm_grouped(hipStream_t s, int M, const cplx* X, cplx* C, const void* W, int w_is_fp32,
                 const GroupDesc* groups_dev, const int* tile_group_dev, const int* tile_n0_dev,
                 int tag, int ntiles_total) {
  if (M <= 0 || ntiles_total <= 0) return TG_OK;
  if (sel().mfma)
    return mfma_gemm_grouped(variant_for(GemmKey{K_GROUPED, tag, 0, M, ntiles_total, 0, 1}), s, M,
                             X, C, W, w_is_fp32, groups_dev, tile_group_dev, tile_n0_dev,
                             ntiles_total);
  dim3 grid(ntiles_total, (M + BM - 1) / BM, 1);
  if (w_is_fp32)
    hipLaunchKernelGGL(gemm_grouped_kernel<float>, grid, dim3(NT), 0, s, M, X, C,
                       (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
  else
    hipLaunchKernelGGL(gemm_grouped_kernel<double>, grid, dim3(NT), 0, s, M, X, C,
                       (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  TG_HIP(hipGetLastError());
  return TG_OK;
}
