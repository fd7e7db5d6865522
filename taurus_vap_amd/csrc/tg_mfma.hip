// FP64 matrix-core GEMMs (v_mfma_f64_16x16x4f64) for gfx950.
//
// C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b], column-major, with complex
// operands split into real and imaginary planes in LDS:
//   complex x complex : Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br  (4 MFMA)
//   real x complex    : Cr += A Br,          Ci += A Bi          (2 MFMA)
//   complex x real    : Cr += Ar B,          Ci += Ai B          (2 MFMA)
// A workgroup of WM x WN waves owns a BM x BN tile of C; each wave owns a
// (BM/WM) x (BN/WN) sub-tile of 16 x 16 MFMA blocks.  K advances in chunks of 16
// staged through LDS; the next chunk is loaded into registers while the
// current one is multiplied.  Each operand is kept in LDS with its contiguous
// global dimension fastest, so the staging stores are unit-stride.
//
// MFMA 16x16x4 f64 operand layout (lane l = 0..63):
//   A(16x4): lane l holds A[l % 16][l / 16];  B(4x16): lane l holds B[l / 16][l % 16];
//   D(16x16): lane l, register r holds D[l / 16 + 4 * r][l % 16].
//
// Batched launches are mapped XCD-aware: workgroup ids are dealt round-robin
// to the 8 XCDs, so all tiles of one batch element are given ids congruent
// modulo 8 and share that XCD's L2.
#include "tg_internal.h"
#include "tg_mfma.h"

namespace {

typedef double __attribute__((ext_vector_type(4))) d4;

__device__ inline d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T> struct IsC { static constexpr bool v = false; };
template <> struct IsC<cplx> { static constexpr bool v = true; };

__device__ inline void split(cplx v, double& r, double& i) { r = v.x; i = v.y; }
__device__ inline void split(double v, double& r, double& i) { r = v; i = 0.0; }
__device__ inline void split(float v, double& r, double& i) { r = (double)v; i = 0.0; }

template <typename T> __device__ inline T zero_v();
template <> __device__ inline double zero_v<double>() { return 0.0; }
template <> __device__ inline float zero_v<float>() { return 0.0f; }
template <> __device__ inline cplx zero_v<cplx>() { return cmk(0.0, 0.0); }

constexpr int XCDS = 8;

// KM = true stages both operands k-major in LDS ([k][i], [k][j], rows of BM / BN
// plus a pad) whatever their layout in memory; false keeps each operand's
// memory order (the row of a k-contiguous operand then holds BK + 1).
template <typename TA, typename TB, int OPA, int OPB, int BM, int BN, int WM, int WN,
          bool KM = false>
struct Tile {
  static constexpr int BK = 16;
  static constexpr int NW = WM * WN, NTH = 64 * NW;
  static constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  static constexpr bool CA = IsC<TA>::v, CB = IsC<TB>::v;
  // A(i,k) in memory: OP_N -> contiguous in i (AG), else contiguous in k
  static constexpr bool AG = (OPA == OP_N);
  // B(k,j) in memory: OP_N -> contiguous in k (BG), else contiguous in j
  static constexpr bool BG = (OPB == OP_N);
  // LDS: A as [k][i] (AK) or [i][k]; B as [j][k] (BJ) or [k][j]
  static constexpr bool AK = AG || KM;
  static constexpr bool BJ = BG && !KM;
  static constexpr int PADR = 4;                   // pad of rows of length BM / BN
  static constexpr int LDA_ = AK ? BM + PADR : BK + 1;
  static constexpr int LDB_ = BJ ? BK + 1 : BN + PADR;
  static constexpr int SA = AK ? BK * LDA_ : BM * LDA_;
  static constexpr int SB = BJ ? BN * LDB_ : BK * LDB_;
  static constexpr int NA = BM * BK / NTH, NB = BN * BK / NTH;
  static_assert(BM % (16 * WM) == 0 && BN % (16 * WN) == 0, "tile");
  static_assert((BM * BK) % NTH == 0 && (BN * BK) % NTH == 0, "staging");

  __device__ static inline int ia(int i, int k) { return AK ? k * LDA_ + i : i * LDA_ + k; }
  __device__ static inline int ib(int k, int j) { return BJ ? j * LDB_ + k : k * LDB_ + j; }

  struct Smem {
    double ar[SA];
    double ai[CA ? SA : 1];
    double br[SB];
    double bi[CB ? SB : 1];
  };

  // element (i, k) of op(A) and (k, j) of op(B) handled by staging slot e
  // (consecutive slots follow the memory order: coalesced loads)
  __device__ static inline void slot_a(int e, int& i, int& k) {
    if (AG) { i = e % BM; k = e / BM; } else { k = e % BK; i = e / BK; }
  }
  __device__ static inline void slot_b(int e, int& k, int& j) {
    if (BG) { k = e % BK; j = e / BK; } else { j = e % BN; k = e / BN; }
  }

  // always inlined: several kernels share one tile body
  __device__ __forceinline__ static void run(int M, int N, int K, cplx alpha, const TA* __restrict__ A, int lda,
                             const TB* __restrict__ B, int ldb, cplx beta, bool use_beta,
                             cplx* __restrict__ C, int ldc, int m0, int n0, Smem& sm) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm0 = (w % WM) * (BM / WM), wn0 = (w / WM) * (BN / WN);
    const int lr = lane & 15, lk = lane >> 4;

    // complex x complex in three real products (Gauss): cr = sum Ar Br, ci = sum
    // Ai Bi, c3 = sum (Ar + Ai)(Br + Bi); C = (cr - ci, c3 - cr - ci) at the end.
    // A real operand takes two products: cr, ci are the real and imaginary parts.
    constexpr bool CC = CA && CB;
    d4 cr[TM][TN], ci[TM][TN], c3[CC ? TM : 1][CC ? TN : 1];
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        cr[x][y] = d4{0.0, 0.0, 0.0, 0.0};
        ci[x][y] = d4{0.0, 0.0, 0.0, 0.0};
        if constexpr (CC) c3[x][y] = d4{0.0, 0.0, 0.0, 0.0};
      }

    TA ra[NA];
    TB rb[NB];
    auto load = [&](int k0) {
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        int i, k;
        slot_a(t + r * NTH, i, k);
        const int gi = m0 + i, gk = k0 + k;
        TA v = zero_v<TA>();
        if (gi < M && gk < K) v = AG ? A[gi + (long long)gk * lda] : A[gk + (long long)gi * lda];
        ra[r] = v;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        int k, j;
        slot_b(t + r * NTH, k, j);
        const int gk = k0 + k, gj = n0 + j;
        TB v = zero_v<TB>();
        if (gk < K && gj < N) v = BG ? B[gk + (long long)gj * ldb] : B[gj + (long long)gk * ldb];
        rb[r] = v;
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        int i, k;
        slot_a(t + r * NTH, i, k);
        double vr, vi;
        split(ra[r], vr, vi);
        sm.ar[ia(i, k)] = vr;
        if (CA) sm.ai[ia(i, k)] = (OPA == OP_C) ? -vi : vi;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        int k, j;
        slot_b(t + r * NTH, k, j);
        double vr, vi;
        split(rb[r], vr, vi);
        sm.br[ib(k, j)] = vr;
        if (CB) sm.bi[ib(k, j)] = (OPB == OP_C) ? -vi : vi;
      }
    };

    load(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
      store();
      __syncthreads();
      if (k0 + BK < K) load(k0 + BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        double ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
        for (int x = 0; x < TM; ++x) {
          const int idx = ia(wm0 + 16 * x + lr, kk + lk);
          ar[x] = sm.ar[idx];
          if (CA) ai[x] = sm.ai[idx];
        }
#pragma unroll
        for (int y = 0; y < TN; ++y) {
          const int idx = ib(kk + lk, wn0 + 16 * y + lr);
          br[y] = sm.br[idx];
          if (CB) bi[y] = sm.bi[idx];
        }
        double sa[CC ? TM : 1], sb[CC ? TN : 1];
        if constexpr (CC) {
#pragma unroll
          for (int x = 0; x < TM; ++x) sa[x] = ar[x] + ai[x];
#pragma unroll
          for (int y = 0; y < TN; ++y) sb[y] = br[y] + bi[y];
        }
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TN; ++y) {
            if constexpr (CC) {
              cr[x][y] = mfma(ar[x], br[y], cr[x][y]);
              ci[x][y] = mfma(ai[x], bi[y], ci[x][y]);
              c3[x][y] = mfma(sa[x], sb[y], c3[x][y]);
            } else if constexpr (CA) {
              cr[x][y] = mfma(ar[x], br[y], cr[x][y]);
              ci[x][y] = mfma(ai[x], br[y], ci[x][y]);
            } else {
              cr[x][y] = mfma(ar[x], br[y], cr[x][y]);
              ci[x][y] = mfma(ar[x], bi[y], ci[x][y]);
            }
          }
      }
      __syncthreads();
    }

    // epilogue: lane holds rows lk + 4*r of column lr of each 16x16 block
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const int gj = n0 + wn0 + 16 * y + lr;
      if (gj >= N) continue;
#pragma unroll
      for (int x = 0; x < TM; ++x) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = m0 + wm0 + 16 * x + lk + 4 * r;
          if (gi >= M) continue;
          cplx* cp = C + gi + (long long)gj * ldc;
          cplx p;
          if constexpr (CC)
            p = cmk(cr[x][y][r] - ci[x][y][r], c3[x][y][r] - cr[x][y][r] - ci[x][y][r]);
          else
            p = cmk(cr[x][y][r], ci[x][y][r]);
          cplx v = cmul(alpha, p);
          if (use_beta) v = cadd(v, cmul(beta, *cp));
          *cp = v;
        }
      }
    }
  }
};

// Batched: grid = XCDS * per workgroups.  BAL = false: batch element b goes to
// XCD b % 8 (per = ceil(batch / 8) * tiles); BAL = true: the batch * tiles
// (element, tile) pairs are cut into 8 contiguous runs of `per`, one per XCD, so
// every XCD gets the same number of tiles.  Either way the tiles of one element
// are consecutive on their XCD.  OCC > 1 bounds the registers for OCC times the
// workgroup's waves per CU.
template <typename TA, typename TB, int OPA, int OPB, int BM, int BN, int WM, int WN, bool BAL,
          int OCC, bool KM>
__global__ void __launch_bounds__(64 * WM * WN * OCC)
mfma_batched_kernel(int M, int N, int K, cplx alpha, const TA* __restrict__ A, int lda,
                    long long sA, const TB* __restrict__ B, int ldb, long long sB, cplx beta,
                    int use_beta, cplx* __restrict__ C, int ldc, long long sC, int batch,
                    int tiles_m, int tiles, long long per, int divA, int divB) {
  typedef Tile<TA, TB, OPA, OPB, BM, BN, WM, WN, KM> T;
  __shared__ typename T::Smem sm;
  const int L = blockIdx.x;
  const int xcd = L % XCDS, q = L / XCDS;
  long long b;
  int tile;
  if (BAL) {
    const long long g = (long long)xcd * per + q;
    if (g >= (long long)batch * tiles) return;
    b = g / tiles;
    tile = (int)(g - b * tiles);
  } else {
    const int slot = q / tiles;
    tile = q % tiles;
    b = xcd + (long long)XCDS * slot;
    if (b >= batch) return;
  }
  const int tm = tile % tiles_m, tn = tile / tiles_m;
  T::run(M, N, K, alpha, A + (b / divA) * sA, lda, B + (b / divB) * sB, ldb, beta, use_beta != 0,
         C + b * sC, ldc, tm * BM, tn * BN, sm);
}

// Real FP64 GEMM on the matrix cores for the real products of the constraint
// machinery and the canonical basis (n x n, batched over states):
// C[b] = alpha op(A[b]) op(B[b]) + beta C[b].  A 32 x 32 tile per workgroup of
// four waves, one 16 x 16 MFMA block per wave, K staged 16 deep in LDS.  Each
// wave computes the transposed block (D = op(B)^T op(A)^T), so that lane l
// holds rows l % 16 of four columns: the stores are unit-stride.
template <int OPA, int OPB>
__global__ void __launch_bounds__(256)
mfma_dd_kernel(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
               long long sA, const double* __restrict__ B, int ldb, long long sB, double beta,
               int use_beta, double* __restrict__ C, int ldc, long long sC, int tiles_m, int divA,
               int divB) {
  constexpr int BT = 32, BK = 16, LD = BT + 1;
  __shared__ double As[BK][LD], Bs[BK][LD];     // As[k][i] = op(A)(i,k), Bs[k][j] = op(B)(k,j)
  const int b = blockIdx.y;
  const double* Ab = A + (long long)(b / divA) * sA;
  const double* Bb = B + (long long)(b / divB) * sB;
  double* Cb = C + (long long)b * sC;
  const int m0 = (blockIdx.x % tiles_m) * BT, n0 = (blockIdx.x / tiles_m) * BT;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 16, wj = (w >> 1) * 16;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += BK) {
    // 512 elements of each operand tile, two per thread, contiguous index fastest
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int e = t + 256 * r;
      int i, k;
      if (OPA == OP_N) { i = e % BT; k = e / BT; } else { k = e % BK; i = e / BK; }
      const int gi = m0 + i, gk = k0 + k;
      double v = 0.0;
      if (gi < M && gk < K) v = (OPA == OP_N) ? Ab[gi + (long long)gk * lda] : Ab[gk + (long long)gi * lda];
      As[k][i] = v;
      int j, k2;
      if (OPB == OP_N) { k2 = e % BK; j = e / BK; } else { j = e % BT; k2 = e / BT; }
      const int gj = n0 + j, gk2 = k0 + k2;
      double u = 0.0;
      if (gj < N && gk2 < K) u = (OPB == OP_N) ? Bb[gk2 + (long long)gj * ldb] : Bb[gj + (long long)gk2 * ldb];
      Bs[k2][j] = u;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const double av = Bs[kk + lk][wj + lr];    // D rows: columns of C
      const double bv = As[kk + lk][wi + lr];    // D columns: rows of C
      acc = mfma(av, bv, acc);
    }
    __syncthreads();
  }
  // lane l, register r: C(m0 + wi + lr, n0 + wj + lk + 4 r)
  const int gi = m0 + wi + lr;
  if (gi < M) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gj = n0 + wj + lk + 4 * r;
      if (gj < N) {
        double* cp = Cb + gi + (long long)gj * ldc;
        double v = alpha * acc[r];
        if (use_beta) v = fma(beta, *cp, v);
        *cp = v;
      }
    }
  }
}

// Grouped contraction: C_g (M x s_g) = X_g (M x s_g) W_g (s_g x s_g), W real (float/double).
template <typename TW, int BM, int BN, int WM, int WN, int OCC = 1>
__global__ void __launch_bounds__(64 * WM * WN * OCC)
mfma_grouped_kernel(int M, const cplx* __restrict__ X, cplx* __restrict__ C,
                    const TW* __restrict__ W, const GroupDesc* __restrict__ groups,
                    const int* __restrict__ tile_group, const int* __restrict__ tile_n0) {
  typedef Tile<cplx, TW, OP_N, OP_N, BM, BN, WM, WN> T;
  __shared__ typename T::Smem sm;
  const int tile = blockIdx.x;
  const GroupDesc g = groups[tile_group[tile]];
  const int n0 = tile_n0[tile];
  T::run(M, g.size, g.size, cmk(1.0, 0.0), X + (long long)g.off * M, M, W + g.woff, g.size,
         cmk(0.0, 0.0), false, C + (long long)g.off * M, M, blockIdx.y * BM, n0, sm);
}

template <typename TA, typename TB, int OPA, int OPB, int BM, int BN, int WM, int WN,
          bool BAL = false, int OCC = 1, bool KM = false>
int launch_batched(hipStream_t s, int M, int N, int K, cplx alpha, const TA* A, int lda,
                   long long sA, const TB* B, int ldb, long long sB, cplx beta, int use_beta,
                   cplx* C, int ldc, long long sC, int batch, int divA, int divB) {
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int tiles = tiles_m * tiles_n;
  const long long per = BAL ? ((long long)batch * tiles + XCDS - 1) / XCDS
                            : (long long)((batch + XCDS - 1) / XCDS) * tiles;
  const long long grid = (long long)XCDS * per;
  if (grid > 0x7fffffffLL) {
    tg_set_error("mfma gemm: grid too large");
    return TG_E_ARG;
  }
  if (grid == 0) return TG_OK;
  hipLaunchKernelGGL((mfma_batched_kernel<TA, TB, OPA, OPB, BM, BN, WM, WN, BAL, OCC, KM>),
                     dim3((unsigned)grid), dim3(64 * WM * WN), 0, s, M, N, K, alpha, A, lda, sA, B,
                     ldb, sB, beta, use_beta, C, ldc, sC, batch, tiles_m, tiles, per, divA, divB);
  TG_HIP(hipGetLastError());
  return TG_OK;
}

template <typename TA, typename TB, int OPA, int OPB>
int launch_variant(int variant, hipStream_t s, int M, int N, int K, cplx alpha, const TA* A,
                   int lda, long long sA, const TB* B, int ldb, long long sB, cplx beta,
                   int use_beta, cplx* C, int ldc, long long sC, int batch, int divA, int divB) {
  switch (variant) {
    case 0:
      return launch_batched<TA, TB, OPA, OPB, 64, 64, 2, 2>(s, M, N, K, alpha, A, lda, sA, B, ldb,
                                                            sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 1:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3>(s, M, N, K, alpha, A, lda, sA, B, ldb,
                                                            sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 2:
      return launch_batched<TA, TB, OPA, OPB, 112, 112, 1, 7>(s, M, N, K, alpha, A, lda, sA, B,
                                                              ldb, sB, beta, use_beta, C, ldc, sC,
                                                              batch, divA, divB);
    case 3:
      return launch_batched<TA, TB, OPA, OPB, 96, 64, 2, 2>(s, M, N, K, alpha, A, lda, sA, B, ldb,
                                                            sB, beta, use_beta, C, ldc, sC, batch,
                                                            divA, divB);
    case 4:
      return launch_batched<TA, TB, OPA, OPB, 112, 112, 1, 7, true>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 5:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3, true>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 6:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3, false, 4>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 7:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3, true, 4>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 8:
      return launch_batched<TA, TB, OPA, OPB, 64, 32, 1, 2, true, 4>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 9:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3, true, 4, true>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
    case 10:
      return launch_batched<TA, TB, OPA, OPB, 48, 48, 1, 3, true, 1, true>(
          s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, use_beta, C, ldc, sC, batch, divA, divB);
  }
  tg_set_error("mfma gemm: bad tile variant");
  return TG_E_ARG;
}

template <typename TA, typename TB>
int dispatch(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
             const TA* A, int lda, long long sA, const TB* B, int ldb, long long sB, cplx beta,
             int use_beta, cplx* C, int ldc, long long sC, int batch, int divA, int divB) {
#define TG_CASE(OA, OB)                                                                        \
  if (opa == OA && opb == OB)                                                                  \
    return launch_variant<TA, TB, OA, OB>(variant, s, M, N, K, alpha, A, lda, sA, B, ldb, sB,  \
                                          beta, use_beta, C, ldc, sC, batch, divA, divB);
  TG_CASE(OP_N, OP_N) TG_CASE(OP_N, OP_T) TG_CASE(OP_T, OP_N) TG_CASE(OP_T, OP_T)
  if (IsC<TA>::v) { TG_CASE(OP_C, OP_N) TG_CASE(OP_C, OP_T) }
  if (IsC<TB>::v) { TG_CASE(OP_N, OP_C) TG_CASE(OP_T, OP_C) }
  if (IsC<TA>::v && IsC<TB>::v) { TG_CASE(OP_C, OP_C) }
#undef TG_CASE
  tg_set_error("mfma gemm: unsupported op combination");
  return TG_E_ARG;
}

}  // namespace

int mfma_gemm_zz(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const cplx* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch, int divA,
                 int divB) {
  return dispatch<cplx, cplx>(variant, s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                              use_beta, C, ldc, sC, batch, divA, divB);
}

int mfma_gemm_dz(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const double* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch, int divA,
                 int divB) {
  if (opa == OP_C) opa = OP_T;
  return dispatch<double, cplx>(variant, s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                                use_beta, C, ldc, sC, batch, divA, divB);
}

int mfma_gemm_zd(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const cplx* A, int lda, long long sA, const double* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch, int divA,
                 int divB) {
  if (opb == OP_C) opb = OP_T;
  return dispatch<cplx, double>(variant, s, opa, opb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,
                                use_beta, C, ldc, sC, batch, divA, divB);
}

int mfma_gemm_grouped(int variant, hipStream_t s, int M, const cplx* X, cplx* C, const void* W,
                      int w_is_fp32, const GroupDesc* groups_dev, const int* tile_group_dev,
                      const int* tile_n0_dev, int ntiles_total) {
  if (M <= 0 || ntiles_total <= 0) return TG_OK;
  // the tile lists are built for 64-column tiles of each group
  if (variant == 0) {
    dim3 grid(ntiles_total, (M + 63) / 64, 1);
    if (w_is_fp32)
      hipLaunchKernelGGL((mfma_grouped_kernel<float, 64, 64, 2, 2>), grid, dim3(256), 0, s, M, X,
                         C, (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
    else
      hipLaunchKernelGGL((mfma_grouped_kernel<double, 64, 64, 2, 2>), grid, dim3(256), 0, s, M, X,
                         C, (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  } else if (variant == 1) {
    dim3 grid(ntiles_total, (M + 127) / 128, 1);
    if (w_is_fp32)
      hipLaunchKernelGGL((mfma_grouped_kernel<float, 128, 64, 4, 1>), grid, dim3(256), 0, s, M, X,
                         C, (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
    else
      hipLaunchKernelGGL((mfma_grouped_kernel<double, 128, 64, 4, 1>), grid, dim3(256), 0, s, M,
                         X, C, (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  } else if (variant == 2) {
    dim3 grid(ntiles_total, (M + 255) / 256, 1);
    if (w_is_fp32)
      hipLaunchKernelGGL((mfma_grouped_kernel<float, 256, 64, 4, 2>), grid, dim3(512), 0, s, M, X,
                         C, (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
    else
      hipLaunchKernelGGL((mfma_grouped_kernel<double, 256, 64, 4, 2>), grid, dim3(512), 0, s, M,
                         X, C, (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  } else if (variant == 3) {
    dim3 grid(ntiles_total, (M + 127) / 128, 1);
    if (w_is_fp32)
      hipLaunchKernelGGL((mfma_grouped_kernel<float, 128, 64, 4, 1, 2>), grid, dim3(256), 0, s, M,
                         X, C, (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
    else
      hipLaunchKernelGGL((mfma_grouped_kernel<double, 128, 64, 4, 1, 2>), grid, dim3(256), 0, s, M,
                         X, C, (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  } else if (variant == 4) {
    dim3 grid(ntiles_total, (M + 63) / 64, 1);
    if (w_is_fp32)
      hipLaunchKernelGGL((mfma_grouped_kernel<float, 64, 64, 2, 2, 3>), grid, dim3(256), 0, s, M, X,
                         C, (const float*)W, groups_dev, tile_group_dev, tile_n0_dev);
    else
      hipLaunchKernelGGL((mfma_grouped_kernel<double, 64, 64, 2, 2, 3>), grid, dim3(256), 0, s, M,
                         X, C, (const double*)W, groups_dev, tile_group_dev, tile_n0_dev);
  } else {
    tg_set_error("mfma grouped gemm: bad tile variant");
    return TG_E_ARG;
  }
  TG_HIP(hipGetLastError());
  return TG_OK;
}

int mfma_gemm_dd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, double alpha,
                 const double* A, int lda, long long sA, const double* B, int ldb, long long sB,
                 double beta, int use_beta, double* C, int ldc, long long sC, int batch, int divA,
                 int divB) {
  if (M <= 0 || N <= 0 || batch <= 0) return TG_OK;
  const int tiles_m = (M + 31) / 32, tiles_n = (N + 31) / 32;
  const dim3 grid(tiles_m * tiles_n, batch);
#define TG_DD(OA, OB)                                                                              \
  hipLaunchKernelGGL((mfma_dd_kernel<OA, OB>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, \
                     B, ldb, sB, beta, use_beta, C, ldc, sC, tiles_m, divA, divB)
  if (opa == OP_N && opb == OP_N) TG_DD(OP_N, OP_N);
  else if (opa == OP_N) TG_DD(OP_N, OP_T);
  else if (opb == OP_N) TG_DD(OP_T, OP_N);
  else TG_DD(OP_T, OP_T);
#undef TG_DD
  TG_HIP(hipGetLastError());
  return TG_OK;
}
