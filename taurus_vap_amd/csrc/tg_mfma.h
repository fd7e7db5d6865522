// FP64 matrix-core GEMM entry points (tg_mfma.hip).  `variant` selects the tile
// shape: 0 = 64x64 (2x2 waves), 1 = 48x48 (3 waves), 2 = 112x112 (7 waves),
// 3 = 96x64 (2x2 waves), 4 / 5 = 112x112 / 48x48 with every XCD given the same
// number of tiles, 6 / 7 = 48x48 bounded to four workgroups per CU, 8 = 64x32
// (2 waves, 256 VGPRs), 9 / 10 = 7 / 5 with both operands staged k-major in LDS;
// grouped:
// 0 = 64x64 (2x2 waves), 1 = 128x64 (4x1), 2 = 256x64 (4x2), 3 = 128x64 bounded to
// two workgroups per CU, 4 = 64x64 bounded to three.  Batch element b reads A + (b / divA) * sA and
// B + (b / divB) * sB (an operand shared by runs of divA consecutive elements).
#pragma once
#include "tg_internal.h"

constexpr int TG_MFMA_VARIANTS = 11;   // autotuned: 0..10
constexpr int TG_MFMA_GROUPED_VARIANTS = 5;

int mfma_gemm_zz(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const cplx* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch,
                 int divA = 1, int divB = 1);
int mfma_gemm_dz(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const double* A, int lda, long long sA, const cplx* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch,
                 int divA = 1, int divB = 1);
int mfma_gemm_zd(int variant, hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, cplx alpha,
                 const cplx* A, int lda, long long sA, const double* B, int ldb, long long sB,
                 cplx beta, int use_beta, cplx* C, int ldc, long long sC, int batch,
                 int divA = 1, int divB = 1);
int mfma_gemm_grouped(int variant, hipStream_t s, int M, const cplx* X, cplx* C, const void* W,
                      int w_is_fp32, const GroupDesc* groups_dev, const int* tile_group_dev,
                      const int* tile_n0_dev, int ntiles_total);
int mfma_gemm_dd(hipStream_t s, TgOp opa, TgOp opb, int M, int N, int K, double alpha,
                 const double* A, int lda, long long sA, const double* B, int ldb, long long sB,
                 double beta, int use_beta, double* C, int ldc, long long sC, int batch,
                 int divA = 1, int divB = 1);
