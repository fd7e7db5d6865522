// The gauge-angle loop of project_wavefunction(iopt=0) (src/module_projection.f90:
// 201-373) on gfx950, batched over angles, and the C ABI entry points.
//
// MI355X-first restatement of the loop body (the numbers are the reference's):
//  * The gauge rotation R = diag(e^{i phi_p} on protons, e^{i phi_n} on
//    neutrons) (generate_rotation_gauge :569-588) has two distinct phases, so
//      U_phi = U^T R U + V^T R* V      (calculate_thouless :694-696)
//      V_phi = V^T R U + U^T R* V      (:699-700)
//      R_n   = D0^T R D0               (calculate_overlap :854-855)
//    are linear combinations of angle-independent real matrices
//    (U^T Pi_p U, ... with Pi_p/Pi_n the proton/neutron projectors) built once
//    per iteration by six real GEMMs; per angle they cost O(n^2).
//  * A_phi = conj(V_phi U_phi^-1), U~ = U + V A_phi, V~ = V + U A_phi,
//    rho = V~ V^T, kappaLR = V~ U^T, kappaRL = V U~^T (:702-716,
//    calculate_densities module_wavefunctions.f90:1073-1087) are batched GEMMs
//    over all angles (angle = batch index).
//  * The TBME contraction is the grouped GEMM of tg_fields.hip.
//  * The overlap uses the reference's Pfaffian formula.  With no fully occupied
//    canonical state (nocc = 0) the Pfaffian of the 2*nbloc matrix M_reg is
//    reduced exactly to one of size nbloc (Schur complement of the v/u block):
//      detR * Pf(M_reg) = Pf(A1) Pf(A1^-1 - R2^T A1 R2)
//                       = prod sgn(x) (-1)^(nbloc/2) Pf(J + Q^T J Q),
//    Q = T R2 T, T = diag(|v/u|^1/2), J = blockdiag(sgn(x_j) [[0,1],[-1,0]]):
//    no LU, no inverse, a Pfaffian 8x cheaper.  With nocc > 0 the reference's
//    own sequence (LU/inverse of R2, M_reg, Pfaffian of size nreg) runs.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>

#include "tg_internal.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

// ---------------------------------------------------------------------------
// errors / buffers
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
void tg_set_error(const std::string& msg) { g_err = msg; }
extern "C" const char* tg_last_error(void) { return g_err.c_str(); }

static unsigned long long g_devbuf_epoch = 0;

int DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && p) return TG_OK;
  release();
  if (nbytes == 0) return TG_OK;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    tg_set_error(std::string("hipMalloc(") + std::to_string(nbytes) + "): " + hipGetErrorString(e));
    return TG_E_NOMEM;
  }
  bytes = nbytes;
  ++g_devbuf_epoch;
  return TG_OK;
}
unsigned long long devbuf_epoch() { return g_devbuf_epoch; }
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

namespace {

constexpr int NT = 256;
inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
// U_phi, V_phi of every item from the six real matrices of its state
// (R6 = [Kp Kn Lp Ln Mp Mn], Kx = U^T Pi_x U, Lx = V^T Pi_x V, Mx = V^T Pi_x U;
// product k of state s at (k * S + s) * n^2).  Item b0 + a is angle (b0 + a) % A
// of state (b0 + a) / A.  A thread takes one element for UV_AG consecutive
// items, so the real matrices (two of them read transposed) are loaded once per
// state rather than once per item.
constexpr int UV_AG = 16;
__global__ void build_uv_phi_kernel(int n, int na, long long b0, int A, int S,
                                    const double* __restrict__ R6, const cplx* __restrict__ ph,
                                    cplx* __restrict__ Uphi, cplx* __restrict__ Vphi,
                                    long long stride) {
  const long long nn = (long long)n * n;
  const long long r = (long long)blockIdx.x * NT + threadIdx.x;
  if (r >= nn) return;
  const int i = (int)(r % n), j = (int)(r / n);
  const long long rt = j + (long long)i * n;
  const long long ks = (long long)S * nn;
  const int a1 = min(na, (int)(blockIdx.y + 1) * UV_AG);
  long long cur = -1;
  double kp = 0, km = 0, kn = 0, knm = 0, mp = 0, mpm = 0, mn = 0, mnm = 0;
  for (int a = blockIdx.y * UV_AG; a < a1; ++a) {
    const long long b = b0 + a;
    const long long st = b / A, ang = b % A;
    if (st != cur) {
      const double* R = R6 + st * nn;
      const double Kp = R[r], Kn = R[ks + r], Lp = R[2 * ks + r], Ln = R[3 * ks + r];
      const double Mp = R[4 * ks + r], Mn = R[5 * ks + r];
      const double Mpt = R[4 * ks + rt], Mnt = R[5 * ks + rt];
      kp = Kp + Lp; km = Kp - Lp; kn = Kn + Ln; knm = Kn - Ln;
      mp = Mp + Mpt; mpm = Mp - Mpt; mn = Mn + Mnt; mnm = Mn - Mnt;
      cur = st;
    }
    const cplx ep = ph[2 * ang], en = ph[2 * ang + 1];
    cplx u, v;
    // e_p Kp + e_n Kn + conj(e_p) Lp + conj(e_n) Ln
    u.x = ep.x * kp + en.x * kn;
    u.y = ep.y * km + en.y * knm;
    // e_p Mp + e_n Mn + conj(e_p) Mp^T + conj(e_n) Mn^T
    v.x = ep.x * mp + en.x * mn;
    v.y = ep.y * mpm + en.y * mnm;
    Uphi[a * stride + r] = u;
    Vphi[a * stride + r] = v;
  }
}

// Ds(i,j) = sgn(i) D(i,j)
__global__ void row_sign_kernel(long long tot, int n, const double* __restrict__ sgn,
                                const double* __restrict__ D, double* __restrict__ Ds) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= tot) return;
  Ds[idx] = sgn[idx % n] * D[idx];
}

// One pass after the A_phi product: A_phi = conj(A_phi), and U~, V~ preset to the
// real U, V of the item's state (state (b0 + a) / A at U + s * nn) for the
// GEMMs that add V A_phi and U A_phi
__global__ void thouless_prep_kernel(long long nn, int na, long long b0, int A,
                                     const double* __restrict__ U, const double* __restrict__ V,
                                     cplx* __restrict__ Aphi, cplx* __restrict__ Ut,
                                     cplx* __restrict__ Vt) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= nn * na) return;
  const long long st = (b0 + idx / nn) / A, e = st * nn + idx % nn;
  Aphi[idx].y = -Aphi[idx].y;
  Ut[idx] = cmk(U[e], 0.0);
  Vt[idx] = cmk(V[e], 0.0);
}

// h_a = Gamma_a + H1
__global__ void add_h1_kernel(long long nn, int na, const cplx* __restrict__ H1,
                              const cplx* __restrict__ G, cplx* __restrict__ h) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= nn * na) return;
  h[idx] = cadd(G[idx], H1[idx % nn]);
}

// H20 = G - G^T per angle, through 32 x 32 LDS tiles (both reads coalesced).
// With G = U~^T (h V~ + fd/2 DeltaLR U~) - fd/2 V~^T DeltaRL V~ this is
// calculate_H20 (module_fields.f90:366-391): U~^T h V~ - (.)^T + fd (U~^T DeltaLR U~
// - V~^T DeltaRL V~), the two Delta terms being skew.
__global__ void __launch_bounds__(256)
skew_part_kernel(int n, const cplx* __restrict__ G, cplx* __restrict__ H) {
  __shared__ cplx ta[32][33], tb[32][33];
  const long long nn = (long long)n * n;
  const long long a = blockIdx.z;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  const cplx* g = G + a * nn;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;           // G(i, j): column j, row i (coalesced in i)
    ta[r][tx] = (i < n && j < n) ? g[i + (long long)j * n] : cmk(0.0, 0.0);
    const int i2 = j0 + tx, j2 = i0 + r;         // G(j0 + tx, i0 + r)
    tb[r][tx] = (i2 < n && j2 < n) ? g[i2 + (long long)j2 * n] : cmk(0.0, 0.0);
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < n && j < n) {
      const cplx x = ta[r][tx], xt = tb[tx][r];   // G(i, j), G(j, i)
      H[a * nn + i + (long long)j * n] = cmk(x.x - xt.x, x.y - xt.y);
    }
  }
}

// Per-angle traces (calculate_particle_number :78-143, calculate_expectval_energy
// module_hamiltonian.f90:1098-1126).  out[a*9 + k]:
// 0 prot, 1 neut, 2 Tr(rho_p^2), 3 Tr(kRL_p^T kLR_p), 4 Tr(rho_n^2),
// 5 Tr(kRL_n^T kLR_n), 6 Tr(H1 rho), 7 Tr(Gamma rho), 8 Tr(DeltaLR kRL^T)
// A workgroup sums the elements (i, j) of one band of 32 rows i of one angle,
// 32 columns j at a time; rho(j, i) comes from a 32 x 32 tile of rho staged in
// LDS, so every global read is coalesced.  part[(a * bands + band) * 9 + k];
// traces_sum_kernel adds the bands.
constexpr int TRT = 32;
__global__ void __launch_bounds__(NT)
traces_kernel(int n, const cplx* __restrict__ H1, const cplx* __restrict__ rho,
              const cplx* __restrict__ kLR, const cplx* __restrict__ kRL,
              const cplx* __restrict__ G, const cplx* __restrict__ dLR, long long stride,
              cplx* __restrict__ part) {
  const int band = blockIdx.x, bands = gridDim.x;
  const long long a = blockIdx.y;
  const int i0 = band * TRT, h = n / 2;
  const int tx = threadIdx.x % TRT, ty = threadIdx.x / TRT;
  const cplx* r_ = rho + a * stride;
  const cplx* kl = kLR + a * stride;
  const cplx* kr = kRL + a * stride;
  const cplx* g_ = G + a * stride;
  const cplx* dl = dLR + a * stride;
  __shared__ cplx T[TRT][TRT + 1];   // T[jj][ii] = rho(j0 + jj, i0 + ii)
  cplx acc[9];
  for (int k = 0; k < 9; ++k) acc[k] = cmk(0.0, 0.0);
  const int i = i0 + tx;
  const bool pi = i < h;
  for (int j0 = 0; j0 < n; j0 += TRT) {
    for (int ii = ty; ii < TRT; ii += NT / TRT) {
      const int jr = j0 + tx, ir = i0 + ii;
      T[tx][ii] = (ir < n && jr < n) ? r_[jr + (long long)ir * n] : cmk(0.0, 0.0);
    }
    __syncthreads();
    if (i < n)
      for (int jj = ty; jj < TRT; jj += NT / TRT) {
        const int j = j0 + jj;
        if (j >= n) break;
        const long long idx = i + (long long)j * n;
        const cplx rji = T[jj][tx];
        const bool pj = j < h;
        if (i == j) acc[pi ? 0 : 1] = cadd(acc[pi ? 0 : 1], r_[idx]);
        if (pi == pj) {
          cfma(acc[pi ? 2 : 4], r_[idx], rji);
          cfma(acc[pi ? 3 : 5], kr[idx], kl[idx]);
        }
        cfma(acc[6], H1[idx], rji);
        cfma(acc[7], g_[idx], rji);
        cfma(acc[8], dl[idx], kr[idx]);
      }
    __syncthreads();
  }
  __shared__ cplx red[NT];
  for (int k = 0; k < 9; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(a * bands + band) * 9 + k] = red[0];
    __syncthreads();
  }
}

// out[a*9 + k] = sum over the bands of part
__global__ void traces_sum_kernel(int na, int bands, const cplx* __restrict__ part,
                                  cplx* __restrict__ out) {
  const int idx = blockIdx.x * NT + threadIdx.x;
  if (idx >= na * 9) return;
  const int a = idx / 9, k = idx % 9;
  cplx s = cmk(0.0, 0.0);
  for (int b = 0; b < bands; ++b) s = cadd(s, part[((long long)a * bands + b) * 9 + k]);
  out[idx] = s;
}

// The reduced overlap's skew matrix per item is K = Q^T (J Q) + J with
// Q = T (e_p Pp + e_n Pn)[nemp:, nemp:] T (T = diag(Tsc), J the sgn-weighted
// symplectic unit: (J Q) row 2j = s_j Q row 2j+1, row 2j+1 = -s_j Q row 2j).
// Q is linear in the two gauge phases with real, angle-independent parts
// Qp = T Pp T, Qn = T Pn T, so
//   K = e_p^2 Qp^T J Qp + e_p e_n (X - X^T) + e_n^2 Qn^T J Qn + J,  X = Qp^T J Qn
// (J^T = -J): three real products per state, then one element-wise pass per item.
// kq_prep_kernel: Qp, Qn, J Qp, J Qn of states st0 .. st0 + nst - 1 (state s of
// the segment at + s nb^2; Pp, Pn of state st at + st n^2, its Tsc and sgn at
// + st * os).
__global__ void kq_prep_kernel(int n, int nemp, int nb, int nst, int st0,
                               const double* __restrict__ Pp0, const double* __restrict__ Pn0,
                               const double* __restrict__ Tsc0, const double* __restrict__ sgn0,
                               long long os, double* __restrict__ Qp, double* __restrict__ Qn,
                               double* __restrict__ JQp, double* __restrict__ JQn) {
  const long long mm = (long long)nb * nb;
  const long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= mm * nst) return;
  const int sl = (int)(idx / mm);
  const long long st = st0 + sl;
  const double* Pp = Pp0 + st * n * n;
  const double* Pn = Pn0 + st * n * n;
  const double* Tsc = Tsc0 + st * os;
  const double* sgn = sgn0 + st * os;
  const int r = (int)(idx % mm);
  const int i = r % nb, j = r / nb;
  auto q = [&](const double* Px, int ii) {
    return Tsc[ii] * Tsc[j] * Px[(ii + nemp) + (long long)(j + nemp) * n];
  };
  Qp[idx] = q(Pp, i);
  Qn[idx] = q(Pn, i);
  const double sj = sgn[i >> 1];
  const int pi = (i & 1) ? i - 1 : i + 1;
  const double f = (i & 1) ? -sj : sj;
  JQp[idx] = f * q(Pp, pi);
  JQn[idx] = f * q(Pn, pi);
}

// Kpn = X - X^T of every state of the segment (nst nb x nb matrices)
__global__ void kq_skew_kernel(int nb, int nst, const double* __restrict__ X,
                               double* __restrict__ Kpn) {
  const long long mm = (long long)nb * nb;
  const long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= mm * nst) return;
  const long long so = (idx / mm) * mm;
  const int r = (int)(idx % mm);
  const int i = r % nb, j = r / nb;
  Kpn[idx] = X[idx] - X[so + j + (long long)i * nb];
}

// K of the segment's items (item b0 + a: angle (b0 + a) % A of state (b0 + a) / A,
// its parts at + ((b0 + a) / A - st0) nb^2), strictly upper triangle only: the
// Pfaffian reads nothing else.
__global__ void kq_combine_kernel(int nb, int na, long long b0, int A, int st0,
                                  const cplx* __restrict__ ph, const double* __restrict__ Kpp,
                                  const double* __restrict__ Knn, const double* __restrict__ Kpn,
                                  const double* __restrict__ sgn0, long long os,
                                  cplx* __restrict__ K) {
  const long long mm = (long long)nb * nb;
  const long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= mm * na) return;
  const long long a = idx / mm;
  const int r = (int)(idx % mm);
  const int i = r % nb, j = r / nb;
  if (i >= j) return;
  const long long st = (b0 + a) / A, ang = (b0 + a) % A;
  const long long so = (st - st0) * mm;
  const cplx ep = ph[2 * ang], en = ph[2 * ang + 1];
  const cplx ep2 = cmul(ep, ep), epn = cmul(ep, en), en2 = cmul(en, en);
  const double kpp = Kpp[so + r], knn = Knn[so + r], kpn = Kpn[so + r];
  cplx v = cmk(ep2.x * kpp + epn.x * kpn + en2.x * knn, ep2.y * kpp + epn.y * kpn + en2.y * knn);
  if (j == i + 1 && (i & 1) == 0) v.x += sgn0[st * os + (i >> 1)];
  K[idx] = v;
}

// Rn2 = (e_p Pp + e_n Pn)[nemp:, nemp:] for the direct overlap
__global__ void build_rn2_kernel(int n, int nemp, int n2, int na, long long b0, int A,
                                 const double* __restrict__ Pp, const double* __restrict__ Pn,
                                 const cplx* __restrict__ ph, cplx* __restrict__ R2) {
  const long long mm = (long long)n2 * n2;
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= mm * na) return;
  const long long a = idx / mm;
  const long long st = (b0 + a) / A, ang = (b0 + a) % A;
  const int r = (int)(idx % mm);
  const int i = r % n2, j = r / n2;
  const cplx ep = ph[2 * ang], en = ph[2 * ang + 1];
  const long long o = (i + nemp) + (long long)(j + nemp) * n + st * n * n;
  R2[idx] = cmk(ep.x * Pp[o] + en.x * Pn[o], ep.y * Pp[o] + en.y * Pn[o]);
}

// M_reg of calculate_overlap (:891-918) from v/u and Rinv (n2 x n2)
// (v/u of the item's state: vou + ((b0 + a) / A) * vs)
__global__ void build_mreg_kernel(int n2, int nb0, int nb1, int nreg, int na, long long b0, int A,
                                  const double* __restrict__ vou0_, const double* __restrict__ vou1_,
                                  long long vs, const cplx* __restrict__ Rinv,
                                  cplx* __restrict__ M) {
  const long long mm = (long long)nreg * nreg;
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= mm * na) return;
  const long long a = idx / mm;
  const long long so = ((b0 + a) / A) * vs;
  const double* vou0 = vou0_ + so;
  const double* vou1 = vou1_ + so;
  const int r = (int)(idx % mm);
  const int i = r % nreg, j = r / nreg;   // 0-based row i, column j
  const cplx* Ri = Rinv + a * (long long)n2 * n2;
  cplx v = cmk(0.0, 0.0);
  if (i < nb1 && j < nb1) {
    // upper-left: M(i,i+1) = vou1(i) for even i, M(i,i-1) = vou1(i) for odd i
    if ((i & 1) == 0 && j == i + 1) v.x = vou1[i];
    if ((i & 1) == 1 && j == i - 1) v.x = vou1[i];
  } else if (i >= nb1 && j >= nb1) {
    const int jj = i - nb1;
    if ((i & 1) == 0 && j == i + 1) v.x = -vou0[jj];
    if ((i & 1) == 1 && j == i - 1) v.x = -vou0[jj];
  } else if (i < nb1 && j >= nb1) {
    const cplx x = Ri[i + (long long)(j - nb1) * n2];
    v = cmk(-x.x, -x.y);
  } else {
    v = Ri[j + (long long)(i - nb1) * n2];
  }
  M[idx] = v;
}

// sum_gauge scalars (:425-458): f1 = w*o, pnp += f1*{1, E, Z, N, Z2, N2}
// per-angle in: over[a], scal[a*9+k]; writes f1/f2 per angle and the 6 sums.
// One workgroup per state (blockIdx.x): items s*na .. s*na + na - 1, sums at pnp + 8 s.
__global__ void __launch_bounds__(NT)
gauge_scalars_kernel(int na, int n, cplx H0, const cplx* __restrict__ w, const cplx* __restrict__ over0,
                     const cplx* __restrict__ sc0, const cplx* __restrict__ pari0,
                     const double* __restrict__ mult, cplx* __restrict__ f120,
                     cplx* __restrict__ ener0, cplx* __restrict__ pnp0) {
  const long long st = blockIdx.x;
  const cplx* over = over0 + st * na;
  const cplx* sc = sc0 + st * na * 9;
  const cplx* pari = pari0 ? pari0 + st * na : nullptr;
  cplx* f12 = f120 + st * 2 * na;
  cplx* ener = ener0 + st * na;
  cplx* pnp = pnp0 + st * 8;
  __shared__ cplx red[7][NT];
  cplx acc[7];
  for (int k = 0; k < 7; ++k) acc[k] = cmk(0.0, 0.0);
  for (int a = threadIdx.x; a < na; a += NT) {
    const cplx* s = sc + (long long)a * 9;
    const cplx prot = s[0], neut = s[1];
    const cplx prot2 = csub(cadd(prot, cmul(prot, prot)), csub(s[2], s[3]));
    const cplx neut2 = csub(cadd(neut, cmul(neut, neut)), csub(s[4], s[5]));
    cplx e = cadd(H0, s[6]);
    e = cadd(e, cscale(cadd(s[7], s[8]), 0.5));
    const cplx f1 = cmul(w[a], over[a]);
    f12[2 * a] = f1;
    f12[2 * a + 1] = cmul(f1, e);
    ener[a] = e;
    // a conjugate pair adds t + conj(t) = 2 Re(t)
    const bool pair = mult[a] > 1.5;
    auto add = [&](cplx& s, cplx t) { s = pair ? cmk(s.x + 2.0 * t.x, s.y) : cadd(s, t); };
    add(acc[0], f1);
    add(acc[1], cmul(e, f1));
    add(acc[2], cmul(prot, f1));
    add(acc[3], cmul(neut, f1));
    add(acc[4], cmul(prot2, f1));
    add(acc[5], cmul(neut2, f1));
    if (pari) add(acc[6], cmul(pari[a], w[a]));   // pnp_pari += rot_pari * weight
  }
  for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = cadd(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 7) pnp[threadIdx.x] = red[threadIdx.x][0];
}

// pnp_A = sum f1 A, pnp_H20 = sum f1 H20, pnp_AH20 = sum f2 A   (angles in order)
// state blockIdx.y: items s*na .., sums at pA/pH/pAH + s * os
__global__ void gauge_matrices_kernel(long long nn, int na, const cplx* __restrict__ f120,
                                      const double* __restrict__ mult,
                                      const cplx* __restrict__ Aphi0, const cplx* __restrict__ H200,
                                      long long stride, cplx* __restrict__ pA0,
                                      cplx* __restrict__ pH0, cplx* __restrict__ pAH0, long long os) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= nn) return;
  const long long st = blockIdx.y;
  const cplx* f12 = f120 + st * 2 * na;
  const cplx* Aphi = Aphi0 + st * na * stride;
  const cplx* H20 = H200 + st * na * stride;
  cplx* pA = pA0 + st * os;
  cplx* pH = pH0 + st * os;
  cplx* pAH = pAH0 + st * os;
  cplx sa = cmk(0, 0), sh = cmk(0, 0), sah = cmk(0, 0);
  for (int a = 0; a < na; ++a) {
    const cplx f1 = f12[2 * a], f2 = f12[2 * a + 1];
    const cplx av = Aphi[a * stride + idx], hv = H20[a * stride + idx];
    if (mult[a] > 1.5) {        // conjugate pair: 2 Re(.)
      const cplx t1 = cmul(av, f1), t2 = cmul(hv, f1), t3 = cmul(av, f2);
      sa.x += 2.0 * t1.x;
      sh.x += 2.0 * t2.x;
      sah.x += 2.0 * t3.x;
    } else {
      cfma(sa, av, f1);
      cfma(sh, hv, f1);
      cfma(sah, av, f2);
    }
  }
  pA[idx] = sa;
  pH[idx] = sh;
  pAH[idx] = sah;
}

// calculate_projected_H20 (:729-759); imag[block] = max Im over the block
__global__ void __launch_bounds__(NT)
projected_h20_kernel(int n, const cplx* __restrict__ pnp0, const cplx* __restrict__ pA0,
                     const cplx* __restrict__ pH0, const cplx* __restrict__ pAH0,
                     double* __restrict__ H20v0, double* __restrict__ imag0, long long os,
                     long long is) {
  // state blockIdx.y: sums at pnp + 8 s, matrices at + s * os (complex), imag at + s * is
  const long long st = blockIdx.y;
  const cplx* pnp = pnp0 + 8 * st;
  const cplx* pA = pA0 + st * os;
  const cplx* pH = pH0 + st * os;
  const cplx* pAH = pAH0 + st * os;
  double* H20v = reinterpret_cast<double*>(reinterpret_cast<cplx*>(H20v0) + st * os);
  double* imag = imag0 + st * is;
  const long long nn = (long long)n * n;
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  const cplx over = pnp[0];
  const cplx oinv = cinv(over);
  const cplx ener = cmul(pnp[1], oinv);
  double im = -INFINITY;
  if (idx < nn) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const long long t = j + (long long)i * n;
    const cplx A = cscale(cmul(csub(pA[idx], pA[t]), oinv), 0.5);
    const cplx H = cscale(cmul(csub(pH[idx], pH[t]), oinv), 0.5);
    const cplx AH = cscale(cmul(csub(pAH[idx], pAH[t]), oinv), 0.5);
    const cplx x = cadd(csub(AH, cmul(ener, A)), H);
    H20v[idx] = x.x;
    const cplx y = cmul(cadd(csub(pAH[idx], cmul(ener, pA[idx])), pH[idx]), oinv);
    im = y.y;
  }
  __shared__ double red[NT];
  red[threadIdx.x] = im;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) imag[blockIdx.x] = red[0];
}

// overlap assembly: o = pre * phase * exp(logabs + lnorm) (lnorm of the item's
// state at lnorm0 + ((b0 + a) / A) * ls)
__global__ void overlap_finish_kernel(int na, long long b0, int A, const double* __restrict__ logabs,
                                      const cplx* __restrict__ phase, const int* __restrict__ zero,
                                      const cplx* __restrict__ pre, const double* __restrict__ lnorm0,
                                      long long ls, cplx* __restrict__ over) {
  int a = blockIdx.x * NT + threadIdx.x;
  if (a >= na) return;
  if (zero[a]) { over[a] = cmk(0.0, 0.0); return; }
  const double m = exp(logabs[a] + lnorm0[((b0 + a) / A) * ls]);
  const cplx z = cmul(pre[a], phase[a]);
  over[a] = cmk(m * z.x, m * z.y);
}

__global__ void fill_cplx_kernel(long long na, cplx v, cplx* __restrict__ x) {
  long long a = (long long)blockIdx.x * NT + threadIdx.x;
  if (a < na) x[a] = v;
}

// x[a] = the value of the item's state (v0 + ((b0 + a) / A) * vs)
__global__ void fill_from_kernel(int na, long long b0, int A, const cplx* __restrict__ v0,
                                 long long vs, cplx* __restrict__ x) {
  int a = blockIdx.x * NT + threadIdx.x;
  if (a < na) x[a] = v0[((b0 + a) / A) * vs];
}

// Pfaffian of an empty block: log|pf| = 0, phase 1
__global__ void unit_pfaffian_kernel(int na, double* __restrict__ logabs, cplx* __restrict__ phase,
                                     int* __restrict__ zero) {
  int a = blockIdx.x * NT + threadIdx.x;
  if (a >= na) return;
  logabs[a] = 0.0;
  phase[a] = cmk(1.0, 0.0);
  zero[a] = 0;
}

__global__ void copy_real_kernel(long long nn, const double* __restrict__ src, double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i < nn) dst[i] = src[i];
}

// pre[a] = signfac * det[a] for the direct overlap
__global__ void scale_det_kernel(int na, double s, const cplx* __restrict__ det, cplx* __restrict__ pre) {
  int a = blockIdx.x * NT + threadIdx.x;
  if (a < na) pre[a] = cscale(det[a], s);
}

// ---- observables of the final projection (project_wavefunction iopt = 1)

// block-wide sum of K complex accumulators into out[0..K) (NT threads)
template <int K>
__device__ void block_sum(cplx (&acc)[K], cplx* __restrict__ out) {
  __shared__ cplx red[NT];
  for (int k = 0; k < K; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// calculate_decompo_energy (module_hamiltonian.f90:1134-1260) of angle blockIdx.x:
// out[12] = 1b_p, 1b_n, PH_pp, PH_pn, PH_np, PH_nn, PP_pp, PP_pn, PP_np, PP_nn,
// E_p, E_n with PH_xy = 1/2 Tr(Gamma_xy rho_yx), PP_xy = 1/2 Tr(Delta_xy kRL_xy^T)
// (x, y = the proton / neutron halves of the single-particle index)
__global__ void __launch_bounds__(NT)
ecomp_kernel(int n, const cplx* __restrict__ H1, const cplx* __restrict__ G0,
             const cplx* __restrict__ rho0, const cplx* __restrict__ dLR0,
             const cplx* __restrict__ kRL0, cplx* __restrict__ out) {
  const long long nn = (long long)n * n, a = blockIdx.x;
  const cplx* G = G0 + a * nn;
  const cplx* rho = rho0 + a * nn;
  const cplx* dLR = dLR0 + a * nn;
  const cplx* kRL = kRL0 + a * nn;
  const int h = n / 2;
  cplx acc[10];   // OB_p, OB_n, PH[2][2], PP[2][2]
  for (int k = 0; k < 10; ++k) acc[k] = cmk(0.0, 0.0);
  for (long long idx = threadIdx.x; idx < nn; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const int bi = i >= h, bj = j >= h;
    const cplx rji = rho[j + (long long)i * n];
    cfma(acc[bi], H1[idx], rji);
    cfma(acc[2 + 2 * bi + bj], G[idx], rji);
    cfma(acc[6 + 2 * bi + bj], dLR[idx], kRL[idx]);
  }
  __shared__ cplx res[10];
  block_sum<10>(acc, res);
  if (threadIdx.x == 0) {
    cplx* o = out + a * 12;
    o[0] = res[0];
    o[1] = res[1];
    for (int k = 0; k < 4; ++k) o[2 + k] = cscale(res[2 + k], 0.5);
    for (int k = 0; k < 4; ++k) o[6 + k] = cscale(res[6 + k], 0.5);
    o[10] = cadd(res[0], cscale(cadd(cadd(res[2], res[3]), cadd(res[6], res[7])), 0.5));
    o[11] = cadd(res[1], cscale(cadd(cadd(res[4], res[5]), cadd(res[8], res[9])), 0.5));
  }
}

// calculate_occupation_number (module_particlenumber.f90:157-176): diagonal of
// rho summed per shell and isospin.  shell[i], iso[i] of each (internal) state;
// out[a * 2 * nsh + 2 * s + t]
__global__ void occnum_kernel(int n, int nsh, const int* __restrict__ shell,
                              const int* __restrict__ iso, const cplx* __restrict__ rho0,
                              cplx* __restrict__ out) {
  const long long nn = (long long)n * n, a = blockIdx.x;
  const cplx* rho = rho0 + a * nn;
  for (int q = threadIdx.x; q < 2 * nsh; q += NT) {
    const int sh = q / 2, t = q % 2;
    cplx sum = cmk(0.0, 0.0);
    for (int i = 0; i < n; ++i)
      if (shell[i] == sh && iso[i] == t) sum = cadd(sum, rho[i + (long long)i * n]);
    out[a * 2 * nsh + q] = sum;
  }
}

// calculate_expectval_obo_cplx (module_operators.f90:146-164) for real operator
// blockIdx.y of angle blockIdx.x: out = (sum_{i in p} (O rho)_ii, sum_{i in n} (O rho)_ii)
__global__ void __launch_bounds__(NT)
obo_trace_kernel(int n, const double* __restrict__ ops, const cplx* __restrict__ rho0,
                 cplx* __restrict__ out) {
  const long long nn = (long long)n * n, a = blockIdx.x, k = blockIdx.y;
  const double* O = ops + k * nn;
  const cplx* rho = rho0 + a * nn;
  const int h = n / 2;
  cplx acc[2] = {cmk(0.0, 0.0), cmk(0.0, 0.0)};
  for (long long idx = threadIdx.x; idx < nn; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const cplx r = rho[j + (long long)i * n];
    const double o = O[idx];
    acc[i >= h].x = fma(o, r.x, acc[i >= h].x);
    acc[i >= h].y = fma(o, r.y, acc[i >= h].y);
  }
  __shared__ cplx res[2];
  block_sum<2>(acc, res);
  if (threadIdx.x == 0) {
    out[(a * gridDim.y + k) * 2] = res[0];
    out[(a * gridDim.y + k) * 2 + 1] = res[1];
  }
}

// traces of calculate_expectval_obos_cplx (module_operators.f90:227-257) for a
// real operator O (scalar factors applied by the caller), angle blockIdx.x:
//   T1 = Tr(O rho), T2 = Tr(O2 rho), T5 = Tr(O rho O rho) = sum (O rho)_ij (O rho)_ji,
//   T6 = Tr(O^T kRL^T O kLR) = sum_ij (kRL O)_ij (O kLR)_ij
// with X = O rho, Y = kRL O, Z = O kLR per angle.
__global__ void __launch_bounds__(NT)
obos_trace_kernel(int n, const cplx* __restrict__ X0, const cplx* __restrict__ Y0,
                  const cplx* __restrict__ Z0, const double* __restrict__ O2,
                  const cplx* __restrict__ rho0, cplx* __restrict__ out) {
  const long long nn = (long long)n * n, a = blockIdx.x;
  const cplx* X = X0 + a * nn;
  const cplx* Y = Y0 + a * nn;
  const cplx* Z = Z0 + a * nn;
  const cplx* rho = rho0 + a * nn;
  cplx acc[4];
  for (int k = 0; k < 4; ++k) acc[k] = cmk(0.0, 0.0);
  for (long long idx = threadIdx.x; idx < nn; idx += NT) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const long long t = j + (long long)i * n;
    if (i == j) acc[0] = cadd(acc[0], X[idx]);
    const cplx r = rho[t];
    acc[1].x = fma(O2[idx], r.x, acc[1].x);
    acc[1].y = fma(O2[idx], r.y, acc[1].y);
    cfma(acc[2], X[idx], X[t]);
    cfma(acc[3], Y[idx], Z[idx]);
  }
  __shared__ cplx res[4];
  block_sum<4>(acc, res);
  if (threadIdx.x < 4) out[a * 4 + threadIdx.x] = res[threadIdx.x];
}

// pnp_rho of sum_gauge (iopt = 1): sum_a f1_a rho_a (f1 = weight x overlap)
__global__ void rho_sum_kernel(long long nn, int na, const cplx* __restrict__ f12,
                               const cplx* __restrict__ rho, cplx* __restrict__ out) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= nn) return;
  cplx s = cmk(0.0, 0.0);
  for (int a = 0; a < na; ++a) cfma(s, rho[a * nn + idx], f12[2 * a]);
  out[idx] = s;
}

}  // namespace

// ===========================================================================
// host orchestration
// ===========================================================================
namespace {

}  // namespace

ScopedTiming::ScopedTiming(tg_ctx* c_, const char* k, hipStream_t st)
    : c(c_), key(k), s(st ? st : c_->stream) {
  if (c->timing && hipEventCreate(&e0) == hipSuccess) (void)hipEventRecord(e0, s);
}
ScopedTiming::~ScopedTiming() {
  if (!c->timing || !e0) return;
  hipEvent_t e1 = nullptr;
  if (hipEventCreate(&e1) != hipSuccess) return;
  (void)hipEventRecord(e1, s);
  c->pending.push_back({std::string(key), e0, e1});
}
int timing_collect(tg_ctx* c) {
  if (c->pending.empty()) return TG_OK;
  TG_HIP(hipStreamSynchronize(c->stream));
  TG_HIP(hipStreamSynchronize(c->stream2));
  TG_HIP(hipStreamSynchronize(c->stream3));
  for (auto& p : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
      c->tsum[p.key] += ms;
      c->tcount[p.key] += 1;
    }
    (void)hipEventDestroy(p.e0);
    (void)hipEventDestroy(p.e1);
  }
  c->pending.clear();
  return TG_OK;
}

namespace {

// calculate_gradient (module_gradient.f90:122-141): Z = -eta (H20v - sum_j
// lambda_j C20_j) + mu Zm1; part[b] = sum Z^2.
__global__ void __launch_bounds__(NT)
gradient_kernel(long long nn, double eta0, double mu0, const double* __restrict__ etamu, int ncons,
                const double* __restrict__ lam0, long long ls, const double* __restrict__ c200,
                long long cs, const double* __restrict__ H20v0, long long hs,
                const double* __restrict__ Zm10, double* __restrict__ Z0,
                double* __restrict__ part0, long long ps) {
  // state blockIdx.y: lambda at + s * ls, constraint_20 at + s * cs, field_H20v at
  // + s * hs, Z / Zm1 at + s * nn, partial sums at + s * ps; eta, mu per state
  // from etamu (gradient_type 1, 2) or the common eta0, mu0
  const long long st = blockIdx.y;
  const double eta = etamu ? etamu[2 * st] : eta0, mu = etamu ? etamu[2 * st + 1] : mu0;
  const double* lam = lam0 + st * ls;
  const double* c20 = c200 + st * cs;
  const double* H20v = H20v0 + st * hs;
  const double* Zm1 = Zm10 + st * nn;
  double* Z = Z0 + st * nn;
  double* part = part0 + st * ps;
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  double z2 = 0.0;
  if (idx < nn) {
    double sg = 0.0;
    for (int j = 0; j < ncons; ++j) sg += lam[j] * c20[(long long)j * nn + idx];
    const double z = -eta * (H20v[idx] - sg) + mu * Zm1[idx];
    Z[idx] = z;
    z2 = z * z;
  }
  __shared__ double red[NT];
  red[threadIdx.x] = z2;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int sign_pow(long long e) { return (e % 2 == 0) ? 1 : -1; }

// per-item complex buffers of the loop body
enum { W_UPHI, W_VPHI, W_UT, W_VT, W_RHO, W_KLR, W_KRL, W_G, W_H, W_DLR, W_DRL, W_T, W_X, W_Y,
       W_Z, W_COUNT };

// Overlap shape of one state (calculate_overlap with state 0 = state 1,
// :803-818): sizes, signfac and the prefactor of the reduced Pfaffian form.
struct OvlShape {
  int nocc = 0, nemp = 0, nbloc = 0, n0 = 0, n2 = 0, signfac = 1;
  bool reduced = false;
  double pre = 1.0;
};

OvlShape ovl_shape(int n, const StateInfo& si) {
  OvlShape o;
  o.nocc = si.nocc;
  o.nemp = si.nemp;
  o.nbloc = n - o.nocc - o.nemp;
  o.reduced = o.nocc == 0 && o.nbloc > 0 && (o.nbloc % 2 == 0);
  o.n0 = n - o.nocc;
  o.n2 = n - o.nemp;
  const long long e1 = (long long)o.n2 * (o.n2 + 1) / 2;
  const long long e2 = (long long)o.nocc * (o.nocc - 1) / 2;
  const long long e3 = (long long)o.nocc * (o.nocc + 1) + (long long)o.nocc * (o.n0 + o.n2) +
                       (long long)o.nocc * o.n0;
  o.signfac = sign_pow(e1) * sign_pow(e2) * sign_pow(e3);
  double prod_sgn = 1.0;
  if (o.reduced)
    for (int q = 0; q < o.nbloc / 2; ++q)
      if (si.voveru[o.nemp + 2 * q] < 0) prod_sgn = -prod_sgn;
  o.pre = o.signfac * prod_sgn * sign_pow(o.nbloc / 2);
  return o;
}

// Device pointers and host scalars of one projection of the resident states
// (pure host arithmetic on the context; no HIP calls).  The batch holds
// B = S * A items, item b being angle b % A of state b / A.
struct RunPlan {
  int n, h, A, S, B, chunk;
  long long nn, os, outs, is;               // strides: overlap set-up, outputs, imag
  bool proj, need_ovl, uniform;
  std::vector<OvlShape> sh;
  cplx *work, *Akeep, *Hkeep;
  cplx *d_ph, *d_tr, *d_over, *d_ener, *d_f12, *d_pnp, *d_pre, *d_det, *d_phase, *d_w, *d_pari;
  cplx* d_trp;   // per-band partial traces of one chunk
  double *d_logabs, *d_imag, *d_mult;
  int *d_info, *d_zero, *d_perm;
  double *d_T, *d_sg, *d_ovac, *d_vou;      // overlap set-up of state 0 (state s at + s * os)
  cplx* d_pre0;                             // state s at + s * os / 2
  size_t work_bytes, scal_bytes, int_bytes, ovl_bytes, pfw_bytes;
};

// Items per launch sequence.  All items at once unless a limit applies (the
// caller's batch limit, or the free-memory chunk of the upload).  A chunk holds
// whole states when the context has several, so that every launch's per-state
// operands are runs of whole states; with one state any chunk size works.
int chunk_items(const tg_ctx* c, int B, int A) {
  int lim = B;
  if (c->batch_limit > 0) lim = std::min(lim, c->batch_limit);
  if (c->mem_chunk > 0) lim = std::min(lim, c->mem_chunk);
  if (lim >= B) return B;
  if (c->ns > 1) return std::max(1, lim / A) * A;
  return std::max(1, lim);
}

RunPlan make_plan(tg_ctx* c) {
  RunPlan P{};
  P.n = c->n;
  P.h = P.n / 2;
  P.A = c->A;
  P.S = c->ns;
  P.B = P.S * P.A;
  P.nn = (long long)P.n * P.n;
  P.chunk = c->plain ? 1 : chunk_items(c, P.B, P.A);
  P.os = 3 * (long long)P.n + 8;
  P.outs = tg_out_stride(P.n);
  P.is = tg_part_len(P.n);
  P.proj = std::max(c->Mphip, c->Mphin) > 1;
  P.need_ovl = P.proj || c->iprint;
  P.uniform = true;
  for (int s = 0; s < P.S; ++s) {
    P.sh.push_back(ovl_shape(P.n, c->st[s]));
    if (P.sh[s].nocc != P.sh[0].nocc || P.sh[s].nemp != P.sh[0].nemp) P.uniform = false;
  }
  const int A = P.A, B = P.B, S = P.S, n = P.n;
  const long long nn = P.nn;
  const size_t wsz = (size_t)W_COUNT * P.chunk * nn;
  P.work_bytes = sizeof(cplx) * (wsz + 2 * (size_t)B * nn);
  // scalars: ph (2A) | traces (9B) | over (B) | ener (B) | f12 (2B) | pnp (8S) | pre (B)
  //          | det (B) | phase (B) | w (A) | pari (B) | logabs (B doubles)
  //          | imag (S * is doubles) | mult (A doubles) | pad to 16 bytes
  //          | partial traces (9 * chunk * bands)
  const size_t bands = (n + TRT - 1) / TRT;
  P.scal_bytes = sizeof(cplx) * (3 * (size_t)A + 17 * (size_t)B + 8 * (size_t)S) +
                 sizeof(double) * ((size_t)B + (size_t)S * P.is + A + 1) +
                 sizeof(cplx) * 9 * (size_t)P.chunk * bands;
  P.int_bytes = sizeof(int) * ((size_t)B * (n + 2) + 8);
  // overlap set-up per state: T (n) | sg (n) | ovac (2) | pre0 (cplx) | v/u (n)
  P.ovl_bytes = sizeof(double) * (size_t)S * P.os;
  // overlap work (its own buffers: the overlap runs on the second stream,
  // concurrently with the Thouless / field / H20 chain), per item of a segment
  //   reduced: K | Pfaffian work, nb^2 complex each, then per state of the
  //            segment Qp | Qn | JQp | JQn | Kpp | Knn | Xpn, nb^2 doubles each
  //   direct:  R2 | M_reg | Pfaffian work | det | info (int) | perm (n int)
  P.pfw_bytes = 0;
  if (P.need_ovl)
    for (const OvlShape& o : P.sh) {
      size_t per;
      if (o.reduced) {
        const size_t mm = (size_t)o.nbloc * o.nbloc;
        per = sizeof(cplx) * 2 * mm;
        const size_t nst = (size_t)std::min(S, P.chunk / A + 2);   // states a segment spans
        P.pfw_bytes = std::max(P.pfw_bytes, per * (size_t)P.chunk + sizeof(double) * 7 * mm * nst + 256);
        continue;
      } else {
        const long long nreg = 2 * (long long)o.nbloc;
        per = sizeof(cplx) * ((size_t)o.n2 * o.n2 + 2 * (size_t)(nreg * nreg) + 1) +
              sizeof(int) * (size_t)(n + 1);
      }
      P.pfw_bytes = std::max(P.pfw_bytes, per * (size_t)P.chunk + 256);
    }
  P.work = c->d_work.as<cplx>();
  P.Akeep = P.work + wsz;
  P.Hkeep = P.Akeep + (size_t)B * nn;
  P.d_ph = c->d_scal.as<cplx>();
  P.d_tr = P.d_ph + 2 * A;
  P.d_over = P.d_tr + 9 * (size_t)B;
  P.d_ener = P.d_over + B;
  P.d_f12 = P.d_ener + B;
  P.d_pnp = P.d_f12 + 2 * (size_t)B;
  P.d_pre = P.d_pnp + 8 * (size_t)S;
  P.d_det = P.d_pre + B;
  P.d_phase = P.d_det + B;
  P.d_w = P.d_phase + B;
  P.d_pari = P.d_w + A;
  P.d_logabs = reinterpret_cast<double*>(P.d_pari + B);
  P.d_imag = P.d_logabs + B;
  P.d_mult = P.d_imag + (size_t)S * P.is;
  {
    const size_t end = (size_t)((const char*)(P.d_mult + A) - (const char*)c->d_scal.p);
    P.d_trp = reinterpret_cast<cplx*>((char*)c->d_scal.p + ((end + 15) / 16) * 16);
  }
  P.d_info = c->d_int.as<int>();
  P.d_zero = P.d_info + B;
  P.d_perm = P.d_zero + B;
  P.d_T = c->d_ovl.as<double>();
  P.d_sg = P.d_T + n;
  P.d_ovac = P.d_sg + n;
  P.d_pre0 = reinterpret_cast<cplx*>(P.d_ovac + 2);
  P.d_vou = P.d_ovac + 4;
  return P;
}

// Allocations, uploads of the state-dependent set-up (angles, weights, overlap
// scaling, ln-norm and prefactor of every state) and, on the first call at a
// size, the GEMM tile autotuning.  Runs eagerly, never under stream capture.
// Overlap set-up of every state from its canonical data: per state T (n) | sg (n)
// | ovac (2) | pre0 (cplx) | v/u (n), as the reduced Pfaffian and the reference
// sequence read them.
std::vector<double> overlap_setup(const tg_ctx* c, const RunPlan& P) {
  const int n = P.n, S = P.S;
  std::vector<double> ovl((size_t)S * P.os, 0.0);
  for (int st = 0; st < S; ++st) {
    const OvlShape& o = P.sh[st];
    const StateInfo& si = c->st[st];
    double* Tsc = ovl.data() + (size_t)st * P.os;
    double* sg = Tsc + n;
    if (o.reduced)
      for (int q = 0; q < o.nbloc / 2; ++q) {
        const double x = si.voveru[o.nemp + 2 * q];
        sg[q] = (x < 0) ? -1.0 : 1.0;
        Tsc[2 * q] = Tsc[2 * q + 1] = sqrt(fabs(x));
      }
    sg[n] = (o.nbloc == 0) ? 0.0 : si.ovac;                             // d_ovac
    sg[n + 2] = o.pre;                                                  // d_pre0 (re, im = 0)
    for (int i = 0; i < o.nbloc; ++i) sg[n + 4 + i] = si.voveru[o.nemp + i];   // d_vou
  }
  return ovl;
}

int project_upload(tg_ctx* c) {
  // When all items do not fit into the free device memory, the per-item work
  // matrices are reused chunk by chunk (A_phi and H20 of every item are kept
  // for sum_gauge): the memory-short mode of the batched path.
  c->mem_chunk = 0;
  if (!c->plain) {
    RunPlan F = make_plan(c);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      const size_t have = free_b + c->d_work.bytes + c->d_pfw.bytes + c->d_blk.bytes;
      const size_t nn = (size_t)F.nn;
      const size_t per_item = sizeof(cplx) * (W_COUNT * nn) + (F.chunk ? F.pfw_bytes / F.chunk : 0) +
                              sizeof(cplx) * 4 * nn;   // contraction layouts (upper bound)
      const size_t fixed = sizeof(cplx) * 2 * (size_t)F.B * nn;
      const size_t budget = (size_t)(0.85 * (double)have);
      if (fixed + per_item * (size_t)F.chunk > budget && budget > fixed + per_item)
        c->mem_chunk = (int)std::max<size_t>(1, (budget - fixed) / per_item);
    }
  }
  RunPlan P = make_plan(c);
  const int n = P.n, A = P.A, S = P.S;
  const long long nn = P.nn;
  hipStream_t s = c->stream;
  TG_TRY(c->d_work.ensure(P.work_bytes));
  TG_TRY(c->d_scal.ensure(P.scal_bytes));
  TG_TRY(c->d_int.ensure(P.int_bytes));
  TG_TRY(c->d_ovl.ensure(P.ovl_bytes));
  TG_TRY(c->d_out.ensure(sizeof(cplx) * (size_t)S * P.outs));
  if (P.pfw_bytes) TG_TRY(c->d_pfw.ensure(P.pfw_bytes));
  if (c->grad_on) TG_TRY(tg_ensure_z(c));
  P = make_plan(c);   // pointers of the (re)allocated buffers

  std::vector<cplx> hph(2 * A), hw(A);
  for (int a = 0; a < A; ++a) {
    hph[2 * a] = cmk(cos(c->phip[a]), sin(c->phip[a]));
    hph[2 * a + 1] = cmk(cos(c->phin[a]), sin(c->phin[a]));
    hw[a] = c->weight[a];
  }
  std::vector<double> ovl = overlap_setup(c, P);
  TG_HIP(hipMemcpyAsync(P.d_ph, hph.data(), sizeof(cplx) * 2 * A, hipMemcpyHostToDevice, s));
  TG_HIP(hipMemcpyAsync(P.d_w, hw.data(), sizeof(cplx) * A, hipMemcpyHostToDevice, s));
  TG_HIP(hipMemcpyAsync(P.d_mult, c->mult.data(), sizeof(double) * A, hipMemcpyHostToDevice, s));
  TG_HIP(hipMemcpyAsync(P.d_T, ovl.data(), sizeof(double) * ovl.size(), hipMemcpyHostToDevice, s));
  TG_HIP(hipStreamSynchronize(s));   // host vectors go out of scope

  // first iteration at this size: autotune the tile shapes of the batched GEMMs
  // of the loop body on zeroed scratch operands (W_T, W_X -> W_Y)
  const int chunk = P.chunk;
  if (c->tuned_n != n || c->tuned_chunk != chunk) {
    auto Wb = [&](int k) { return P.work + (size_t)k * chunk * nn; };
    const double* U = c->d_U.as<double>();
    cplx* Ta = Wb(W_T);
    cplx* Xa = Wb(W_X);
    cplx* Ya = Wb(W_Y);
    const long long tot = nn * chunk;
    hipLaunchKernelGGL(fill_cplx_kernel, dim3(nblk(tot)), dim3(NT), 0, s, tot, cmk(0, 0), Ta);
    hipLaunchKernelGGL(fill_cplx_kernel, dim3(nblk(tot)), dim3(NT), 0, s, tot, cmk(0, 0), Xa);
    TG_HIP(hipGetLastError());
    TG_TRY(gemm_tune_zz(s, OP_N, OP_N, n, n, n, Xa, n, nn, Ta, n, nn, Ya, n, nn, chunk));
    TG_TRY(gemm_tune_zz(s, OP_T, OP_N, n, n, n, Xa, n, nn, Ta, n, nn, Ya, n, nn, chunk));
    TG_TRY(gemm_tune_dz(s, OP_N, OP_N, n, n, n, U, n, 0, Xa, n, nn, Ya, n, nn, chunk));
    TG_TRY(gemm_tune_dz(s, OP_N, OP_T, n, n, n, U, n, 0, Xa, n, nn, Ya, n, nn, chunk));
    TG_TRY(gemm_tune_zd(s, OP_N, OP_T, n, n, n, Xa, n, nn, U, n, 0, Ya, n, nn, chunk));
    c->tuned_n = n;
    c->tuned_chunk = chunk;
  }
  c->state_dirty = 0;
  return TG_OK;
}

}  // namespace

// The overlap set-up of new canonical data on stream s (the deferred overlap of
// tg_iterate_canonical): the overlap work buffer for the new shapes and the
// per-state scaling; the rest of the projection's set-up is unchanged.
int upload_overlap_setup(tg_ctx* c, hipStream_t s) {
  RunPlan P = make_plan(c);
  if (P.pfw_bytes) TG_TRY(c->d_pfw.ensure(P.pfw_bytes));
  P = make_plan(c);
  std::vector<double> ovl = overlap_setup(c, P);
  TG_HIP(hipMemcpyAsync(P.d_T, ovl.data(), sizeof(double) * ovl.size(), hipMemcpyHostToDevice, s));
  TG_HIP(hipStreamSynchronize(s));   // the host vector goes out of scope
  return TG_OK;
}

namespace {

// The projection proper: kernel launches only (capturable into a hipGraph).
// Buffers and the uploaded set-up come from project_upload.
int project_run(tg_ctx* c) {
  const RunPlan P = make_plan(c);
  const int n = P.n, h = P.h, A = P.A, S = P.S, B = P.B;
  const long long nn = P.nn;
  hipStream_t s = c->stream;
  const int chunk = P.chunk;
  const bool proj = P.proj;
  const double* U = c->d_U.as<double>();
  const double* V = c->d_V.as<double>();
  double* R6 = c->d_real6.as<double>();      // Kp Kn Lp Ln Mp Mn Pp Pn [parity Pp Pn] [Ds]
  const long long ks = (long long)S * nn;    // stride of one product over the S states
  const cplx one = cmk(1, 0), zero = cmk(0, 0);
  if (chunk != B && S != 1 && chunk % A != 0) {
    tg_set_error("projection: a chunk must hold whole states");
    return TG_E_ARG;
  }
  // tg_iterate_canonical: the canonical basis of the state runs beside the
  // Thouless / field / H20 chain; everything that reads D0 or the overlap shapes
  // (the D0 products, the overlap chain) follows it on the second stream
  const bool defer = c->defer_overlap && c->defer_hook && chunk == B && !c->plain;

  // D0^T Pi D0 per species (R_n = D0^T R D0 of calculate_overlap :854-855) and,
  // when printing, D0^T diag((-1)^l Pi) D0 (generate_rotation_parity :607-625)
  auto d_products = [&](hipStream_t st) -> int {
    const double* D = c->d_D.as<double>();   // written by a deferred canonical basis
    for (int sp = 0; sp < 2; ++sp) {
      const int r0 = sp * h;
      TG_TRY(gemm_dd(st, OP_T, OP_N, n, n, h, 1.0, D + r0, n, nn, D + r0, n, nn, 0.0, R6 + (6 + sp) * ks, n, nn, S));
    }
    if (c->iprint) {
      double* Ds = R6 + 10 * ks;
      hipLaunchKernelGGL(row_sign_kernel, dim3(nblk(ks)), dim3(NT), 0, st, ks, n,
                         c->d_psign.as<double>(), D, Ds);
      TG_HIP(hipGetLastError());
      for (int sp = 0; sp < 2; ++sp) {
        const int r0 = sp * h;
        TG_TRY(gemm_dd(st, OP_T, OP_N, n, n, h, 1.0, D + r0, n, nn, Ds + r0, n, nn, 0.0, R6 + (8 + sp) * ks, n, nn, S));
      }
    }
    return TG_OK;
  };

  // ---- angle-independent real products (once per iteration, every state)
  {
    ScopedTiming tm(c, "gemm");
    for (int sp = 0; sp < 2; ++sp) {
      const int r0 = sp * h;
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, h, 1.0, U + r0, n, nn, U + r0, n, nn, 0.0, R6 + (0 + sp) * ks, n, nn, S));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, h, 1.0, V + r0, n, nn, V + r0, n, nn, 0.0, R6 + (2 + sp) * ks, n, nn, S));
      TG_TRY(gemm_dd(s, OP_T, OP_N, n, n, h, 1.0, V + r0, n, nn, U + r0, n, nn, 0.0, R6 + (4 + sp) * ks, n, nn, S));
    }
    if (!defer) TG_TRY(d_products(s));
  }
  const double* Pp = R6 + 6 * ks;
  const double* Pn = R6 + 7 * ks;
  // the overlaps depend only on R6 and the set-up: fork them to the second stream
  hipStream_t s2 = c->stream2;
  TG_HIP(hipEventRecord(c->ev_fork, s));
  TG_HIP(hipStreamWaitEvent(s2, c->ev_fork, 0));
  // the canonical basis's device work goes first on the second stream, so that
  // its single-workgroup eigen-solver runs beside the inverse
  if (defer) TG_TRY(c->defer_hook(0));
  const RunPlan* PP = &P;   // the plan of the overlap chain (the new one when deferred)
  RunPlan P2;

  auto Wb = [&](int k) { return P.work + (size_t)k * chunk * nn; };
  cplx* Akeep = P.Akeep;                     // A_phi of every item
  cplx* Hkeep = P.Hkeep;                     // H20_phi of every item
  const long long os = P.os;

  // overlaps <Phi|R|Phi> for the gauge rotation (rot_over) and, when printing,
  // for parity x gauge (rot_pari, :279-284).  R_n = D0^T R D0 = ep*Pa + en*Pb.
  // Items are processed in segments of one overlap shape: the whole chunk when
  // every state has the same nocc / nemp, else one segment per state.
  auto overlap_segment = [&](const double* Pa, const double* Pb, long long a0, long long b0, int cnt,
                             const OvlShape& o, cplx* dst) -> int {
    const RunPlan& P = *PP;
    const int lo = (int)(b0 - a0);         // chunk-local offset of the segment
    if (o.reduced) {
      ScopedTiming tm(c, "overlap", s2);
      const int nb = o.nbloc;
      const long long mm = (long long)nb * nb;
      cplx* K = c->d_pfw.as<cplx>();
      cplx* PW = K + (size_t)cnt * mm;
      // the segment's states and their angle-independent parts of K
      const int st0 = (int)(b0 / A), nst = (int)((b0 + cnt - 1) / A) - st0 + 1;
      double* Qp = reinterpret_cast<double*>(PW + (size_t)cnt * mm);
      double* Qn = Qp + (size_t)nst * mm;
      double* JQp = Qn + (size_t)nst * mm;
      double* JQn = JQp + (size_t)nst * mm;
      double* Kpp = JQn + (size_t)nst * mm;
      double* Knn = Kpp + (size_t)nst * mm;
      double* Xpn = Knn + (size_t)nst * mm;
      {
        ScopedTiming tz(c, "kernel:k_parts", s2);
        hipLaunchKernelGGL(kq_prep_kernel, dim3(nblk(mm * nst)), dim3(NT), 0, s2, n, o.nemp, nb, nst,
                           st0, Pa, Pb, P.d_T, P.d_sg, os, Qp, Qn, JQp, JQn);
        TG_HIP(hipGetLastError());
        TG_TRY(gemm_dd(s2, OP_T, OP_N, nb, nb, nb, 1.0, Qp, nb, mm, JQp, nb, mm, 0.0, Kpp, nb, mm, nst));
        TG_TRY(gemm_dd(s2, OP_T, OP_N, nb, nb, nb, 1.0, Qn, nb, mm, JQn, nb, mm, 0.0, Knn, nb, mm, nst));
        TG_TRY(gemm_dd(s2, OP_T, OP_N, nb, nb, nb, 1.0, Qp, nb, mm, JQn, nb, mm, 0.0, Xpn, nb, mm, nst));
        double* Kpn = JQp;   // free after the products
        hipLaunchKernelGGL(kq_skew_kernel, dim3(nblk(mm * nst)), dim3(NT), 0, s2, nb, nst, Xpn, Kpn);
        hipLaunchKernelGGL(kq_combine_kernel, dim3(nblk(mm * cnt)), dim3(NT), 0, s2, nb, cnt, b0, A,
                           st0, P.d_ph, Kpp, Knn, Kpn, P.d_sg, os, K);
        TG_HIP(hipGetLastError());
      }
      {
        ScopedTiming tk(c, "kernel:pfaffian", s2);
        TG_TRY(batched_pfaffian(s2, nb, K, nb, mm, cnt, PW, P.d_logabs + lo, P.d_phase + lo,
                                P.d_zero + lo));
      }
      hipLaunchKernelGGL(fill_from_kernel, dim3(nblk(cnt)), dim3(NT), 0, s2, cnt, b0, A, P.d_pre0,
                         os / 2, P.d_pre + lo);
      hipLaunchKernelGGL(overlap_finish_kernel, dim3(nblk(cnt)), dim3(NT), 0, s2, cnt, b0, A,
                         P.d_logabs + lo, P.d_phase + lo, P.d_zero + lo, P.d_pre + lo, P.d_ovac,
                         os, dst + lo);
      TG_HIP(hipGetLastError());
      return TG_OK;
    }
    // reference sequence: R2 -> inverse + det -> M_reg -> Pfaffian(nreg)
    ScopedTiming tm(c, "overlap", s2);
    const int nreg = 2 * o.nbloc, n2 = o.n2;
    const long long m2 = (long long)n2 * n2, mr = (long long)nreg * nreg;
    cplx* R2 = c->d_pfw.as<cplx>();
    cplx* M = R2 + (size_t)cnt * m2;
    cplx* PW = M + (size_t)cnt * mr;
    cplx* det2 = PW + (size_t)cnt * mr;
    int* info2 = reinterpret_cast<int*>(det2 + cnt);
    int* perm2 = info2 + cnt;
    hipLaunchKernelGGL(build_rn2_kernel, dim3(nblk(m2 * cnt)), dim3(NT), 0, s2, n, o.nemp, n2, cnt,
                       b0, A, Pa, Pb, P.d_ph, R2);
    TG_HIP(hipGetLastError());
    TG_TRY(batched_inverse(s2, n2, R2, n2, m2, cnt, det2, info2, perm2));
    if (o.nbloc > 0) {
      hipLaunchKernelGGL(build_mreg_kernel, dim3(nblk(mr * cnt)), dim3(NT), 0, s2, n2, o.nbloc,
                         o.nbloc, nreg, cnt, b0, A, P.d_vou, P.d_vou, os, R2, M);
      TG_HIP(hipGetLastError());
      TG_TRY(batched_pfaffian(s2, nreg, M, nreg, mr, cnt, PW, P.d_logabs + lo, P.d_phase + lo,
                              P.d_zero + lo));
    } else {
      hipLaunchKernelGGL(unit_pfaffian_kernel, dim3(nblk(cnt)), dim3(NT), 0, s2, cnt,
                         P.d_logabs + lo, P.d_phase + lo, P.d_zero + lo);
      TG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(scale_det_kernel, dim3(nblk(cnt)), dim3(NT), 0, s2, cnt,
                       (double)o.signfac, det2, P.d_pre + lo);
    hipLaunchKernelGGL(overlap_finish_kernel, dim3(nblk(cnt)), dim3(NT), 0, s2, cnt, b0, A,
                       P.d_logabs + lo, P.d_phase + lo, P.d_zero + lo, P.d_pre + lo, P.d_ovac, os,
                       dst + lo);
    TG_HIP(hipGetLastError());
    return TG_OK;
  };
  auto overlap = [&](const double* Pa, const double* Pb, long long a0, int na, cplx* dst) -> int {
    const RunPlan& P = *PP;
    if (P.uniform) return overlap_segment(Pa, Pb, a0, a0, na, P.sh[0], dst);
    for (int st = (int)(a0 / A); st <= (int)((a0 + na - 1) / A); ++st) {
      const long long b0 = std::max<long long>((long long)st * A, a0);
      const long long b1 = std::min<long long>((long long)(st + 1) * A, a0 + na);
      TG_TRY(overlap_segment(Pa, Pb, a0, b0, (int)(b1 - b0), P.sh[st], dst));
    }
    return TG_OK;
  };
  // ---- loop over item chunks (one chunk of all items unless plain)
  for (int a0 = 0; a0 < B; a0 += chunk) {
    const int na = std::min(chunk, B - a0);
    const long long tot = nn * na;
    cplx* Uphi = Wb(W_UPHI);
    cplx* Vphi = Wb(W_VPHI);
    cplx* Aphi = Akeep + (size_t)a0 * nn;
    cplx* Ut = Wb(W_UT);
    cplx* Vt = Wb(W_VT);
    cplx* rho = Wb(W_RHO);
    cplx* kLR = Wb(W_KLR);
    cplx* kRL = Wb(W_KRL);
    cplx* G = Wb(W_G);
    cplx* hsp = Wb(W_H);
    cplx* dLR = Wb(W_DLR);
    cplx* dRL = Wb(W_DRL);
    cplx* X = Wb(W_X);
    cplx* Y = Wb(W_Y);
    cplx* Z = Wb(W_Z);
    cplx* H20 = Hkeep + (size_t)a0 * nn;

    if (!proj) {
      hipLaunchKernelGGL(fill_cplx_kernel, dim3(nblk(na)), dim3(NT), 0, s, na, cmk(1.0, 0.0),
                         P.d_over + a0);
      TG_HIP(hipGetLastError());
    } else if (!defer) {
      TG_TRY(overlap(Pp, Pn, a0, na, P.d_over + a0));
    }
    if (c->iprint && !defer) TG_TRY(overlap(R6 + 8 * ks, R6 + 9 * ks, a0, na, P.d_pari + a0));

    // Thouless matrix A_phi (calculate_thouless :684-718)
    {
      ScopedTiming tm(c, "thouless_inverse");
      hipLaunchKernelGGL(build_uv_phi_kernel, dim3(nblk(nn), (na + UV_AG - 1) / UV_AG), dim3(NT), 0,
                         s, n, na, (long long)a0,
                         A, S, R6, P.d_ph, Uphi, Vphi, nn);
      TG_HIP(hipGetLastError());
      ScopedTiming tk(c, "kernel:gj_inverse");
      TG_TRY(batched_inverse(s, n, Uphi, n, nn, na, P.d_det, P.d_info, P.d_perm));
    }
    // the real U, V of an item's state: batch element b of a launch over the
    // chunk reads Uc + (b / A) * nn, Uc = U of the chunk's first state (a chunk
    // holds whole states, or S = 1)
    {
      ScopedTiming tm(c, "gemm");
      const double* Uc = U + (size_t)(a0 / A) * nn;
      const double* Vc = V + (size_t)(a0 / A) * nn;
      {
        ScopedTiming tz(c, "kernel:zgemm");
        TG_TRY(gemm_zz(s, OP_N, OP_N, n, n, n, one, Vphi, n, nn, Uphi, n, nn, zero, Aphi, n, nn, na));
      }
      hipLaunchKernelGGL(thouless_prep_kernel, dim3(nblk(tot)), dim3(NT), 0, s, nn, na,
                         (long long)a0, A, U, V, Aphi, Ut, Vt);
      TG_HIP(hipGetLastError());
      TG_TRY(gemm_dz(s, OP_N, OP_N, n, n, n, one, Vc, n, nn, Aphi, n, nn, one, Ut, n, nn, na, A, 1));
      TG_TRY(gemm_dz(s, OP_N, OP_N, n, n, n, one, Uc, n, nn, Aphi, n, nn, one, Vt, n, nn, na, A, 1));
      // densities (calculate_densities with conj(U~), conj(V~))
      TG_TRY(gemm_zd(s, OP_N, OP_T, n, n, n, one, Vt, n, nn, Vc, n, nn, zero, rho, n, nn, na, 1, A));
      TG_TRY(gemm_zd(s, OP_N, OP_T, n, n, n, one, Vt, n, nn, Uc, n, nn, zero, kLR, n, nn, na, 1, A));
      TG_TRY(gemm_dz(s, OP_N, OP_T, n, n, n, one, Vc, n, nn, Ut, n, nn, zero, kRL, n, nn, na, A, 1));
    }
    // fields; in the projection the pp channel's scatters (third stream) are joined
    // after the first H20 product, which needs only h
    const bool split = proj;
    {
      ScopedTiming tm(c, "contraction");
      TG_TRY(contraction_run(c, na, rho, kLR, proj ? kRL : nullptr, G, dLR, dRL, proj ? 0 : 1,
                             c->d_H1.as<cplx>(), hsp, split));
    }
    // H20 (calculate_H20 with U~, V~) as G - G^T,
    //   X1 = h V~ + fd/2 DeltaLR U~,  X2 = -fd/2 DeltaRL V~,  G = U~^T X1 + V~^T X2
    const cplx hfd = cmk(0.5 * c->factor_delta, 0.0), mhfd = cmk(-0.5 * c->factor_delta, 0.0);
    auto zg = [&](TgOp oa, const cplx* Aa, cplx al, const cplx* Bb, cplx be, cplx* Cc) -> int {
      ScopedTiming tz(c, "kernel:zgemm");
      return gemm_zz(s, oa, OP_N, n, n, n, al, Aa, n, nn, Bb, n, nn, be, Cc, n, nn, na);
    };
    if (split) {
      {
        ScopedTiming tm(c, "gemm_h20");
        TG_TRY(zg(OP_N, hsp, one, Vt, zero, X));
      }
      TG_TRY(contraction_join(c));
    }
    // particle numbers and energy
    {
      const int bands = (n + TRT - 1) / TRT;
      hipLaunchKernelGGL(traces_kernel, dim3(bands, na), dim3(NT), 0, s, n, c->d_H1.as<cplx>(), rho,
                         kLR, kRL, G, dLR, nn, P.d_trp);
      hipLaunchKernelGGL(traces_sum_kernel, dim3(nblk(9LL * na)), dim3(NT), 0, s, na, bands, P.d_trp,
                         P.d_tr + 9 * (size_t)a0);
    }
    TG_HIP(hipGetLastError());
    {
      ScopedTiming tm(c, "gemm_h20");
      if (!split) TG_TRY(zg(OP_N, hsp, one, Vt, zero, X));
      TG_TRY(zg(OP_N, dLR, hfd, Ut, one, X));
      TG_TRY(zg(OP_N, dRL, mhfd, Vt, zero, Y));
      TG_TRY(zg(OP_T, Ut, one, X, zero, Z));
      TG_TRY(zg(OP_T, Vt, one, Y, one, Z));
      hipLaunchKernelGGL(skew_part_kernel, dim3((n + 31) / 32, (n + 31) / 32, na), dim3(256), 0, s,
                         n, Z, H20);
      TG_HIP(hipGetLastError());
    }
  }

  if (defer) {
    // the canonical basis (second stream, host waits on that stream only), the
    // overlap set-up of the new basis, then the overlap chain with its shapes
    TG_TRY(c->defer_hook(1));
    P2 = make_plan(c);
    PP = &P2;
    TG_TRY(d_products(s2));
    if (proj) TG_TRY(overlap(Pp, Pn, 0, B, P.d_over));
    if (c->iprint) TG_TRY(overlap(R6 + 8 * ks, R6 + 9 * ks, 0, B, P.d_pari));
  }

  // join the overlaps of the second stream
  TG_HIP(hipEventRecord(c->ev_join, s2));
  TG_HIP(hipStreamWaitEvent(s, c->ev_join, 0));

  // ---- sum_gauge and calculate_projected_H20, per state
  cplx* pA = c->d_out.as<cplx>();
  cplx* pH = pA + nn;
  cplx* pAH = pH + nn;
  double* H20v = reinterpret_cast<double*>(pAH + nn);
  hipLaunchKernelGGL(gauge_scalars_kernel, dim3(S), dim3(NT), 0, s, A, n, c->H0, P.d_w, P.d_over,
                     P.d_tr, c->iprint ? P.d_pari : nullptr, P.d_mult, P.d_f12, P.d_ener, P.d_pnp);
  hipLaunchKernelGGL(gauge_matrices_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, A, P.d_f12,
                     P.d_mult, Akeep, Hkeep, nn, pA, pH, pAH, P.outs);
  hipLaunchKernelGGL(projected_h20_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, n, P.d_pnp, pA, pH,
                     pAH, H20v, P.d_imag, P.outs, P.is);
  TG_HIP(hipGetLastError());
  if (c->grad_on) {
    // Z buffers: [Zi | Zim1 | lambda | partial sums], S states each; the new Z
    // becomes the momentum term of the next step (gradient_Zim1 = gradient_Zi)
    double* Zb = c->d_Z.as<double>();
    double* Zcur = Zb;
    double* Zprev = Zb + ks;
    double* lam = Zb + 2 * ks;
    double* part = lam + (size_t)S * TG_MAX_CONS;
    const long long ls = c->c20_shared ? 0 : TG_MAX_CONS;
    const long long cs = c->c20_shared ? 0 : (long long)c->grad_ncons * nn;
    hipLaunchKernelGGL(gradient_kernel, dim3(nblk(nn), S), dim3(NT), 0, s, nn, c->grad_eta,
                       c->grad_mu, (const double*)nullptr, c->grad_ncons, lam, ls,
                       c->d_c20.as<double>(), cs, H20v,
                       2 * P.outs, Zprev, Zcur, part, (long long)tg_part_len(n));
    TG_HIP(hipGetLastError());
    hipLaunchKernelGGL(copy_real_kernel, dim3(nblk(ks)), dim3(NT), 0, s, ks, Zcur, Zprev);
    TG_HIP(hipGetLastError());
  }
  return TG_OK;
}

// Graph key: everything project_run passes by value or sizes by, plus the
// allocation epoch (a reallocated buffer invalidates captured pointers).
GraphKey graph_key(const tg_ctx* c) {
  GraphKey k{};
  double fd = c->factor_delta, eta = c->grad_eta, mu = c->grad_mu;
  long long b1, b2, b3;
  std::memcpy(&b1, &fd, 8);
  std::memcpy(&b2, &eta, 8);
  std::memcpy(&b3, &mu, 8);
  unsigned long long shapes = 1469598103934665603ULL;   // FNV-1a over (nocc, nemp) of the states
  for (const StateInfo& si : c->st)
    for (int x : {si.nocc, si.nemp}) {
      shapes ^= (unsigned long long)(unsigned)x;
      shapes *= 1099511628211ULL;
    }
  const long long vals[20] = {c->n, c->A, c->Mphip, c->Mphin, c->plain, c->ns, (long long)shapes,
                              c->iprint, c->grad_on, c->grad_ncons, b1, b2, b3,
                              (long long)devbuf_epoch(), c->w_fp32, c->c20_shared, 1,
                              chunk_items(c, c->ns * c->A, c->A), 0, 0};
  for (int i = 0; i < 20; ++i) k.v[i] = vals[i];
  return k;
}

void graph_release(tg_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = 0;
}

// One projection of the resident state.  With TG_FLAG_GRAPH the kernel
// sequence is captured into a hipGraph the second time a configuration is seen
// (the first, eager run allocates and autotunes) and replayed afterwards; a new
// state of the same configuration only re-uploads its set-up.
int project_device(tg_ctx* c) {
  if (c->state_dirty) TG_TRY(project_upload(c));
  if (!c->use_graph || c->timing) return project_run(c);
  const GraphKey k = graph_key(c);
  if (c->graph_valid && k == c->graph_key_v) {
    TG_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    return TG_OK;
  }
  if (!(c->graph_seen && k == c->graph_seen_key)) {
    c->graph_seen = 1;
    c->graph_seen_key = k;
    return project_run(c);
  }
  graph_release(c);
  TG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int r = project_run(c);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (r != TG_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (r != TG_OK) return r;
    tg_set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return TG_E_HIP;
  }
  c->graph = g;
  if (hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
    c->graph_exec = nullptr;
    graph_release(c);
    tg_set_error("hipGraphInstantiate failed");
    return TG_E_HIP;
  }
  c->graph_valid = 1;
  c->graph_key_v = k;
  if (std::getenv("TG_GRAPH_VERBOSE")) {
    size_t nodes = 0;
    (void)hipGraphGetNodes(g, nullptr, &nodes);
    fprintf(stderr, "[taurus_gpu] projection captured: %zu graph nodes\n", nodes);
  }
  TG_HIP(hipGraphLaunch(c->graph_exec, c->stream));
  return TG_OK;
}

// dst(i, j) = src(perm i, perm j): a single-particle matrix into the internal order
void permute_in(const tg_ctx* c, const cplx* src, cplx* dst) {
  const int n = c->n;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      dst[i + (size_t)j * n] = src[c->sp_perm[i] + (size_t)c->sp_perm[j] * n];
}
// dst(perm i, perm j) = src(i, j): back to the reference order
void permute_out(const tg_ctx* c, const cplx* src, cplx* dst) {
  const int n = c->n;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      dst[c->sp_perm[i] + (size_t)c->sp_perm[j] * n] = src[i + (size_t)j * n];
}

void assign_angles(tg_ctx* c, bool conj);

// S states of this basis (state k: U0 + k n^2, V0 + k n^2, D0 + k n^2,
// voveru0 + k n holding n - nocc0[k] values) and the common projection set-up.
int upload_states(tg_ctx* c, int S, const double* U0, const double* V0, const double* D0,
                  const int* nocc0, const int* nemp0, const double* ovac0, const double* voveru0,
                  int Mphip, int Mphin, double facpi, double valZ, double valN, double factor_delta,
                  int iprint) {
  const int n = c->n;
  bool ok = S >= 1 && U0 && V0 && D0 && nocc0 && nemp0 && ovac0 && Mphip >= 1 && Mphin >= 1;
  for (int k = 0; ok && k < S; ++k)
    ok = nocc0[k] >= 0 && nemp0[k] >= 0 && nocc0[k] + nemp0[k] <= n &&
         (n - nocc0[k] == 0 || voveru0);
  if (!ok) {
    tg_set_error("tg_set_states: invalid argument");
    return TG_E_ARG;
  }
  if (c->plain && S > 1) {
    tg_set_error("tg_set_states: the plain per-angle mode holds one state");
    return TG_E_ARG;
  }
  const size_t nn = (size_t)n * n;
  TG_TRY(c->d_U.ensure(sizeof(double) * nn * S));
  TG_TRY(c->d_V.ensure(sizeof(double) * nn * S));
  TG_TRY(c->d_D.ensure(sizeof(double) * nn * S));
  TG_TRY(c->d_real6.ensure(sizeof(double) * 11 * nn * S));
  // rows (single-particle index) into the internal order
  std::vector<double> h((size_t)3 * nn * S);
  double* hU = h.data();
  double* hV = hU + nn * S;
  double* hD = hV + nn * S;
  for (int k = 0; k < S; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t src = k * nn + (size_t)c->sp_perm[i] + (size_t)j * n;
        const size_t dst = k * nn + (size_t)i + (size_t)j * n;
        hU[dst] = U0[src];
        hV[dst] = V0[src];
        hD[dst] = D0[src];
      }
  TG_HIP(hipMemcpyAsync(c->d_U.p, hU, sizeof(double) * nn * S, hipMemcpyHostToDevice, c->stream));
  TG_HIP(hipMemcpyAsync(c->d_V.p, hV, sizeof(double) * nn * S, hipMemcpyHostToDevice, c->stream));
  TG_HIP(hipMemcpyAsync(c->d_D.p, hD, sizeof(double) * nn * S, hipMemcpyHostToDevice, c->stream));
  TG_HIP(hipStreamSynchronize(c->stream));
  c->st.assign(S, StateInfo());
  for (int k = 0; k < S; ++k) {
    StateInfo& si = c->st[k];
    si.nocc = nocc0[k];
    si.nemp = nemp0[k];
    si.ovac = ovac0[k];
    if (n - nocc0[k] > 0) si.voveru.assign(voveru0 + (size_t)k * n, voveru0 + (size_t)k * n + (n - nocc0[k]));
  }
  c->ns = S;
  c->Mphip = Mphip;
  c->Mphin = Mphin;
  c->facpi = facpi;
  c->valZ = valZ;
  c->valN = valN;
  c->factor_delta = factor_delta;
  c->iprint = iprint ? 1 : 0;
  return tg_states_ready(c);
}

// assign_angles (:112-185) for the context's Mphip, Mphin, facpi and valence
// numbers.  With `conj` (TG_FLAG_CONJ_PAIRS) and the full [0, 2 pi) grid, the
// grid is closed under phi -> -phi (mod 2 pi): for real U, V every per-angle
// quantity at -phi is the complex conjugate of that at phi, and so is the
// weight, so only one angle of each pair is computed and enters the gauge sums
// twice, as 2 Re(.); self-conjugate angles enter once.
void assign_angles(tg_ctx* c, bool conj) {
  const int Mphip = c->Mphip, Mphin = c->Mphin;
  const double facpi = c->facpi, valZ = c->valZ, valN = c->valN;
  const bool pairs = conj && Mphip * Mphin > 1 &&
                     fabs(facpi - 2.0 * M_PI) <= 1e-12 * 2.0 * M_PI;
  auto partner = [](int j, int M) { return (M % 2 == 1) ? (j == 1 ? 1 : M - j + 2) : M - j + 1; };
  c->phip.clear();
  c->phin.clear();
  c->weight.clear();
  c->mult.clear();
  for (int jp = 1; jp <= Mphip; ++jp)
    for (int jn = 1; jn <= Mphin; ++jn) {
      double mult = 1.0;
      if (pairs) {
        const int i0 = (jp - 1) * Mphin + (jn - 1);
        const int i1 = (partner(jp, Mphip) - 1) * Mphin + (partner(jn, Mphin) - 1);
        if (i1 < i0) continue;        // computed as the partner of i1
        if (i1 > i0) mult = 2.0;
      }
      const double pp = (Mphip % 2 == 1) ? facpi * (jp - 1.0) / Mphip : facpi * (jp - 0.5) / Mphip;
      const double pn = (Mphin % 2 == 1) ? facpi * (jn - 1.0) / Mphin : facpi * (jn - 0.5) / Mphin;
      c->phip.push_back(pp);
      c->phin.push_back(pn);
      const cplx wp = cmk(cos(-pp * valZ) / Mphip, sin(-pp * valZ) / Mphip);
      const cplx wn = cmk(cos(-pn * valN) / Mphin, sin(-pn * valN) / Mphin);
      c->weight.push_back(cmul(wp, wn));
      c->mult.push_back(mult);
    }
  c->A = (int)c->phip.size();
}

int fetch_results(tg_ctx* c, int state, double* scal, double* pA, double* pH, double* pAH,
                  double* H20v, double* imag) {
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  if (state < 0 || state >= c->ns) {
    tg_set_error("results: state index out of range");
    return TG_E_ARG;
  }
  TG_HIP(hipStreamSynchronize(c->stream));
  const RunPlan P = make_plan(c);
  const cplx* d_out = c->d_out.as<cplx>() + (size_t)state * P.outs;
  const cplx* d_pnp = P.d_pnp + 8 * (size_t)state;
  const double* d_imag = P.d_imag + (size_t)state * P.is;
  if (scal) TG_HIP(hipMemcpy(scal, d_pnp, sizeof(cplx) * 7, hipMemcpyDeviceToHost));
  if (pA) TG_HIP(hipMemcpy(pA, d_out, sizeof(cplx) * nn, hipMemcpyDeviceToHost));
  if (pH) TG_HIP(hipMemcpy(pH, d_out + nn, sizeof(cplx) * nn, hipMemcpyDeviceToHost));
  if (pAH) TG_HIP(hipMemcpy(pAH, d_out + 2 * nn, sizeof(cplx) * nn, hipMemcpyDeviceToHost));
  if (H20v)
    TG_HIP(hipMemcpy(H20v, reinterpret_cast<const double*>(d_out + 3 * nn), sizeof(double) * nn,
                     hipMemcpyDeviceToHost));
  const int nb = (int)((nn + NT - 1) / NT);
  std::vector<double> im(nb);
  TG_HIP(hipMemcpy(im.data(), d_imag, sizeof(double) * nb, hipMemcpyDeviceToHost));
  double mx = -INFINITY;
  for (double v : im) mx = std::max(mx, v);
  if (imag) *imag = mx;
  if (mx > 1.0e-8) {
    tg_set_error("The gradient has non-zero imaginary part: " + std::to_string(mx));
    return TG_E_IMAG_GRADIENT;
  }
  return TG_OK;
}

int get_device_ready() {
  int cnt = 0;
  hipError_t e = hipGetDeviceCount(&cnt);
  if (e != hipSuccess || cnt < 1) {
    tg_set_error("no HIP device available");
    return TG_E_NODEVICE;
  }
  return TG_OK;
}

}  // namespace

// entry points for the other translation units of the library
int tg_project_device(tg_ctx* c) { return project_device(c); }
int tg_project_upload(tg_ctx* c) { return c->state_dirty ? project_upload(c) : TG_OK; }
int tg_project_run(tg_ctx* c) { return project_run(c); }
int tg_states_ready(tg_ctx* c) {
  if (c->grad_on) TG_TRY(tg_ensure_z(c));
  assign_angles(c, c->conj_pairs != 0);
  c->have_state = 1;
  c->evolved = 0;
  c->state_dirty = 1;
  return TG_OK;
}
int tg_ensure_z(tg_ctx* c) {
  const long long nn = (long long)c->n * c->n;
  const int S = c->ns;
  const size_t need = sizeof(double) * ((size_t)2 * S * nn + (size_t)S * TG_MAX_CONS +
                                        (size_t)S * tg_part_len(c->n));
  if (c->z_states == S && c->d_Z.bytes >= need) return TG_OK;
  TG_TRY(c->d_Z.ensure(need));
  TG_HIP(hipMemsetAsync(c->d_Z.p, 0, need, c->stream));
  // multipliers common to every state (tg_set_gradient) in the slot of state 0
  if (c->c20_shared && !c->grad_lambda.empty())
    TG_HIP(hipMemcpyAsync(c->d_Z.as<double>() + 2 * S * nn, c->grad_lambda.data(),
                          sizeof(double) * c->grad_lambda.size(), hipMemcpyHostToDevice, c->stream));
  TG_HIP(hipStreamSynchronize(c->stream));
  c->z_states = S;
  return TG_OK;
}
int tg_gradient_step(tg_ctx* c, const double* c20, const double* lam) {
  const int n = c->n, S = c->ns;
  const long long nn = (long long)n * n;
  double* Zb = c->d_Z.as<double>();
  double* part = Zb + 2 * S * nn + (size_t)S * TG_MAX_CONS;
  const double* H20v = reinterpret_cast<const double*>(c->d_out.as<cplx>() + 3 * nn);
  // per-state eta, mu of the gradient recipes (gradient_type 1, 2), else the given ones
  const double* em = nullptr;
  if (c->grad_type > 0 && c->etamu.size() == 2 * (size_t)S) {
    TG_TRY(c->d_etamu.ensure(sizeof(double) * 2 * S));
    TG_HIP(hipMemcpyAsync(c->d_etamu.p, c->etamu.data(), sizeof(double) * 2 * S,
                          hipMemcpyHostToDevice, c->stream));
    em = c->d_etamu.as<double>();
  }
  hipLaunchKernelGGL(gradient_kernel, dim3(nblk(nn), S), dim3(NT), 0, c->stream, nn, c->grad_eta,
                     c->grad_mu, em, c->grad_ncons, lam, (long long)TG_MAX_CONS, c20,
                     (long long)c->grad_ncons * nn, H20v, 2 * tg_out_stride(n), Zb + S * nn, Zb,
                     part, tg_part_len(n));
  TG_HIP(hipGetLastError());
  return TG_OK;
}
int tg_fetch_results(tg_ctx* c, int state, double* scal, double* pA, double* pH, double* pAH,
                     double* H20v, double* imag) {
  return fetch_results(c, state, scal, pA, pH, pAH, H20v, imag);
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int tg_init(tg_ctx** out, int ndim, const int32_t* sp_2mj, const int32_t* sp_2mt,
            const int32_t* sp_l, const int32_t* sp_tr, double H0_re, double H0_im,
            const double* H1, int64_t ntb, const int16_t* abcd, const void* H2, int h2_bytes,
            const int8_t* trperm, int flags) {
  if (!out || ndim <= 0 || ndim % 2 != 0 || !sp_2mj || !sp_2mt || !sp_l || !sp_tr || !H1 ||
      ntb < 0 || (ntb > 0 && (!abcd || !H2 || !trperm)) || (h2_bytes != 4 && h2_bytes != 8)) {
    tg_set_error("tg_init: invalid argument");
    return TG_E_ARG;
  }
  TG_TRY(get_device_ready());
  tg_ctx* c = new tg_ctx();
  c->n = ndim;
  c->flags = flags;
  c->plain = (flags & TG_FLAG_PLAIN) ? 1 : 0;
  c->w_fp32 = (flags & TG_FLAG_TBME_FP32) ? 1 : 0;
  c->use_graph = (flags & TG_FLAG_GRAPH) ? 1 : 0;
  c->conj_pairs = (flags & TG_FLAG_CONJ_PAIRS) ? 1 : 0;
  // Internal single-particle order: by (2mt, 2mj, l mod 2, reference index).
  // Every class of the contraction blocks then occupies contiguous runs of
  // rows in each column of the densities and fields, and the protons stay in
  // the first half.  U, V, D0, H1 and the TBME indices are permuted on the way
  // in and the single-particle fields on the way out; quasiparticle-basis
  // quantities (A, H20, gradient) are unaffected.
  for (int i = 0; i < ndim; ++i)
    if (sp_tr[i] < 1 || sp_tr[i] > ndim) {
      delete c;
      tg_set_error("tg_init: sp_tr out of range");
      return TG_E_ARG;
    }
  c->sp_perm.resize(ndim);
  for (int i = 0; i < ndim; ++i) c->sp_perm[i] = i;
  std::stable_sort(c->sp_perm.begin(), c->sp_perm.end(), [&](int x, int y) {
    if (sp_2mt[x] != sp_2mt[y]) return sp_2mt[x] < sp_2mt[y];
    if (sp_2mj[x] != sp_2mj[y]) return sp_2mj[x] < sp_2mj[y];
    return (sp_l[x] & 1) < (sp_l[y] & 1);
  });
  c->sp_iperm.resize(ndim);
  for (int i = 0; i < ndim; ++i) c->sp_iperm[c->sp_perm[i]] = i;
  int nprot = 0;
  for (int i = 0; i < ndim; ++i) nprot += sp_2mt[i] < 0;
  if (2 * nprot != ndim) {
    delete c;
    tg_set_error("tg_init: the basis must hold ndim/2 proton states");
    return TG_E_ARG;
  }
  c->sp_2mj.resize(ndim);
  c->sp_2mt.resize(ndim);
  c->sp_l.resize(ndim);
  c->sp_tr.resize(ndim);
  for (int i = 0; i < ndim; ++i) {
    const int o = c->sp_perm[i];
    c->sp_2mj[i] = sp_2mj[o];
    c->sp_2mt[i] = sp_2mt[o];
    c->sp_l[i] = sp_l[o];
    c->sp_tr[i] = c->sp_iperm[sp_tr[o] - 1];
  }
  c->H0 = cmk(H0_re, H0_im);
  int r = TG_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork3, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join3, hipEventDisableTiming) != hipSuccess) {
    tg_finalize(c);
    tg_set_error("tg_init: stream / event creation failed");
    return TG_E_HIP;
  }
  r = c->d_H1.ensure(sizeof(cplx) * (size_t)ndim * ndim);
  if (r == TG_OK) {
    std::vector<cplx> h1((size_t)ndim * ndim);
    permute_in(c, reinterpret_cast<const cplx*>(H1), h1.data());
    if (hipMemcpy(c->d_H1.p, h1.data(), sizeof(cplx) * h1.size(), hipMemcpyHostToDevice) !=
        hipSuccess) r = TG_E_HIP;
  }
  if (r == TG_OK) {
    std::vector<double> ps(ndim);
    for (int i = 0; i < ndim; ++i) ps[i] = (c->sp_l[i] % 2 == 0) ? 1.0 : -1.0;
    r = c->d_psign.ensure(sizeof(double) * ndim);
    if (r == TG_OK && hipMemcpy(c->d_psign.p, ps.data(), sizeof(double) * ndim,
                                hipMemcpyHostToDevice) != hipSuccess) r = TG_E_HIP;
  }
  if (r == TG_OK) r = contraction_build(c, ntb, abcd, H2, h2_bytes, trperm);
  if (r != TG_OK) {
    tg_finalize(c);
    return r;
  }
  *out = c;
  return TG_OK;
}

int tg_finalize(tg_ctx* c) {
  if (!c) return TG_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)timing_collect(c);
  graph_release(c);
  for (DevBuf* b : {&c->d_H1, &c->d_U, &c->d_V, &c->d_D, &c->d_ang, &c->d_real6, &c->d_work,
                    &c->d_blk, &c->d_scal, &c->d_out, &c->d_ovl, &c->d_int, &c->d_pfw, &c->d_c20,
                    &c->d_Z, &c->d_psign, &c->d_cHO, &c->d_it, &c->d_obs, &c->d_canon, &c->d_cq})
    b->release();
  for (Channel* ch : {&c->ph, &c->pp})
    for (DevBuf* b : {&ch->d_inv, &ch->d_groups, &ch->d_W, &ch->d_tile_group, &ch->d_tile_n0})
      b->release();
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->stream3) (void)hipStreamSynchronize(c->stream3);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_fork3) (void)hipEventDestroy(c->ev_fork3);
  if (c->ev_join3) (void)hipEventDestroy(c->ev_join3);
  if (c->stream3) (void)hipStreamDestroy(c->stream3);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TG_OK;
}

int tg_set_batch_limit(tg_ctx* c, int max_items) {
  if (!c || max_items < 0) {
    tg_set_error("tg_set_batch_limit: invalid argument");
    return TG_E_ARG;
  }
  c->batch_limit = max_items;
  c->state_dirty = 1;
  return TG_OK;
}

int tg_set_mode(tg_ctx* c, int plain) {
  if (!c) return TG_E_ARG;
  c->plain = plain ? 1 : 0;
  return TG_OK;
}

int tg_calculate_fields(tg_ctx* c, const double* rhoLR, const double* kappaLR,
                        const double* kappaRL, double* gammaLR, double* hspLR, double* deltaLR,
                        double* deltaRL) {
  if (!c || !rhoLR || !kappaLR || !kappaRL || !gammaLR || !hspLR || !deltaLR || !deltaRL) {
    tg_set_error("tg_calculate_fields: null argument");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  TG_TRY(c->d_work.ensure(sizeof(cplx) * 8 * nn));
  cplx* w = c->d_work.as<cplx>();
  std::vector<cplx> h(4 * nn);
  permute_in(c, reinterpret_cast<const cplx*>(rhoLR), h.data());
  permute_in(c, reinterpret_cast<const cplx*>(kappaLR), h.data() + nn);
  permute_in(c, reinterpret_cast<const cplx*>(kappaRL), h.data() + 2 * nn);
  TG_HIP(hipMemcpyAsync(w, h.data(), sizeof(cplx) * 3 * nn, hipMemcpyHostToDevice, c->stream));
  TG_TRY(contraction_run(c, 1, w, w + nn, w + 2 * nn, w + 3 * nn, w + 5 * nn, w + 6 * nn, 0));
  hipLaunchKernelGGL(add_h1_kernel, dim3(nblk(nn)), dim3(NT), 0, c->stream, (long long)nn, 1,
                     c->d_H1.as<cplx>(), w + 3 * nn, w + 4 * nn);
  TG_HIP(hipGetLastError());
  TG_HIP(hipStreamSynchronize(c->stream));
  TG_HIP(hipMemcpy(h.data(), w + 3 * nn, sizeof(cplx) * 4 * nn, hipMemcpyDeviceToHost));
  permute_out(c, h.data(), reinterpret_cast<cplx*>(gammaLR));
  permute_out(c, h.data() + nn, reinterpret_cast<cplx*>(hspLR));
  permute_out(c, h.data() + 2 * nn, reinterpret_cast<cplx*>(deltaLR));
  permute_out(c, h.data() + 3 * nn, reinterpret_cast<cplx*>(deltaRL));
  return TG_OK;
}

int tg_calculate_fields_diag(tg_ctx* c, const double* rhoLR, const double* kappaLR,
                             double* gammaLR, double* hspLR, double* deltaLR, double* deltaRL) {
  if (!c || !rhoLR || !kappaLR || !gammaLR || !hspLR || !deltaLR) {
    tg_set_error("tg_calculate_fields_diag: null argument");
    return TG_E_ARG;
  }
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  TG_TRY(c->d_work.ensure(sizeof(cplx) * 8 * nn));
  cplx* w = c->d_work.as<cplx>();
  std::vector<cplx> h(4 * nn);
  permute_in(c, reinterpret_cast<const cplx*>(rhoLR), h.data());
  permute_in(c, reinterpret_cast<const cplx*>(kappaLR), h.data() + nn);
  TG_HIP(hipMemcpyAsync(w, h.data(), sizeof(cplx) * 2 * nn, hipMemcpyHostToDevice, c->stream));
  TG_TRY(contraction_run(c, 1, w, w + nn, nullptr, w + 3 * nn, w + 5 * nn, w + 6 * nn, 1));
  hipLaunchKernelGGL(add_h1_kernel, dim3(nblk(nn)), dim3(NT), 0, c->stream, (long long)nn, 1,
                     c->d_H1.as<cplx>(), w + 3 * nn, w + 4 * nn);
  TG_HIP(hipGetLastError());
  TG_HIP(hipStreamSynchronize(c->stream));
  TG_HIP(hipMemcpy(h.data(), w + 3 * nn, sizeof(cplx) * 4 * nn, hipMemcpyDeviceToHost));
  permute_out(c, h.data(), reinterpret_cast<cplx*>(gammaLR));
  permute_out(c, h.data() + nn, reinterpret_cast<cplx*>(hspLR));
  permute_out(c, h.data() + 2 * nn, reinterpret_cast<cplx*>(deltaLR));
  if (deltaRL) permute_out(c, h.data() + 3 * nn, reinterpret_cast<cplx*>(deltaRL));
  return TG_OK;
}

int tg_inverse(int n, int count, double* A, int lda, double* det, int* info) {
  if (n < 0 || count < 0 || lda < std::max(1, n) || (count > 0 && (!A || !det || !info))) {
    tg_set_error("tg_inverse: invalid argument");
    return TG_E_ARG;
  }
  if (count == 0 || n == 0) return TG_OK;
  TG_TRY(get_device_ready());
  const size_t m = (size_t)lda * n;
  cplx *dA = nullptr, *dDet = nullptr;
  int *dInfo = nullptr, *dPerm = nullptr;
  int r = TG_OK;
  if (hipMalloc(&dA, sizeof(cplx) * m * count) != hipSuccess ||
      hipMalloc(&dDet, sizeof(cplx) * count) != hipSuccess ||
      hipMalloc(&dInfo, sizeof(int) * count) != hipSuccess ||
      hipMalloc(&dPerm, sizeof(int) * (size_t)n * count) != hipSuccess) {
    tg_set_error("tg_inverse: hipMalloc failed");
    r = TG_E_NOMEM;
  }
  if (r == TG_OK && hipMemcpy(dA, A, sizeof(cplx) * m * count, hipMemcpyHostToDevice) != hipSuccess)
    r = TG_E_HIP;
  if (r == TG_OK) r = batched_inverse(0, n, dA, lda, (long long)m, count, dDet, dInfo, dPerm);
  if (r == TG_OK && (hipDeviceSynchronize() != hipSuccess ||
                     hipMemcpy(A, dA, sizeof(cplx) * m * count, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(det, dDet, sizeof(cplx) * count, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(info, dInfo, sizeof(int) * count, hipMemcpyDeviceToHost) != hipSuccess))
    r = TG_E_HIP;
  if (dA) (void)hipFree(dA);
  if (dDet) (void)hipFree(dDet);
  if (dInfo) (void)hipFree(dInfo);
  if (dPerm) (void)hipFree(dPerm);
  return r;
}

int tg_zskpfa(int n, int count, const double* A, int lda, double* pfaff, int* info) {
  if (n < 0 || count < 0 || lda < std::max(1, n) || (count > 0 && (!A || !pfaff || !info))) {
    tg_set_error("tg_zskpfa: invalid argument");
    return TG_E_ARG;
  }
  if (count == 0) return TG_OK;
  if (n == 0 || n % 2 == 1) {
    for (int k = 0; k < count; ++k) {
      pfaff[2 * k] = (n == 0) ? 1.0 : 0.0;
      pfaff[2 * k + 1] = 0.0;
      info[k] = 0;
    }
    return TG_OK;
  }
  TG_TRY(get_device_ready());
  const size_t m = (size_t)lda * n;
  cplx *dA = nullptr, *dW = nullptr, *dPh = nullptr;
  double* dL = nullptr;
  int* dZ = nullptr;
  int r = TG_OK;
  if (hipMalloc(&dA, sizeof(cplx) * m * count) != hipSuccess ||
      hipMalloc(&dW, sizeof(cplx) * (size_t)n * n * count) != hipSuccess ||
      hipMalloc(&dPh, sizeof(cplx) * count) != hipSuccess ||
      hipMalloc(&dL, sizeof(double) * count) != hipSuccess ||
      hipMalloc(&dZ, sizeof(int) * count) != hipSuccess) {
    tg_set_error("tg_zskpfa: hipMalloc failed");
    r = TG_E_NOMEM;
  }
  if (r == TG_OK && hipMemcpy(dA, A, sizeof(cplx) * m * count, hipMemcpyHostToDevice) != hipSuccess)
    r = TG_E_HIP;
  if (r == TG_OK) r = batched_pfaffian(0, n, dA, lda, (long long)m, count, dW, dL, dPh, dZ);
  std::vector<double> L(count);
  std::vector<cplx> Ph(count);
  std::vector<int> Zr(count);
  if (r == TG_OK && (hipDeviceSynchronize() != hipSuccess ||
                     hipMemcpy(L.data(), dL, sizeof(double) * count, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(Ph.data(), dPh, sizeof(cplx) * count, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(Zr.data(), dZ, sizeof(int) * count, hipMemcpyDeviceToHost) != hipSuccess))
    r = TG_E_HIP;
  if (r == TG_OK)
    for (int k = 0; k < count; ++k) {
      const double mag = Zr[k] ? 0.0 : exp(L[k]);
      pfaff[2 * k] = mag * Ph[k].x;
      pfaff[2 * k + 1] = mag * Ph[k].y;
      info[k] = 0;
    }
  if (dA) (void)hipFree(dA);
  if (dW) (void)hipFree(dW);
  if (dPh) (void)hipFree(dPh);
  if (dL) (void)hipFree(dL);
  if (dZ) (void)hipFree(dZ);
  if (r == TG_E_HIP) tg_set_error("tg_zskpfa: HIP error");
  return r;
}

int tg_calculate_overlap(int ndim, int nocc0, int nemp0, double ovac0, const double* voveru0,
                         const double* D0, int nocc1, int nemp1, double ovac1,
                         const double* voveru1, const double* D1, const double* ROT,
                         double* overlap) {
  const int n = ndim;
  if (n <= 0 || !D0 || !D1 || !ROT || !overlap || nocc0 < 0 || nocc1 < 0 || nemp0 < 0 ||
      nemp1 < 0 || nocc0 + nemp0 > n || nocc1 + nemp1 > n) {
    tg_set_error("tg_calculate_overlap: invalid argument");
    return TG_E_ARG;
  }
  // dimensions and phase factor (:803-818)
  const int nempm = std::min(nemp0, nemp1);
  const int nbloc0 = n - nocc0 - nempm, nbloc1 = n - nocc1 - nempm;
  const int nreg = nbloc0 + nbloc1;
  const int n0 = n - nocc0, n1 = n - nocc1, n2 = n - nempm;
  const long long e1 = (long long)n2 * (n2 + 1) / 2;
  const long long e2 = (long long)nocc0 * (nocc0 - 1) / 2;
  const long long e3 = (long long)nocc0 * (nocc0 + 1) / 2 + (long long)nocc1 * (nocc1 + 1) / 2 +
                       (long long)nocc0 * (n0 + n2) + (long long)nocc1 * n1;
  const double signfac = sign_pow(e1) * sign_pow(e2) * sign_pow(e3);
  const double normvac = 0.5 * (ovac0 + ovac1);
  const int isla0 = (n0 - nemp0 == 0) ? 1 : 0, isla1 = (n1 - nemp1 == 0) ? 2 : 0;
  cplx ov = cmk(0, 0);
  switch (isla0 + isla1) {
    case 1: if (nocc0 > n - nemp1) { overlap[0] = overlap[1] = 0; return TG_OK; } break;
    case 2: if (nocc1 > n - nemp0) { overlap[0] = overlap[1] = 0; return TG_OK; } break;
    case 3:
      if (nocc0 == nocc1) ov = cmk(1, 0);
      else { overlap[0] = overlap[1] = 0; return TG_OK; }
      break;
    default: break;
  }
  if ((nbloc0 > 0 && !voveru0) || (nbloc1 > 0 && !voveru1)) {
    tg_set_error("tg_calculate_overlap: voveru missing");
    return TG_E_ARG;
  }
  TG_TRY(get_device_ready());
  const size_t nn = (size_t)n * n, m2 = (size_t)n2 * n2, mr = (size_t)nreg * nreg;
  cplx *dD0 = nullptr, *dD1 = nullptr, *dR = nullptr, *dT = nullptr, *dRn = nullptr, *dM = nullptr,
       *dPW = nullptr, *dDet = nullptr, *dPh = nullptr;
  double *dL = nullptr, *dv0 = nullptr, *dv1 = nullptr;
  int *dInfo = nullptr, *dPerm = nullptr, *dZ = nullptr;
  int r = TG_OK;
  auto al = [&](void** p, size_t b) {
    if (r == TG_OK && hipMalloc(p, std::max<size_t>(b, 16)) != hipSuccess) r = TG_E_NOMEM;
  };
  al((void**)&dD0, sizeof(cplx) * nn); al((void**)&dD1, sizeof(cplx) * nn);
  al((void**)&dR, sizeof(cplx) * nn); al((void**)&dT, sizeof(cplx) * nn);
  al((void**)&dRn, sizeof(cplx) * nn); al((void**)&dM, sizeof(cplx) * mr);
  al((void**)&dPW, sizeof(cplx) * mr); al((void**)&dDet, sizeof(cplx));
  al((void**)&dPh, sizeof(cplx)); al((void**)&dL, sizeof(double));
  al((void**)&dv0, sizeof(double) * (size_t)std::max(nbloc0, 1));
  al((void**)&dv1, sizeof(double) * (size_t)std::max(nbloc1, 1));
  al((void**)&dInfo, sizeof(int)); al((void**)&dPerm, sizeof(int) * (size_t)n2 + 16);
  al((void**)&dZ, sizeof(int));
  hipStream_t s = 0;
  const cplx one = cmk(1, 0), zero = cmk(0, 0);
  if (r == TG_OK) {
    if (hipMemcpy(dD0, D0, sizeof(cplx) * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dD1, D1, sizeof(cplx) * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dR, ROT, sizeof(cplx) * nn, hipMemcpyHostToDevice) != hipSuccess ||
        (nbloc0 > 0 && hipMemcpy(dv0, voveru0 + nempm, sizeof(double) * nbloc0, hipMemcpyHostToDevice) != hipSuccess) ||
        (nbloc1 > 0 && hipMemcpy(dv1, voveru1 + nempm, sizeof(double) * nbloc1, hipMemcpyHostToDevice) != hipSuccess))
      r = TG_E_HIP;
  }
  // Rn = D0^H (ROT D1); R2 = Rn(nempm:, nempm:)
  if (r == TG_OK) r = gemm_zz(s, OP_N, OP_N, n, n, n, one, dR, n, 0, dD1, n, 0, zero, dT, n, 0, 1);
  if (r == TG_OK) r = gemm_zz(s, OP_C, OP_N, n2, n2, n, one, dD0 + (size_t)nempm * n, n, 0,
                              dT + (size_t)nempm * n, n, 0, zero, dRn, n2, 0, 1);
  if (r == TG_OK) r = batched_inverse(s, n2, dRn, n2, 0, 1, dDet, dInfo, dPerm);
  int info = 0;
  cplx det = cmk(0, 0);
  if (r == TG_OK && (hipDeviceSynchronize() != hipSuccess ||
                     hipMemcpy(&info, dInfo, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(&det, dDet, sizeof(cplx), hipMemcpyDeviceToHost) != hipSuccess))
    r = TG_E_HIP;
  if (r == TG_OK && info != 0) {
    tg_set_error("In calculate_overlap got info = " + std::to_string(info) + " from zgetrf");
    r = TG_E_SINGULAR;
  }
  if (r == TG_OK && (isla0 + isla1) != 3) {
    if (nreg > 0 && nreg % 2 == 0) {
      hipLaunchKernelGGL(build_mreg_kernel, dim3(nblk((long long)mr)), dim3(NT), 0, s, n2, nbloc0,
                         nbloc1, nreg, 1, 0LL, 1, dv0, dv1, 0LL, dRn, dM);
      if (hipGetLastError() != hipSuccess) r = TG_E_HIP;
      if (r == TG_OK) r = batched_pfaffian(s, nreg, dM, nreg, 0, 1, dPW, dL, dPh, dZ);
      double la = 0;
      cplx ph = cmk(0, 0);
      int z = 0;
      if (r == TG_OK && (hipDeviceSynchronize() != hipSuccess ||
                         hipMemcpy(&la, dL, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                         hipMemcpy(&ph, dPh, sizeof(cplx), hipMemcpyDeviceToHost) != hipSuccess ||
                         hipMemcpy(&z, dZ, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
        r = TG_E_HIP;
      if (r == TG_OK) {
        if (z) {
          ov = cmk(0, 0);
        } else {
          const double m = exp(la + normvac);
          const cplx d = cscale(det, signfac);
          ov = cscale(cmul(d, ph), m);
        }
      }
    } else {
      ov = cmk((nreg == 0) ? 1.0 : 0.0, 0.0);
      if (nreg == 0) ov = cscale(cmul(cscale(det, signfac), ov), exp(normvac));
    }
  } else if (r == TG_OK) {
    // both Slater determinants: overlap = detR * signfac * exp(normvac)
    ov = cscale(cscale(det, signfac), exp(normvac));
  }
  for (void* p : {(void*)dD0, (void*)dD1, (void*)dR, (void*)dT, (void*)dRn, (void*)dM, (void*)dPW,
                  (void*)dDet, (void*)dPh, (void*)dL, (void*)dv0, (void*)dv1, (void*)dInfo,
                  (void*)dPerm, (void*)dZ})
    if (p) (void)hipFree(p);
  if (r == TG_E_HIP) tg_set_error("tg_calculate_overlap: HIP error");
  if (r == TG_E_NOMEM) tg_set_error("tg_calculate_overlap: hipMalloc failed");
  if (r == TG_OK) {
    overlap[0] = ov.x;
    overlap[1] = ov.y;
  }
  return r;
}

int tg_set_state(tg_ctx* c, const double* U0, const double* V0, const double* D0, int nocc0,
                 int nemp0, double ovac0, const double* voveru0, int Mphip, int Mphin,
                 double facpi, double valence_Z, double valence_N, double factor_delta,
                 int iprint) {
  if (!c) return TG_E_ARG;
  const int n = c->n;
  std::vector<double> vou(n, 0.0);
  if (voveru0 && nocc0 >= 0 && nocc0 <= n) std::copy(voveru0, voveru0 + (n - nocc0), vou.begin());
  return upload_states(c, 1, U0, V0, D0, &nocc0, &nemp0, &ovac0, voveru0 ? vou.data() : nullptr,
                       Mphip, Mphin, facpi, valence_Z, valence_N, factor_delta, iprint);
}

int tg_set_states(tg_ctx* c, int nstates, const double* U0, const double* V0, const double* D0,
                  const int* nocc0, const int* nemp0, const double* ovac0,
                  const double* voveru0, int Mphip, int Mphin, double facpi, double valence_Z,
                  double valence_N, double factor_delta, int iprint) {
  if (!c) return TG_E_ARG;
  return upload_states(c, nstates, U0, V0, D0, nocc0, nemp0, ovac0, voveru0, Mphip, Mphin, facpi,
                       valence_Z, valence_N, factor_delta, iprint);
}

int tg_project_resident(tg_ctx* c) {
  if (!c || !c->have_state) {
    tg_set_error("tg_project_resident: no state");
    return TG_E_ARG;
  }
  ScopedTiming tm(c, "total");
  return project_device(c);
}

int tg_sync(tg_ctx* c) {
  if (!c) return TG_E_ARG;
  TG_HIP(hipStreamSynchronize(c->stream));
  return TG_OK;
}

int tg_get_results(tg_ctx* c, double* pnp_scalars, double* pnp_A, double* pnp_H20,
                   double* pnp_AH20, double* field_H20v, double* imag_grad) {
  if (!c || !c->have_state) return TG_E_ARG;
  return fetch_results(c, 0, pnp_scalars, pnp_A, pnp_H20, pnp_AH20, field_H20v, imag_grad);
}

int tg_get_results_state(tg_ctx* c, int state, double* pnp_scalars, double* pnp_A,
                         double* pnp_H20, double* pnp_AH20, double* field_H20v,
                         double* imag_grad) {
  if (!c || !c->have_state) return TG_E_ARG;
  return fetch_results(c, state, pnp_scalars, pnp_A, pnp_H20, pnp_AH20, field_H20v, imag_grad);
}

int tg_project_wavefunction(tg_ctx* c, int iprint, const double* U0, const double* V0,
                            const double* D0, int nocc0, int nemp0, double ovac0,
                            const double* voveru0, int Mphip, int Mphin, double facpi,
                            double valence_Z, double valence_N, double factor_delta,
                            double* pnp_scalars, double* pnp_A, double* pnp_H20,
                            double* pnp_AH20, double* field_H20v, double* imag_grad) {
  if (!c) return TG_E_ARG;
  TG_TRY(tg_set_state(c, U0, V0, D0, nocc0, nemp0, ovac0, voveru0, Mphip, Mphin, facpi, valence_Z,
                      valence_N, factor_delta, iprint));
  TG_TRY(tg_project_resident(c));
  return fetch_results(c, 0, pnp_scalars, pnp_A, pnp_H20, pnp_AH20, field_H20v, imag_grad);
}

int tg_calculate_gradient(int ndim, int gradient_type, double gradient_eta, double gradient_mu,
                          int info_H11, int info_hsp, const double* eigen_H11,
                          const double* eigen_hsp, const double* field_H20v, int constraint_dim,
                          const double* constraint_20, const double* lagrange_lambda1,
                          const double* gradient_Zim1, double* gradient_Zi,
                          double* gradient_norm) {
  const int n = ndim;
  if (n <= 0 || !field_H20v || !gradient_Zim1 || !gradient_Zi || !gradient_norm ||
      constraint_dim < 0 || (constraint_dim > 0 && (!constraint_20 || !lagrange_lambda1))) {
    tg_set_error("tg_calculate_gradient: invalid argument");
    return TG_E_ARG;
  }
  int opt = gradient_type;
  if (info_H11 != 0 && gradient_type == 1) opt = 0;
  if (info_hsp != 0 && gradient_type == 2) opt = 0;
  double eta = gradient_eta, mu = gradient_mu;
  if (opt == 1 || opt == 2) {
    const double* ev = (opt == 1) ? eigen_H11 : eigen_hsp;
    if (!ev) {
      tg_set_error("tg_calculate_gradient: eigenvalues missing");
      return TG_E_ARG;
    }
    double amax = 0, amin = INFINITY;
    for (int i = 0; i < n; ++i) {
      amax = std::max(amax, fabs(ev[i]));
      amin = std::min(amin, fabs(ev[i]));
    }
    const double emax = (opt == 1 ? 4.0 : 2.0) * amax, emin = 2.0 * amin;
    const double cond = emax / emin;
    eta = pow(2.0 / (sqrt(emax) + sqrt(emin)), 2);
    mu = pow((sqrt(cond) - 1) / (sqrt(cond) + 1), 2);
  }
  // Z_i = -eta (H20 - sum lambda C20) + mu Z_{i-1}; ||Z||_F / eta
  const size_t n2 = (size_t)n * n;
  double sum = 0;
  for (size_t i = 0; i < n2; ++i) {
    double sg = 0;
    for (int j = 0; j < constraint_dim; ++j) sg += lagrange_lambda1[j] * constraint_20[j * n2 + i];
    const double z = -eta * (field_H20v[i] - sg) + mu * gradient_Zim1[i];
    gradient_Zi[i] = z;
    sum += z * z;
  }
  *gradient_norm = sqrt(sum) / eta;
  return TG_OK;
}

int tg_enable_timing(tg_ctx* c, int on) {
  if (!c) return TG_E_ARG;
  TG_TRY(timing_collect(c));
  c->timing = on ? 1 : 0;
  return TG_OK;
}

int tg_reset_timing(tg_ctx* c) {
  if (!c) return TG_E_ARG;
  TG_TRY(timing_collect(c));
  c->tsum.clear();
  c->tcount.clear();
  return TG_OK;
}

int tg_get_timing(tg_ctx* c, const char* key, double* ms_sum, int* count) {
  if (!c || !key) return TG_E_ARG;
  TG_TRY(timing_collect(c));
  auto it = c->tsum.find(key);
  if (ms_sum) *ms_sum = (it == c->tsum.end()) ? 0.0 : it->second;
  if (count) *count = (it == c->tsum.end()) ? 0 : c->tcount[key];
  return TG_OK;
}

int tg_set_gradient(tg_ctx* c, double gradient_eta, double gradient_mu, int constraint_dim,
                    const double* constraint_20, const double* lagrange_lambda1) {
  if (!c || constraint_dim < 0 || constraint_dim > TG_MAX_CONS ||
      (constraint_dim > 0 && (!constraint_20 || !lagrange_lambda1))) {
    tg_set_error("tg_set_gradient: invalid argument");
    return TG_E_ARG;
  }
  const size_t nn = (size_t)c->n * c->n;
  c->grad_on = 1;
  c->grad_eta = gradient_eta;
  c->grad_mu = gradient_mu;
  c->grad_ncons = constraint_dim;
  c->c20_shared = 1;           // one constraint_20 / lambda set for every state
  c->grad_lambda.assign(lagrange_lambda1, lagrange_lambda1 + constraint_dim);
  c->z_states = 0;             // zero Zi, Zim1 of every state
  TG_TRY(tg_ensure_z(c));      // (and uploads the multipliers)
  TG_TRY(c->d_c20.ensure(sizeof(double) * std::max<size_t>(1, (size_t)constraint_dim * nn)));
  if (constraint_dim > 0) {
    TG_HIP(hipMemcpy(c->d_c20.p, constraint_20, sizeof(double) * constraint_dim * nn,
                     hipMemcpyHostToDevice));
  }
  return TG_OK;
}

int tg_get_gradient(tg_ctx* c, double* gradient_Zi, double* gradient_norm) {
  return tg_get_gradient_state(c, 0, gradient_Zi, gradient_norm);
}

}  // extern "C"

namespace {
// Observables of the angles of the current grid (c->A of them, all resident in
// the work buffers after project_device): records a0 .. a0 + A - 1 of per_angle,
// and the rho term of sum_gauge added to rho_acc (complex n^2, internal order).
int observables_grid(tg_ctx* c, int nobo, const double* obo_ops, int nobos,
                     const double* obos_ops, int nshell, const int* sp_shell, int a0,
                     double* per_angle, cplx* rho_acc) {
  const int n = c->n, h = n / 2;
  const long long nn = (long long)n * n;
  const RunPlan P = make_plan(c);
  const int A = P.A;
  hipStream_t s = c->stream;
  if (P.chunk != P.B) {
    tg_set_error("tg_project_observables: the angles of a block do not fit into device memory");
    return TG_E_NOMEM;
  }
  // device layout: ops (nobo n^2) | obos (2 nobos n^2) | ecomp (12 A) | occ (2 nshell A)
  //               | obo (2 nobo A) | obos (4 nobos A) | rho sum (n^2) [complex] | shell, iso (int)
  const size_t nd_ops = (size_t)(nobo + 2 * nobos) * nn;
  const size_t nc_out = (size_t)A * (12 + 2 * nshell + 2 * nobo + 4 * nobos) + nn;
  TG_TRY(c->d_obs.ensure(sizeof(double) * nd_ops + sizeof(cplx) * nc_out + sizeof(int) * 2 * n + 256));
  double* d_ops = c->d_obs.as<double>();
  double* d_obos = d_ops + (size_t)nobo * nn;
  cplx* d_ec = reinterpret_cast<cplx*>(d_ops + nd_ops);
  cplx* d_occ = d_ec + 12 * (size_t)A;
  cplx* d_obo = d_occ + 2 * (size_t)nshell * A;
  cplx* d_obs2 = d_obo + 2 * (size_t)nobo * A;
  cplx* d_rho = d_obs2 + 4 * (size_t)nobos * A;
  int* d_shell = reinterpret_cast<int*>(d_rho + nn);
  int* d_iso = d_shell + n;
  {
    std::vector<double> hops(nd_ops);
    auto perm_op = [&](const double* src, double* dst) {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          dst[i + (size_t)j * n] = src[c->sp_perm[i] + (size_t)c->sp_perm[j] * n];
    };
    for (int k = 0; k < nobo; ++k) perm_op(obo_ops + k * nn, hops.data() + k * nn);
    for (int k = 0; k < 2 * nobos; ++k) perm_op(obos_ops + k * nn, hops.data() + (nobo + k) * nn);
    std::vector<int> hint(2 * n, 0);
    for (int i = 0; i < n; ++i) {
      hint[i] = nshell > 0 ? sp_shell[c->sp_perm[i]] : 0;
      hint[n + i] = c->sp_perm[i] >= h ? 1 : 0;
    }
    if (nd_ops) TG_HIP(hipMemcpyAsync(d_ops, hops.data(), sizeof(double) * nd_ops,
                                      hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(d_shell, hint.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, s));
    TG_HIP(hipStreamSynchronize(s));
  }
  auto Wb = [&](int k) { return P.work + (size_t)k * P.chunk * nn; };
  const cplx* rho = Wb(W_RHO);
  const cplx* kLR = Wb(W_KLR);
  const cplx* kRL = Wb(W_KRL);
  const cplx one = cmk(1, 0), zero = cmk(0, 0);
  hipLaunchKernelGGL(ecomp_kernel, dim3(A), dim3(NT), 0, s, n, c->d_H1.as<cplx>(), Wb(W_G), rho,
                     Wb(W_DLR), kRL, d_ec);
  if (nshell > 0)
    hipLaunchKernelGGL(occnum_kernel, dim3(A), dim3(NT), 0, s, n, nshell, d_shell, d_iso, rho, d_occ);
  if (nobo > 0)
    hipLaunchKernelGGL(obo_trace_kernel, dim3(A, nobo), dim3(NT), 0, s, n, d_ops, rho, d_obo);
  TG_HIP(hipGetLastError());
  for (int k = 0; k < nobos; ++k) {
    const double* O = d_obos + (size_t)2 * k * nn;
    const double* O2 = O + nn;
    cplx* X = Wb(W_X);
    cplx* Y = Wb(W_Y);
    cplx* Z = Wb(W_Z);
    TG_TRY(gemm_dz(s, OP_N, OP_N, n, n, n, one, O, n, 0, rho, n, nn, zero, X, n, nn, A));
    TG_TRY(gemm_zd(s, OP_N, OP_N, n, n, n, one, kRL, n, nn, O, n, 0, zero, Y, n, nn, A));
    TG_TRY(gemm_dz(s, OP_N, OP_N, n, n, n, one, O, n, 0, kLR, n, nn, zero, Z, n, nn, A));
    hipLaunchKernelGGL(obos_trace_kernel, dim3(A), dim3(NT), 0, s, n, X, Y, Z, O2, rho,
                       d_obs2 + (size_t)4 * k * A);
    TG_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(rho_sum_kernel, dim3(nblk(nn)), dim3(NT), 0, s, nn, A, P.d_f12, rho, d_rho);
  TG_HIP(hipGetLastError());
  TG_HIP(hipStreamSynchronize(s));
  // per-angle records
  std::vector<cplx> ov(A), pa(A), tr(9 * (size_t)A), en(A), out(nc_out);
  TG_HIP(hipMemcpy(ov.data(), P.d_over, sizeof(cplx) * A, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(pa.data(), P.d_pari, sizeof(cplx) * A, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(tr.data(), P.d_tr, sizeof(cplx) * 9 * A, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(en.data(), P.d_ener, sizeof(cplx) * A, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(out.data(), d_ec, sizeof(cplx) * nc_out, hipMemcpyDeviceToHost));
  const size_t R = 19 + 2 * (size_t)nshell + 2 * (size_t)nobo + 4 * (size_t)nobos;
  cplx* rec = reinterpret_cast<cplx*>(per_angle);
  const cplx* o_ec = out.data();
  const cplx* o_occ = o_ec + 12 * (size_t)A;
  const cplx* o_obo = o_occ + 2 * (size_t)nshell * A;
  const cplx* o_obs = o_obo + 2 * (size_t)nobo * A;
  for (int a = 0; a < A; ++a) {
    cplx* q = rec + R * (a0 + a);
    const cplx* t = tr.data() + 9 * (size_t)a;
    q[0] = ov[a];
    q[1] = pa[a];
    q[2] = t[0];
    q[3] = t[1];
    // calculate_particle_number (module_particlenumber.f90:78-143): <Z^2>, <N^2>
    q[4] = csub(cadd(t[0], cmul(t[0], t[0])), csub(t[2], t[3]));
    q[5] = csub(cadd(t[1], cmul(t[1], t[1])), csub(t[4], t[5]));
    q[6] = en[a];
    for (int k = 0; k < 12; ++k) q[7 + k] = o_ec[12 * (size_t)a + k];
    for (int k = 0; k < 2 * nshell; ++k) q[19 + k] = o_occ[2 * (size_t)nshell * a + k];
    for (int k = 0; k < 2 * nobo; ++k) q[19 + 2 * nshell + k] = o_obo[2 * (size_t)nobo * a + k];
    for (int k = 0; k < nobos; ++k)
      for (int m = 0; m < 4; ++m)
        q[19 + 2 * nshell + 2 * nobo + 4 * k + m] = o_obs[(size_t)4 * k * A + 4 * a + m];
  }
  for (long long i = 0; i < nn; ++i) rho_acc[i] = cadd(rho_acc[i], out[nc_out - nn + i]);
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_project_observables(tg_ctx* c, int nobo, const double* obo_ops, int nobos,
                           const double* obos_ops, int nshell, const int* sp_shell,
                           double* per_angle, double* pnp_rho) {
  if (!c || !c->have_state || c->ns != 1 || nobo < 0 || nobos < 0 || nshell < 0 ||
      (nobo > 0 && !obo_ops) || (nobos > 0 && !obos_ops) || (nshell > 0 && !sp_shell) ||
      !per_angle) {
    tg_set_error("tg_project_observables: invalid argument (one resident state)");
    return TG_E_ARG;
  }
  const int n = c->n, h = n / 2;
  const long long nn = (long long)n * n;
  for (int i = 0; i < n && nshell > 0; ++i)
    if (sp_shell[i] < 0 || sp_shell[i] >= nshell) {
      tg_set_error("tg_project_observables: shell index out of range");
      return TG_E_ARG;
    }
  // every angle of the grid (no pairing of conjugate angles) with the parity
  // overlaps (iopt + iprint > 0, :279-284), in blocks of angles: all at once when
  // they fit, one angle at a time in the plain per-angle mode, blocks of the
  // memory-short chunk size otherwise.  The caller's angle set-up, iprint and
  // gradient step are restored afterwards.
  const std::vector<double> sv_phip = c->phip, sv_phin = c->phin, sv_mult = c->mult;
  const std::vector<cplx> sv_weight = c->weight;
  const int sv_A = c->A, sv_iprint = c->iprint, sv_grad = c->grad_on, sv_graph = c->use_graph;
  auto restore = [&]() {
    c->phip = sv_phip;
    c->phin = sv_phin;
    c->mult = sv_mult;
    c->weight = sv_weight;
    c->A = sv_A;
    c->iprint = sv_iprint;
    c->grad_on = sv_grad;
    c->use_graph = sv_graph;
    c->state_dirty = 1;
  };
  assign_angles(c, false);
  const std::vector<double> all_phip = c->phip, all_phin = c->phin, all_mult = c->mult;
  const std::vector<cplx> all_weight = c->weight;
  const int Aall = c->A;
  c->iprint = 1;
  c->grad_on = 0;
  c->use_graph = 0;
  std::vector<cplx> rho_acc((size_t)nn, cmk(0.0, 0.0));
  int blk = c->plain ? 1 : Aall;
  int r = TG_OK;
  for (int a0 = 0; a0 < Aall && r == TG_OK;) {
    const int na = std::min(blk, Aall - a0);
    c->phip.assign(all_phip.begin() + a0, all_phip.begin() + a0 + na);
    c->phin.assign(all_phin.begin() + a0, all_phin.begin() + a0 + na);
    c->mult.assign(all_mult.begin() + a0, all_mult.begin() + a0 + na);
    c->weight.assign(all_weight.begin() + a0, all_weight.begin() + a0 + na);
    c->A = na;
    c->state_dirty = 1;
    r = project_device(c);
    if (r != TG_OK) break;
    const RunPlan P = make_plan(c);
    if (P.chunk != P.B) {
      // memory-short: the device holds P.chunk angles' work matrices at once
      if (P.chunk >= na) { r = TG_E_NOMEM; break; }
      blk = P.chunk;
      continue;
    }
    r = observables_grid(c, nobo, obo_ops, nobos, obos_ops, nshell, sp_shell, a0, per_angle,
                         rho_acc.data());
    a0 += na;
  }
  restore();
  TG_TRY(r);
  if (pnp_rho) permute_out(c, rho_acc.data(), reinterpret_cast<cplx*>(pnp_rho));
  return TG_OK;
}

int tg_get_gradient_state(tg_ctx* c, int state, double* gradient_Zi, double* gradient_norm) {
  if (!c || !c->grad_on || c->z_states < 1) {
    tg_set_error("tg_get_gradient: tg_set_gradient was not called");
    return TG_E_ARG;
  }
  if (state < 0 || state >= c->z_states) {
    tg_set_error("tg_get_gradient: state index out of range");
    return TG_E_ARG;
  }
  const size_t nn = (size_t)c->n * c->n;
  const int S = c->z_states;
  TG_HIP(hipStreamSynchronize(c->stream));
  const double* Zb = c->d_Z.as<double>();
  if (gradient_Zi)
    TG_HIP(hipMemcpy(gradient_Zi, Zb + state * nn, sizeof(double) * nn, hipMemcpyDeviceToHost));
  const size_t nb = (nn + NT - 1) / NT;
  std::vector<double> part(nb);
  TG_HIP(hipMemcpy(part.data(), Zb + 2 * nn * S + (size_t)S * TG_MAX_CONS + state * tg_part_len(c->n),
                   sizeof(double) * nb, hipMemcpyDeviceToHost));
  double sum = 0;
  for (double v : part) sum += v;
  const double eta = (c->grad_type > 0 && c->etamu.size() == 2 * (size_t)S) ? c->etamu[2 * state]
                                                                            : c->grad_eta;
  if (gradient_norm) *gradient_norm = sqrt(sum) / eta;
  return TG_OK;
}

}  // extern "C"
