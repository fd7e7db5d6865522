// Two-body contraction of calculate_fields (src/module_fields.f90:99-189) as
// grouped dense GEMMs.
//
// The reference walks the stored TBME list (one representative per symmetry
// class, src/module_hamiltonian.f90:330-378) and scatters 8 Gamma and 4 Delta
// updates per element and time-reversal pass (module_fields.f90:128-157,
// find_timerev :1028-1089).  Here the same list is re-laid out once, at
// tg_init, into two block-diagonal matrices:
//   ph channel  Gamma_ac = sum_bd W[(a,c),(d,b)] rho_db,
//               blocks labelled by (2m_a-2m_c, 2t_a-2t_c, l_a+l_c mod 2);
//   pp channel  Delta_xy = sum_{u>v} P[(x,y),(u,v)] kappa_uv  (x>y),
//               blocks labelled by (2m_x+2m_y, 2t_x+2t_y, l_x+l_y mod 2).
// W and P are filled by adding exactly the reference's 8 + 4 updates, so every
// element keeps its value (FP32 when the TBMEs are FP32: no rounding added) and
// the angles of a batch become the M dimension of one GEMM per block: the
// scattered, atomic-prone loop turns into dense FP64 arithmetic with each
// matrix element read once per batch.
#include <algorithm>
#include <map>
#include <tuple>

#include "tg_internal.h"

namespace {

constexpr int NT = 256;

// Tiled transposes between the angle-major matrices (angle a at src + a*nn,
// element (x, y) at linear index x + y*n) and the position-major contraction
// layouts (X[row0 + a + p*ld], position p of a channel).  A workgroup moves a
// tile of TL linear indices x TA angles through LDS: the matrix side is read or
// written as TA runs of TL consecutive elements, the contraction side as TL runs
// of TA consecutive angles, so both sides are coalesced.  inv maps linear
// indices to positions (Channel::inv).
constexpr int TL = 64;   // linear indices per tile
constexpr int TA = 32;   // angles per tile

// X[row0 + a + inv[lin]*ld] = src_a[lin] for every lin with inv[lin] >= 0
__global__ void __launch_bounds__(NT)
gather_tiled_kernel(int na, long long nn, const int* __restrict__ inv,
                    const cplx* __restrict__ src, int ld, int row0, cplx* __restrict__ X) {
  __shared__ cplx tile[TL][TA + 1];
  const long long lin0 = (long long)blockIdx.x * TL;
  const int a0 = blockIdx.y * TA;
  const int t = threadIdx.x;
  {
    const int l = t % TL;
    const long long lin = lin0 + l;
    for (int r = t / TL; r < TA; r += NT / TL) {
      const int a = a0 + r;
      if (a < na && lin < nn) tile[l][r] = src[(long long)a * nn + lin];
    }
  }
  __syncthreads();
  const int r = t % TA;
  for (int l = t / TA; l < TL; l += NT / TA) {
    const long long lin = lin0 + l;
    const int a = a0 + r;
    if (lin < nn && a < na) {
      const int q = inv[lin];
      if (q >= 0) X[row0 + a + (long long)q * ld] = tile[l][r];
    }
  }
}

// dst_a[lin] = field at lin from C[row0 + a + q*ld]:  q = inv[lin] >= 0: C;
// pp (q <= -2): -C at position -2 - q; the diagonal (q == -1): 0.  With hsp:
// hsp_a[lin] = dst_a[lin] + H1[lin].
__global__ void __launch_bounds__(NT)
scatter_tiled_kernel(int na, long long nn, const int* __restrict__ inv,
                     const cplx* __restrict__ C, int ld, int row0, cplx* __restrict__ dst,
                     const cplx* __restrict__ H1, cplx* __restrict__ hsp) {
  __shared__ cplx tile[TL][TA + 1];
  const long long lin0 = (long long)blockIdx.x * TL;
  const int a0 = blockIdx.y * TA;
  const int t = threadIdx.x;
  {
    const int r = t % TA;
    for (int l = t / TA; l < TL; l += NT / TA) {
      const long long lin = lin0 + l;
      const int a = a0 + r;
      if (lin < nn && a < na) {
        const int q = inv[lin];
        cplx v = cmk(0.0, 0.0);
        if (q >= 0) {
          v = C[row0 + a + (long long)q * ld];
        } else if (q <= -2) {
          v = C[row0 + a + (long long)(-2 - q) * ld];
          v = cmk(-v.x, -v.y);
        }
        tile[l][r] = v;
      }
    }
  }
  __syncthreads();
  const int l = t % TL;
  const long long lin = lin0 + l;
  if (lin >= nn) return;
  const cplx h = hsp ? H1[lin] : cmk(0.0, 0.0);
  for (int r = t / TL; r < TA; r += NT / TL) {
    const int a = a0 + r;
    if (a < na) {
      const cplx v = tile[l][r];
      dst[(long long)a * nn + lin] = v;
      if (hsp) hsp[(long long)a * nn + lin] = cadd(v, h);
    }
  }
}

// hsp_a = Gamma_a + H1
__global__ void add_h1_kernel(long long nn, int na, const cplx* __restrict__ H1,
                              const cplx* __restrict__ G, cplx* __restrict__ h) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  if (idx >= nn * na) return;
  h[idx] = cadd(G[idx], H1[idx % nn]);
}

// calculate_fields_diag: Gamma(i,j) = conj(Gamma(j,i)) for i > j; DeltaRL = conj(DeltaLR)
__global__ void hermitize_kernel(int n, int na, long long nn, cplx* __restrict__ G,
                                 const cplx* __restrict__ dLR, cplx* __restrict__ dRL) {
  long long idx = (long long)blockIdx.x * NT + threadIdx.x;
  long long tot = (long long)na * nn;
  if (idx >= tot) return;
  long long a = idx / nn;
  int r = (int)(idx % nn);
  int i = r % n, j = r / n;
  if (G != nullptr && i > j) G[a * nn + r] = cconj(G[a * nn + j + (long long)i * n]);
  if (dRL != nullptr) dRL[idx] = cconj(dLR[idx]);
}

inline unsigned nblk(long long tot) { return (unsigned)((tot + NT - 1) / NT); }

struct KeyCmp {
  bool operator()(const std::tuple<int, int, int>& a, const std::tuple<int, int, int>& b) const {
    return a < b;
  }
};

// Assigns each element (x,y) of a channel to its block and position.
void plan_channel(Channel& ch, int n, const std::vector<int>& m, const std::vector<int>& t,
                  const std::vector<int>& l, bool pp, std::vector<int>& pos_of) {
  std::map<std::tuple<int, int, int>, std::vector<int>, KeyCmp> classes;
  pos_of.assign((size_t)n * n, -1);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      if (pp && x <= y) continue;
      std::tuple<int, int, int> key = pp ? std::make_tuple(m[x] + m[y], t[x] + t[y], (l[x] + l[y]) & 1)
                                         : std::make_tuple(m[x] - m[y], t[x] - t[y], (l[x] + l[y]) & 1);
      classes[key].push_back(x + y * n);
    }
  ch.groups.clear();
  ch.perm.clear();
  long long woff = 0;
  int off = 0;
  for (auto& kv : classes) {
    GroupDesc g;
    g.off = off;
    g.size = (int)kv.second.size();
    g.woff = woff;
    ch.groups.push_back(g);
    for (size_t k = 0; k < kv.second.size(); ++k) {
      pos_of[kv.second[k]] = off + (int)k;
      ch.perm.push_back(kv.second[k]);
    }
    off += g.size;
    woff += (long long)g.size * g.size;
  }
  ch.npos = off;
  ch.ngroups = (int)ch.groups.size();
  ch.wsize = woff;
  ch.inv = pos_of;
  if (pp)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < y; ++x) ch.inv[x + (size_t)y * n] = -2 - pos_of[y + (size_t)x * n];
}

}  // namespace

int contraction_build(tg_ctx* c, int64_t ntb, const int16_t* abcd, const void* H2, int h2_bytes,
                      const int8_t* trperm) {
  const int n = c->n;
  std::vector<int> pos_ph, pos_pp;
  plan_channel(c->ph, n, c->sp_2mj, c->sp_2mt, c->sp_l, false, pos_ph);
  plan_channel(c->pp, n, c->sp_2mj, c->sp_2mt, c->sp_l, true, pos_pp);
  std::vector<double> Wph((size_t)c->ph.wsize, 0.0), Wpp((size_t)c->pp.wsize, 0.0);

  // group index of each position
  std::vector<int> gph(c->ph.npos), gpp(c->pp.npos);
  for (int g = 0; g < c->ph.ngroups; ++g)
    for (int k = 0; k < c->ph.groups[g].size; ++k) gph[c->ph.groups[g].off + k] = g;
  for (int g = 0; g < c->pp.ngroups; ++g)
    for (int k = 0; k < c->pp.groups[g].size; ++k) gpp[c->pp.groups[g].off + k] = g;

  bool bad = false;
  // W(q,p) of the group holding both positions, stored at woff + q + p*s
  auto add_ph = [&](int a, int cc, int d, int b, double v) {
    int po = pos_ph[a + cc * n], pi = pos_ph[d + b * n];
    int g = gph[po];
    if (gph[pi] != g) { bad = true; return; }
    const GroupDesc& G = c->ph.groups[g];
    Wph[(size_t)(G.woff + (long long)(pi - G.off) + (long long)(po - G.off) * G.size)] += v;
  };
  auto add_pp = [&](int x, int y, int u, int w, double v) {
    // stored pairs have x > y and u > w; Delta and kappa are skew, so a pair
    // in the other order enters with the opposite sign
    if (x < y) { std::swap(x, y); v = -v; }
    if (u < w) { std::swap(u, w); v = -v; }
    int po = pos_pp[x + y * n], pi = pos_pp[u + w * n];
    if (po < 0 || pi < 0) { bad = true; return; }
    int g = gpp[po];
    if (gpp[pi] != g) { bad = true; return; }
    const GroupDesc& G = c->pp.groups[g];
    Wpp[(size_t)(G.woff + (long long)(pi - G.off) + (long long)(po - G.off) * G.size)] += v;
  };

  const std::vector<int>& tr = c->sp_tr;
  for (int64_t kk = 0; kk < ntb; ++kk) {
    int ia = abcd[4 * kk] - 1, ib = abcd[4 * kk + 1] - 1, ic = abcd[4 * kk + 2] - 1,
        id = abcd[4 * kk + 3] - 1;
    if (ia < 0 || ib < 0 || ic < 0 || id < 0 || ia >= n || ib >= n || ic >= n || id >= n) {
      tg_set_error("tg_init: TBME index out of range");
      return TG_E_ARG;
    }
    // internal single-particle order
    ia = c->sp_iperm[ia];
    ib = c->sp_iperm[ib];
    ic = c->sp_iperm[ic];
    id = c->sp_iperm[id];
    double h = (h2_bytes == 4) ? (double)((const float*)H2)[kk] : ((const double*)H2)[kk];
    const int perm = trperm[kk];
    for (int it = 0; it < 2; ++it) {
      if (it == 1) {
        if (c->sp_2mj[ia] + c->sp_2mj[ib] == 0) break;
        // find_timerev (module_hamiltonian.f90:1028-1089)
        int ta = tr[ia], tb = tr[ib], tc = tr[ic], td = tr[id];
        int code = perm < 0 ? perm + 8 : perm;
        switch (code) {
          case 0: ia = ta; ib = tb; ic = tc; id = td; break;
          case 1: ia = tb; ib = ta; ic = tc; id = td; break;
          case 2: ia = ta; ib = tb; ic = td; id = tc; break;
          case 3: ia = tb; ib = ta; ic = td; id = tc; break;
          case 4: ia = tc; ib = td; ic = ta; id = tb; break;
          case 5: ia = tc; ib = td; ic = tb; id = ta; break;
          case 6: ia = td; ib = tc; ic = ta; id = tb; break;
          case 7: ia = td; ib = tc; ic = tb; id = ta; break;
          default: tg_set_error("tg_init: bad trperm code"); return TG_E_ARG;
        }
        h = (perm < 0) ? -h : h;
      }
      const double f = (ia == ic && ib == id) ? 0.0 : h;
      add_ph(ia, ic, id, ib, h);
      add_ph(ia, id, ic, ib, -h);
      add_ph(ib, ic, id, ia, -h);
      add_ph(ib, id, ic, ia, h);
      add_ph(ic, ia, ib, id, f);
      add_ph(ic, ib, ia, id, -f);
      add_ph(id, ia, ib, ic, -f);
      add_ph(id, ib, ia, ic, f);
      add_pp(ib, ia, id, ic, h);
      add_pp(id, ic, ib, ia, f);
    }
  }
  if (bad) {
    tg_set_error("tg_init: TBME list violates the M / T_z / parity selection rules");
    return TG_E_ARG;
  }

  // upload
  for (Channel* ch : {&c->ph, &c->pp}) {
    std::vector<double>& W = (ch == &c->ph) ? Wph : Wpp;
    TG_TRY(ch->d_inv.ensure(sizeof(int) * ch->inv.size()));
    TG_HIP(hipMemcpy(ch->d_inv.p, ch->inv.data(), sizeof(int) * ch->inv.size(),
                     hipMemcpyHostToDevice));
    TG_TRY(ch->d_groups.ensure(sizeof(GroupDesc) * ch->groups.size()));
    TG_HIP(hipMemcpy(ch->d_groups.p, ch->groups.data(), sizeof(GroupDesc) * ch->groups.size(),
                     hipMemcpyHostToDevice));
    if (c->w_fp32) {
      std::vector<float> Wf(W.size());
      for (size_t i = 0; i < W.size(); ++i) Wf[i] = (float)W[i];
      TG_TRY(ch->d_W.ensure(sizeof(float) * Wf.size()));
      TG_HIP(hipMemcpy(ch->d_W.p, Wf.data(), sizeof(float) * Wf.size(), hipMemcpyHostToDevice));
    } else {
      TG_TRY(ch->d_W.ensure(sizeof(double) * W.size()));
      TG_HIP(hipMemcpy(ch->d_W.p, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice));
    }
    // tile table: (group, n0) for every 64-column tile of every group
    std::vector<int> tg, tn;
    for (int g = 0; g < ch->ngroups; ++g)
      for (int n0 = 0; n0 < ch->groups[g].size; n0 += 64) {
        tg.push_back(g);
        tn.push_back(n0);
      }
    ch->ntiles_total = (int)tg.size();
    TG_TRY(ch->d_tile_group.ensure(sizeof(int) * tg.size()));
    TG_TRY(ch->d_tile_n0.ensure(sizeof(int) * tn.size()));
    TG_HIP(hipMemcpy(ch->d_tile_group.p, tg.data(), sizeof(int) * tg.size(), hipMemcpyHostToDevice));
    TG_HIP(hipMemcpy(ch->d_tile_n0.p, tn.data(), sizeof(int) * tn.size(), hipMemcpyHostToDevice));
  }
  return TG_OK;
}

int contraction_run(tg_ctx* c, int na, const cplx* rho, const cplx* kLR, const cplx* kRL,
                    cplx* gamma, cplx* dLR, cplx* dRL, int hermitian, const cplx* H1, cplx* hsp,
                    bool defer_join) {
  const int n = c->n;
  const long long nn = (long long)n * n;
  hipStream_t s = c->stream;
  const int nph = c->ph.npos, npp = c->pp.npos;
  const int two = (kRL != nullptr && !hermitian) ? 2 : 1;   // pp right-hand sides per angle
  // layout: Xph (na x nph), Cph, Xpp (two*na x npp), Cpp
  size_t need = sizeof(cplx) * ((size_t)2 * na * nph + (size_t)2 * two * na * npp);
  TG_TRY(c->d_blk.ensure(need));
  cplx* Xph = c->d_blk.as<cplx>();
  cplx* Cph = Xph + (size_t)na * nph;
  cplx* Xpp = Cph + (size_t)na * nph;
  cplx* Cpp = Xpp + (size_t)two * na * npp;

  const int* iph = c->ph.d_inv.as<int>();
  const int* ipp = c->pp.d_inv.as<int>();
  const dim3 tiles((unsigned)((nn + TL - 1) / TL), (unsigned)((na + TA - 1) / TA));
  // the pp channel (gathers, product, scatters) on the third stream beside the ph
  // one: the groups are of very uneven size, so each product's tail leaves CUs the
  // other can use, and the gathers and scatters (HBM-bound) overlap products
  hipStream_t s3 = c->stream3;
  TG_HIP(hipEventRecord(c->ev_fork3, s));
  TG_HIP(hipStreamWaitEvent(s3, c->ev_fork3, 0));
  hipLaunchKernelGGL(gather_tiled_kernel, tiles, dim3(NT), 0, s, na, nn, iph, rho, na, 0, Xph);
  hipLaunchKernelGGL(gather_tiled_kernel, tiles, dim3(NT), 0, s3, na, nn, ipp, kLR, two * na, 0,
                     Xpp);
  if (two == 2)
    hipLaunchKernelGGL(gather_tiled_kernel, tiles, dim3(NT), 0, s3, na, nn, ipp, kRL, two * na, na,
                       Xpp);
  TG_HIP(hipGetLastError());
  if (!gemm_grouped_tuned(na, c->ph.ntiles_total, 0) ||
      !gemm_grouped_tuned(two * na, c->pp.ntiles_total, 1)) {
    // first call for this batch size: both channels' tile shapes are timed on the
    // context stream alone (outputs recomputed below)
    TG_HIP(hipEventRecord(c->ev_join3, s3));
    TG_HIP(hipStreamWaitEvent(s, c->ev_join3, 0));
    TG_TRY(gemm_tune_grouped(s, na, Xph, Cph, c->ph.d_W.p, c->w_fp32,
                             c->ph.d_groups.as<GroupDesc>(), c->ph.d_tile_group.as<int>(),
                             c->ph.d_tile_n0.as<int>(), c->ph.ntiles_total, 0));
    TG_TRY(gemm_tune_grouped(s, two * na, Xpp, Cpp, c->pp.d_W.p, c->w_fp32,
                             c->pp.d_groups.as<GroupDesc>(), c->pp.d_tile_group.as<int>(),
                             c->pp.d_tile_n0.as<int>(), c->pp.ntiles_total, 1));
    TG_HIP(hipEventRecord(c->ev_fork3, s));
    TG_HIP(hipStreamWaitEvent(s3, c->ev_fork3, 0));
  }
  {
    ScopedTiming tk(c, "kernel:grouped_pp", s3);
    TG_TRY(gemm_grouped(s3, two * na, Xpp, Cpp, c->pp.d_W.p, c->w_fp32,
                        c->pp.d_groups.as<GroupDesc>(), c->pp.d_tile_group.as<int>(),
                        c->pp.d_tile_n0.as<int>(), 1, c->pp.ntiles_total));
  }
  hipLaunchKernelGGL(scatter_tiled_kernel, tiles, dim3(NT), 0, s3, na, nn, ipp, Cpp, two * na, 0,
                     dLR, (const cplx*)nullptr, (cplx*)nullptr);
  if (two == 2)
    hipLaunchKernelGGL(scatter_tiled_kernel, tiles, dim3(NT), 0, s3, na, nn, ipp, Cpp, two * na, na,
                       dRL, (const cplx*)nullptr, (cplx*)nullptr);
  TG_HIP(hipGetLastError());
  {
    ScopedTiming tk(c, "kernel:grouped_ph");
    TG_TRY(gemm_grouped(s, na, Xph, Cph, c->ph.d_W.p, c->w_fp32, c->ph.d_groups.as<GroupDesc>(),
                        c->ph.d_tile_group.as<int>(), c->ph.d_tile_n0.as<int>(), 0,
                        c->ph.ntiles_total));
  }
  // Gamma (and h = Gamma + H1 when no hermitian completion follows), DeltaLR, DeltaRL
  const bool fuse_h = hsp != nullptr && !hermitian;
  hipLaunchKernelGGL(scatter_tiled_kernel, tiles, dim3(NT), 0, s, na, nn, iph, Cph, na, 0, gamma,
                     fuse_h ? H1 : nullptr, fuse_h ? hsp : nullptr);
  TG_HIP(hipGetLastError());
  if (defer_join && two == 2 && fuse_h) return TG_OK;
  TG_TRY(contraction_join(c));
  if (hermitian)
    hipLaunchKernelGGL(hermitize_kernel, dim3(nblk((long long)na * nn)), dim3(NT), 0, s, n, na, nn,
                       gamma, dLR, dRL);
  else if (two == 1 && dRL != nullptr)
    hipLaunchKernelGGL(hermitize_kernel, dim3(nblk((long long)na * nn)), dim3(NT), 0, s, n, na, nn,
                       (cplx*)nullptr, dLR, dRL);
  if (hsp != nullptr && !fuse_h)
    hipLaunchKernelGGL(add_h1_kernel, dim3(nblk((long long)na * nn)), dim3(NT), 0, s, nn, na, H1,
                       gamma, hsp);
  TG_HIP(hipGetLastError());
  return TG_OK;
}

int contraction_join(tg_ctx* c) {
  TG_HIP(hipEventRecord(c->ev_join3, c->stream3));
  TG_HIP(hipStreamWaitEvent(c->stream, c->ev_join3, 0));
  return TG_OK;
}
