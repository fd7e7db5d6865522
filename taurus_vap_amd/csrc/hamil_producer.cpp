// m-scheme two-body matrix elements in the reference's storage convention.
//
// Restates the producer side of the TBME list consumed by the contraction:
//   * the symmetry reconstruction of set_hamiltonian_2body (src/module_hamiltonian.
//     f90:284-319): every stored coupled element, in the reference's loop order,
//     copied to its seven index permutations with the exchange phases, the
//     isospin channel of a J-scheme element following the permutation (find_iso,
//     :911-959);
//   * the m-scheme loop (:330-378): one representative per symmetry class, a<b,
//     c>=a, d>c (d<=b when c==a), M = m_a+m_b >= 0, parity / T_z / M
//     conservation, cut |V| > 1e-16, stored in loop order;
//   * uncouple_JT (:828-885) for JT-scheme (ANTOINE, types 1-2) interactions and
//     uncouple_J (:747-814) for J-scheme (types 3-4) ones;
//   * reconstruct_2body_timerev (:970-1018): the int8 permutation/phase code of
//     the time-reversed element.
// Clebsch-Gordan coefficients are the Racah formula on log-gamma values,
// tabulated once for all (j1, j2, J, m1, m2).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/taurus_hamil.h"

namespace {

// <j1 m1 j2 m2 | j3 m3>, all arguments twice their values: the Racah formula
// (j1,m1,j2,m2|j3,m3) = sqrt(2 j3 + 1) D(j1 j2 j3) [factorials of j +- m]^1/2
// sum_l (-1)^l / [factorials], with the factorials as log-gamma values and the
// sum as a ratio recurrence of its terms from the first non-vanishing one.
// The operation order is that of ClebschGordan (src/module_mathmethods.f90:
// 281-342), so the stored FP64 elements are bit-identical to the reference's.
double cg_lgamma(int j1, int j2, int j3, int m1, int m2, int m3) {
  const int n1 = 1 + (j1 + j2 - j3) / 2, n2 = 1 + (j2 + j3 - j1) / 2, n3 = 1 + (j3 + j1 - j2) / 2;
  const int n4 = 1 + (j1 - m1) / 2, n5 = 1 + (j2 + m2) / 2, n6 = 1 + (j1 + m1) / 2;
  const int n7 = 1 + (j2 - m2) / 2, n8 = 1 + (j3 + m3) / 2, n9 = 1 + (j3 - m3) / 2;
  const int n10 = n1 + n2 + n3 - 1;
  int nmin = n1;
  for (int v : {n2, n3, n4, n5, n6, n7, n8, n9}) nmin = v < nmin ? v : nmin;
  if (nmin < 1 || m1 + m2 != m3) return 0.0;
  const double p = lgamma((double)n1) + lgamma((double)n2) + lgamma((double)n3) +
                   lgamma((double)n4) + lgamma((double)n5) + lgamma((double)n6) +
                   lgamma((double)n7) + lgamma((double)n8) + lgamma((double)n9) -
                   lgamma((double)n10);
  int k1 = n1, k2 = n4, k3 = n5, k4 = n4 - n3, k5 = n5 - n2;
  int l1 = 0;
  for (int v : {k4, k5}) l1 = v > l1 ? v : l1;
  int l2 = k1;
  for (int v : {k2, k3}) l2 = v < l2 ? v : l2;
  double h = 1.0, hl = 1.0;
  for (int l = l1 + 1; l <= l2; ++l) {
    const double hlm1 = hl;
    hl = hlm1 * (double)(l - k1) * (double)(l - k2) * (double)(l - k3) /
         (double)((l - k4) * (l - k5) * l);
    h = h + hl;
  }
  k1 -= l1;
  k2 -= l1;
  k3 -= l1;
  k4 = l1 + 1 - k4;
  k5 = l1 + 1 - k5;
  const int k6 = l1 + 1;
  const double q = lgamma((double)k1) + lgamma((double)k2) + lgamma((double)k3) +
                   lgamma((double)k4) + lgamma((double)k5) + lgamma((double)k6);
  const double sgn = (l1 & 1) ? -1.0 : 1.0;
  return sqrt((double)j3 + 1.0) * sgn * exp(0.5 * p - q) * h;
}

double cg_racah(int j1, int j2, int J, int m1, int m2) { return cg_lgamma(j1, j2, J, m1, m2, m1 + m2); }

struct CGTable {
  int jmax2;   // largest 2j
  int Jmax;    // largest J
  std::vector<double> v;
  int dj, dJ, dm;
  void build(int jm2) {
    jmax2 = jm2;
    Jmax = jm2;
    dj = jm2 + 1;
    dJ = Jmax + 1;
    dm = 2 * jm2 + 1;
    v.assign((size_t)dj * dj * dJ * dm * dm, 0.0);
    for (int j1 = 1; j1 <= jm2; j1 += 2)
      for (int j2 = 1; j2 <= jm2; j2 += 2)
        for (int J = 0; J <= Jmax; ++J)
          for (int m1 = -j1; m1 <= j1; m1 += 2)
            for (int m2 = -j2; m2 <= j2; m2 += 2) at(j1, j2, J, m1, m2) = cg_racah(j1, j2, 2 * J, m1, m2);
  }
  double& at(int j1, int j2, int J, int m1, int m2) {
    return v[(((size_t)j1 * dj + j2) * dJ + J) * dm * dm + (size_t)(m1 + jmax2) * dm + (m2 + jmax2)];
  }
};


// Isospin channel of a J-scheme element after an index permutation of the
// symmetry reconstruction.  Channels: 0 = pp pp, 1 = pn pn, 2 = pn np,
// 3 = np pn, 4 = np np, 5 = nn nn.  Permutations 1..7 = abdc, bacd, badc, cdab,
// cdba, dcab, dcba: position k of the new element holds old position P[perm][k];
// the new channel follows from which of the moved particles are protons.
int iso_after_perm(int t, int perm) {
  if (t == 0 || t == 5) return t;
  static const int P[8][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {1, 0, 2, 3}, {1, 0, 3, 2},
                              {2, 3, 0, 1}, {2, 3, 1, 0}, {3, 2, 0, 1}, {3, 2, 1, 0}};
  const int bra_np = (t == 3 || t == 4), ket_np = (t == 2 || t == 4);
  const bool prot[4] = {!bra_np, (bool)bra_np, !ket_np, (bool)ket_np};
  const int nbra = prot[P[perm][0]] ? 0 : 1, nket = prot[P[perm][2]] ? 0 : 1;
  return 1 + 2 * nbra + nket;
}

// Channel of the m-scheme element (ta, tb, tc, td) (2 t_z, protons -1) in the
// J-scheme ordering; -1 when charge is not conserved.
int jscheme_channel(int ta, int tb, int tc, int td) {
  if (ta == -1 && tb == -1 && tc == -1 && td == -1) return 0;
  if (ta == -1 && tb == 1 && tc == -1 && td == 1) return 1;
  if (ta == -1 && tb == 1 && tc == 1 && td == -1) return 2;
  if (ta == 1 && tb == -1 && tc == -1 && td == 1) return 3;
  if (ta == 1 && tb == -1 && tc == 1 && td == -1) return 4;
  if (ta == 1 && tb == 1 && tc == 1 && td == 1) return 5;
  return -1;
}

// The m-scheme loop of set_hamiltonian_2body; `uncouple(a, b, c, d)` gives the
// m-scheme element of 0-based single-particle states.
template <class F>
int64_t mscheme_loop(int n, const int32_t* sp_2mj, const int32_t* sp_2mt, const int32_t* sp_l,
                     double scaling_2b, int64_t capacity, int16_t* abcd, double* H2, F uncouple) {
  const double Vcut = 1.0e-16;
  int64_t cnt = 0;
  for (int a = 0; a < n; ++a) {
    const int la = sp_l[a], ma = sp_2mj[a], ta = sp_2mt[a];
    for (int b = a + 1; b < n; ++b) {
      const int lb = sp_l[b], mb = sp_2mj[b], tb = sp_2mt[b];
      if (ma + mb < 0) continue;
      for (int c = a; c < n; ++c) {
        const int lc = sp_l[c], mc = sp_2mj[c], tc = sp_2mt[c];
        const int dmax = (c == a) ? b : n - 1;
        for (int d = c + 1; d <= dmax; ++d) {
          const int ld = sp_l[d], md = sp_2mj[d], td = sp_2mt[d];
          if (((la + lb) & 1) != ((lc + ld) & 1)) continue;
          if (ta + tb != tc + td) continue;
          if (ma + mb != mc + md) continue;
          if (!abcd) {   // counting pass: selection rules only (an upper bound)
            ++cnt;
            continue;
          }
          const double sum = uncouple(a, b, c, d) * scaling_2b;
          if (fabs(sum) > Vcut) {
            if (cnt >= capacity) return -2;
            abcd[4 * cnt] = (int16_t)(a + 1);
            abcd[4 * cnt + 1] = (int16_t)(b + 1);
            abcd[4 * cnt + 2] = (int16_t)(c + 1);
            abcd[4 * cnt + 3] = (int16_t)(d + 1);
            H2[cnt] = sum;
            ++cnt;
          }
        }
      }
    }
  }
  return cnt;
}

}  // namespace

extern "C" {

double th_clebsch_gordan(int j1, int j2, int j3, int m1, int m2, int m3) {
  return cg_lgamma(j1, j2, j3, m1, m2, m3);
}

int64_t th_mscheme(int n, const int32_t* sp_2j, const int32_t* sp_2mj, const int32_t* sp_2mt,
                   const int32_t* sp_l, const int32_t* sp_sh, int nsh, int jmax,
                   const double* Vcpd, double scaling_2b, int64_t capacity, int16_t* abcd,
                   double* H2) {
  if (n <= 0 || nsh <= 0 || jmax < 0 || (abcd && (!Vcpd || !H2))) return -1;
  int jm2 = 0;
  for (int i = 0; i < n; ++i) jm2 = sp_2j[i] > jm2 ? sp_2j[i] : jm2;
  CGTable cg;
  if (abcd) cg.build(jm2);
  // isospin CG <1/2 t1 1/2 t2 | T MT>
  auto cgt = [](int T, int t1, int t2) { return cg_racah(1, 1, 2 * T, t1, t2); };
  auto V = [&](int mmt, int J, int a, int b, int c, int d) {
    // hamil_H2cpd(0:3, 0:jmax, nsh, nsh, nsh, nsh), shells 1-based
    return Vcpd[(size_t)mmt + 4 * ((size_t)J + (size_t)(jmax + 1) *
                                   ((size_t)(a - 1) + nsh * ((size_t)(b - 1) + nsh *
                                   ((size_t)(c - 1) + nsh * (size_t)(d - 1)))))];
  };
  // uncouple_JT
  auto uncouple = [&](int a, int b, int c, int d) {
    const int ja = sp_2j[a], jb = sp_2j[b], jc = sp_2j[c], jd = sp_2j[d];
    const int ma = sp_2mj[a], mb = sp_2mj[b], mc = sp_2mj[c], md = sp_2mj[d];
    const int ta = sp_2mt[a], tb = sp_2mt[b], tc = sp_2mt[c], td = sp_2mt[d];
    const int sha = sp_sh[a], shb = sp_sh[b], shc = sp_sh[c], shd = sp_sh[d];
    const int MJ = ma + mb, MT = ta + tb;
    const double dab = (sha == shb) ? 1.0 : 0.0, dcd = (shc == shd) ? 1.0 : 0.0;
    double sum = 0.0;
    const int J0 = abs(MJ) / 2, J1 = ((ja + jb) / 2 < (jc + jd) / 2) ? (ja + jb) / 2 : (jc + jd) / 2;
    for (int J = J0; J <= J1; ++J) {
      const double cg1 = cg.at(ja, jb, J, ma, mb);
      const double cg2 = cg.at(jc, jd, J, mc, md);
      for (int T = abs(MT) / 2; T <= 1; ++T) {
        const int mmt = MT / 2 + 2 * T;
        const double v = V(mmt, J, sha, shb, shc, shd);
        if (fabs(v) < 1.0e-12) continue;
        const double ph = ((J + T) & 1) ? -1.0 : 1.0;
        const double nab = sqrt(1.0 - dab * ph) / (1.0 + dab);
        const double ncd = sqrt(1.0 - dcd * ph) / (1.0 + dcd);
        if (nab <= 1.0e-12 || ncd <= 1.0e-12) continue;
        const double ninv = 1.0 / (nab * ncd);
        sum += ninv * cg1 * cg2 * cgt(T, ta, tb) * cgt(T, tc, td) * v;
      }
    }
    return sum;
  };
  return mscheme_loop(n, sp_2mj, sp_2mt, sp_l, scaling_2b, capacity, abcd, H2, uncouple);
}

int64_t th_mscheme_j(int n, const int32_t* sp_2j, const int32_t* sp_2mj, const int32_t* sp_2mt,
                     const int32_t* sp_l, const int32_t* sp_sh, int nsh, int jmax,
                     const double* Vcpd, double scaling_2b, int64_t capacity, int16_t* abcd,
                     double* H2) {
  if (n <= 0 || nsh <= 0 || jmax < 0 || (abcd && (!Vcpd || !H2))) return -1;
  int jm2 = 0;
  for (int i = 0; i < n; ++i) jm2 = sp_2j[i] > jm2 ? sp_2j[i] : jm2;
  CGTable cg;
  if (abcd) cg.build(jm2);
  auto V = [&](int t, int J, int a, int b, int c, int d) {
    // hamil_H2cpd(0:5, 0:jmax, nsh, nsh, nsh, nsh), shells 1-based
    return Vcpd[(size_t)t + 6 * ((size_t)J + (size_t)(jmax + 1) *
                                 ((size_t)(a - 1) + nsh * ((size_t)(b - 1) + nsh *
                                 ((size_t)(c - 1) + nsh * (size_t)(d - 1)))))];
  };
  bool bad = false;
  // uncouple_J
  auto uncouple = [&](int a, int b, int c, int d) {
    const int ja = sp_2j[a], jb = sp_2j[b], jc = sp_2j[c], jd = sp_2j[d];
    const int ma = sp_2mj[a], mb = sp_2mj[b], mc = sp_2mj[c], md = sp_2mj[d];
    const int sha = sp_sh[a], shb = sp_sh[b], shc = sp_sh[c], shd = sp_sh[d];
    const int T = jscheme_channel(sp_2mt[a], sp_2mt[b], sp_2mt[c], sp_2mt[d]);
    if (T < 0) {
      bad = true;
      return 0.0;
    }
    double dab = 0.0, dcd = 0.0;
    if (T == 0 || T == 5) {
      dab = (sha == shb) ? 1.0 : 0.0;
      dcd = (shc == shd) ? 1.0 : 0.0;
    }
    double sum = 0.0;
    const int J0 = abs(ma + mb) / 2, J1 = ((ja + jb) / 2 < (jc + jd) / 2) ? (ja + jb) / 2 : (jc + jd) / 2;
    for (int J = J0; J <= J1; ++J) {
      const double v = V(T, J, sha, shb, shc, shd);
      if (fabs(v) < 1.0e-12) continue;
      const double ph = (J & 1) ? 1.0 : -1.0;   // (-1)^(J+1)
      const double nab = sqrt(1.0 - dab * ph) / (1.0 + dab);
      const double ncd = sqrt(1.0 - dcd * ph) / (1.0 + dcd);
      if (nab <= 1.0e-12 || ncd <= 1.0e-12) continue;
      const double ninv = 1.0 / (nab * ncd);
      sum += ninv * cg.at(ja, jb, J, ma, mb) * cg.at(jc, jd, J, mc, md) * v;
    }
    return sum;
  };
  const int64_t cnt = mscheme_loop(n, sp_2mj, sp_2mt, sp_l, scaling_2b, capacity, abcd, H2, uncouple);
  return bad ? -1 : cnt;
}

int th_reconstruct_2body(int htype, int nsh, int jmax, const int32_t* sh_2j, double* Vcpd) {
  if (nsh <= 0 || jmax < 0 || !sh_2j || !Vcpd || htype < 1 || htype > 4) return -1;
  const bool jt = htype <= 2;
  const int nch = jt ? 4 : 6;
  auto at = [&](int t, int J, int a, int b, int c, int d) -> double& {
    return Vcpd[(size_t)t + nch * ((size_t)J + (size_t)(jmax + 1) *
                                  ((size_t)a + nsh * ((size_t)b + nsh * ((size_t)c + nsh * (size_t)d))))];
  };
  // every element, in the reference's loop order (a outermost, the channel
  // innermost), is copied to its seven images; an image written here is read
  // again when the loop reaches it
  for (int a = 0; a < nsh; ++a)
    for (int b = 0; b < nsh; ++b)
      for (int c = 0; c < nsh; ++c)
        for (int d = 0; d < nsh; ++d)
          for (int J = 0; J <= jmax; ++J)
            for (int t = 0; t < nch; ++t) {
              const double v = at(t, J, a, b, c, d);
              if (!(fabs(v) > 1.0e-12)) continue;
              // (-1)^(ja + jb + J + x): x = 0 for the T = 0 channel of a JT
              // interaction, else 1
              const int x = (jt && t == 0) ? 0 : 1;
              const double pab = (((sh_2j[a] + sh_2j[b]) / 2 + J + x) & 1) ? -1.0 : 1.0;
              const double pcd = (((sh_2j[c] + sh_2j[d]) / 2 + J + x) & 1) ? -1.0 : 1.0;
              auto ch = [&](int perm) { return jt ? t : iso_after_perm(t, perm); };
              at(ch(1), J, a, b, d, c) = v * pcd;
              at(ch(2), J, b, a, c, d) = v * pab;
              at(ch(3), J, b, a, d, c) = v * pab * pcd;
              at(ch(4), J, c, d, a, b) = v;
              at(ch(5), J, c, d, b, a) = v * pab;
              at(ch(6), J, d, c, a, b) = v * pcd;
              at(ch(7), J, d, c, b, a) = v * pab * pcd;
            }
  return 0;
}

int th_timerev(int n, const int32_t* sp_2j, const int32_t* sp_tr, int64_t ntb,
               const int16_t* abcd, int8_t* trperm) {
  if (n <= 0 || ntb < 0 || (ntb > 0 && (!abcd || !trperm))) return -1;
  for (int64_t k = 0; k < ntb; ++k) {
    const int ia = abcd[4 * k] - 1, ib = abcd[4 * k + 1] - 1, ic = abcd[4 * k + 2] - 1,
              id = abcd[4 * k + 3] - 1;
    if (ia < 0 || ib < 0 || ic < 0 || id < 0 || ia >= n || ib >= n || ic >= n || id >= n) return -1;
    int code = 0;
    double phase = (((sp_2j[ia] + sp_2j[ib] + sp_2j[ic] + sp_2j[id]) / 2) & 1) ? -1.0 : 1.0;
    int ta = sp_tr[ia], tb = sp_tr[ib], tc = sp_tr[ic], td = sp_tr[id];
    if (ta > tb) { int t = ta; ta = tb; tb = t; phase = -phase; code += 1; }
    if (tc > td) { int t = tc; tc = td; td = t; phase = -phase; code += 2; }
    if (ta > tc || (ta == tc && tb > td)) code += 4;
    if (phase < 0.0) code -= 8;
    trperm[k] = (int8_t)code;
  }
  return 0;
}

}  // extern "C"
