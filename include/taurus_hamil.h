/*
 * taurus_hamil.h -- host C ABI of the m-scheme TBME producer (libtaurus_hamil.so).
 *
 * Produces the reduced two-body list that tg_init consumes, in the reference's
 * storage convention, from a JT-coupled (ANTOINE, hamil_type 1-2) or a J-coupled
 * (hamil_type 3-4) interaction.
 */
#ifndef TAURUS_HAMIL_H
#define TAURUS_HAMIL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* <j1 m1 j2 m2 | j3 m3>, arguments twice their values; replaces ClebschGordan
 * (src/module_mathmethods.f90:281-342). */
double th_clebsch_gordan(int j1, int j2, int j3, int m1, int m2, int m3);

/*
 * th_mscheme -- m-scheme loop of set_hamiltonian_2body with uncouple_JT
 * (src/module_hamiltonian.f90:330-378, 828-885).
 *   sp_*       single-particle quantum numbers (reference ordering, 2j, 2m, 2t, l,
 *              1-based shell index), n of them
 *   Vcpd       hamil_H2cpd(0:3, 0:jmax, nsh, nsh, nsh, nsh) in Fortran order,
 *              after the symmetry reconstruction of :284-319
 *   abcd/H2    output, 4*capacity int16 (1-based) and capacity doubles, in the
 *              reference's loop order; with abcd == NULL only the selection-rule
 *              count (an upper bound) is returned.
 * Returns the number of stored elements, -1 on bad arguments, -2 if capacity
 * is too small.
 */
int64_t th_mscheme(int n, const int32_t* sp_2j, const int32_t* sp_2mj, const int32_t* sp_2mt,
                   const int32_t* sp_l, const int32_t* sp_sh, int nsh, int jmax,
                   const double* Vcpd, double scaling_2b, int64_t capacity, int16_t* abcd,
                   double* H2);

/*
 * th_mscheme_j -- the same loop with uncouple_J (src/module_hamiltonian.f90:
 * 747-814) for J-scheme interactions: Vcpd is hamil_H2cpd(0:5, 0:jmax, nsh, nsh,
 * nsh, nsh), channels 0 pp pp, 1 pn pn, 2 pn np, 3 np pn, 4 np np, 5 nn nn, after
 * the symmetry reconstruction.  Arguments and returns as th_mscheme.
 */
int64_t th_mscheme_j(int n, const int32_t* sp_2j, const int32_t* sp_2mj, const int32_t* sp_2mt,
                     const int32_t* sp_l, const int32_t* sp_sh, int nsh, int jmax,
                     const double* Vcpd, double scaling_2b, int64_t capacity, int16_t* abcd,
                     double* H2);

/*
 * th_reconstruct_2body -- the symmetry reconstruction of set_hamiltonian_2body
 * (src/module_hamiltonian.f90:284-319) with find_iso (:911-959), in place on
 * hamil_H2cpd(0:nch-1, 0:jmax, nsh, nsh, nsh, nsh) (Fortran order; nch = 4 for
 * htype 1-2, 6 for htype 3-4), in the reference's loop order.  sh_2j: 2j of
 * each shell.  Returns 0, or -1 on bad arguments.
 */
int th_reconstruct_2body(int htype, int nsh, int jmax, const int32_t* sh_2j, double* Vcpd);

/*
 * th_timerev -- reconstruct_2body_timerev (src/module_hamiltonian.f90:970-1018):
 * int8 permutation code (bits 1 swap ab, 2 swap cd, 4 swap bra/ket; -8 phase -1).
 * sp_tr is 1-based as HOsp_tr.  Returns 0, or -1 on bad arguments.
 */
int th_timerev(int n, const int32_t* sp_2j, const int32_t* sp_tr, int64_t ntb,
               const int16_t* abcd, int8_t* trperm);

#ifdef __cplusplus
}
#endif

#endif
