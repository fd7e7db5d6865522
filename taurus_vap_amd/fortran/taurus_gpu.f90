!==============================================================================!
! MODULE TaurusGPU                                                             !
!                                                                              !
! iso_c_binding interface of libtaurus_gpu.so (include/taurus_gpu.h) and the   !
! device context of a TAURUS_vap run.  The context is created on first use     !
! from the module arrays of the run (HOsp_*, hamil_*), so the Hamiltonian is   !
! uploaded once; TBMEs stay FP32 on the device when rH2 = real32.              !
!                                                                              !
! Environment: TAURUS_GPU_PLAIN=1 selects one gauge angle per launch sequence. !
!==============================================================================!
MODULE TaurusGPU

use iso_c_binding
use Constants
use Basis
use Hamiltonian

implicit none
public

integer(c_int), parameter :: TG_OK = 0, TG_E_IMAG_GRADIENT = -5, TG_E_NOMEM = -3, TG_E_UNSUPPORTED = -7, &
                             TG_FLAG_PLAIN = 1, TG_FLAG_TBME_FP32 = 2, &
                             TG_FLAG_GRAPH = 4, TG_FLAG_CONJ_PAIRS = 8
type(c_ptr) :: taurus_gpu_ctx = c_null_ptr
integer(c_int) :: taurus_gpu_flags = 0

interface
  function tg_last_error() bind(C, name='tg_last_error') result(p)
    import
    type(c_ptr) :: p
  end function

  integer(c_int) function tg_init(ctx, ndim, sp_2mj, sp_2mt, sp_l, sp_tr, H0_re, H0_im, &
                                  H1, ntb, abcd, H2, h2_bytes, trperm, flags) &
                                  bind(C, name='tg_init')
    import
    type(c_ptr), intent(out) :: ctx
    integer(c_int), value :: ndim
    integer(c_int32_t), intent(in) :: sp_2mj(*), sp_2mt(*), sp_l(*), sp_tr(*)
    real(c_double), value :: H0_re, H0_im
    complex(c_double_complex), intent(in) :: H1(*)
    integer(c_int64_t), value :: ntb
    integer(c_int16_t), intent(in) :: abcd(*)
    type(*), intent(in) :: H2(*)
    integer(c_int), value :: h2_bytes
    integer(c_int8_t), intent(in) :: trperm(*)
    integer(c_int), value :: flags
  end function

  integer(c_int) function tg_finalize(ctx) bind(C, name='tg_finalize')
    import
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function tg_calculate_fields(ctx, rhoLR, kappaLR, kappaRL, gammaLR, &
                                              hspLR, deltaLR, deltaRL) &
                                              bind(C, name='tg_calculate_fields')
    import
    type(c_ptr), value :: ctx
    complex(c_double_complex), intent(in) :: rhoLR(*), kappaLR(*), kappaRL(*)
    complex(c_double_complex), intent(out) :: gammaLR(*), hspLR(*), deltaLR(*), deltaRL(*)
  end function

  integer(c_int) function tg_calculate_fields_diag(ctx, rhoLR, kappaLR, gammaLR, hspLR, &
                                                   deltaLR, deltaRL) &
                                                   bind(C, name='tg_calculate_fields_diag')
    import
    type(c_ptr), value :: ctx
    complex(c_double_complex), intent(in) :: rhoLR(*), kappaLR(*)
    complex(c_double_complex), intent(out) :: gammaLR(*), hspLR(*), deltaLR(*), deltaRL(*)
  end function

  integer(c_int) function tg_zskpfa(n, count, A, lda, pfaff, info) bind(C, name='tg_zskpfa')
    import
    integer(c_int), value :: n, count, lda
    complex(c_double_complex), intent(in) :: A(*)
    complex(c_double_complex), intent(out) :: pfaff(*)
    integer(c_int), intent(out) :: info(*)
  end function

  integer(c_int) function tg_calculate_overlap(ndim, nocc0, nemp0, ovac0, voveru0, D0, &
                                               nocc1, nemp1, ovac1, voveru1, D1, ROT, &
                                               overlap) bind(C, name='tg_calculate_overlap')
    import
    integer(c_int), value :: ndim, nocc0, nemp0, nocc1, nemp1
    real(c_double), value :: ovac0, ovac1
    real(c_double), intent(in) :: voveru0(*), voveru1(*)
    complex(c_double_complex), intent(in) :: D0(*), D1(*), ROT(*)
    complex(c_double_complex), intent(out) :: overlap
  end function

  integer(c_int) function tg_project_wavefunction(ctx, iprint, U0, V0, D0, nocc0, nemp0, &
                                                  ovac0, voveru0, Mphip, Mphin, facpi, &
                                                  valence_Z, valence_N, factor_delta, &
                                                  pnp_scalars, pnp_A, pnp_H20, pnp_AH20, &
                                                  field_H20v, imag_grad) &
                                                  bind(C, name='tg_project_wavefunction')
    import
    type(c_ptr), value :: ctx
    integer(c_int), value :: iprint, nocc0, nemp0, Mphip, Mphin
    real(c_double), intent(in) :: U0(*), V0(*), D0(*), voveru0(*)
    real(c_double), value :: ovac0, facpi, valence_Z, valence_N, factor_delta
    complex(c_double_complex), intent(out) :: pnp_scalars(*), pnp_A(*), pnp_H20(*), pnp_AH20(*)
    real(c_double), intent(out) :: field_H20v(*), imag_grad
  end function

  integer(c_int) function tg_calculate_gradient(ndim, gradient_type, gradient_eta, &
                                                gradient_mu, info_H11, info_hsp, eigen_H11, &
                                                eigen_hsp, field_H20v, constraint_dim, &
                                                constraint_20, lagrange_lambda1, &
                                                gradient_Zim1, gradient_Zi, gradient_norm) &
                                                bind(C, name='tg_calculate_gradient')
    import
    integer(c_int), value :: ndim, gradient_type, info_H11, info_hsp, constraint_dim
    real(c_double), value :: gradient_eta, gradient_mu
    real(c_double), intent(in) :: eigen_H11(*), eigen_hsp(*), field_H20v(*), &
                                  constraint_20(*), lagrange_lambda1(*), gradient_Zim1(*)
    real(c_double), intent(out) :: gradient_Zi(*), gradient_norm
  end function

  ! whole iterations on the device (loop body of src/taurus_vap.f90:77-113,
  ! gradient_type 0): constraints once, then per iteration
  ! tg_set_state -> tg_iterate -> tg_get_state
  integer(c_int) function tg_set_state(ctx, U0, V0, D0, nocc0, nemp0, ovac0, voveru0, &
                                       Mphip, Mphin, facpi, valence_Z, valence_N, &
                                       factor_delta, iprint) bind(C, name='tg_set_state')
    import
    type(c_ptr), value :: ctx
    integer(c_int), value :: nocc0, nemp0, Mphip, Mphin, iprint
    real(c_double), intent(in) :: U0(*), V0(*), D0(*), voveru0(*)
    real(c_double), value :: ovac0, facpi, valence_Z, valence_N, factor_delta
  end function

  integer(c_int) function tg_set_gradient(ctx, gradient_eta, gradient_mu, constraint_dim, &
                                          constraint_20, lagrange_lambda1) &
                                          bind(C, name='tg_set_gradient')
    import
    type(c_ptr), value :: ctx
    real(c_double), value :: gradient_eta, gradient_mu
    integer(c_int), value :: constraint_dim
    real(c_double), intent(in) :: constraint_20(*), lagrange_lambda1(*)
  end function

  integer(c_int) function tg_set_constraints(ctx, constraint_dim, constraint_pair, &
                                             constraint_HO, nsets, constraint_val, &
                                             constraint_eps) bind(C, name='tg_set_constraints')
    import
    type(c_ptr), value :: ctx
    integer(c_int), value :: constraint_dim, constraint_pair, nsets
    real(c_double), intent(in) :: constraint_HO(*), constraint_val(*)
    real(c_double), value :: constraint_eps
  end function

  integer(c_int) function tg_iterate(ctx, gradient_eps, pnp_scalars, gradient_norm, &
                                     lagrange_lambda1, adjust_iterations, converged) &
                                     bind(C, name='tg_iterate')
    import
    type(c_ptr), value :: ctx
    real(c_double), value :: gradient_eps
    complex(c_double_complex), intent(out) :: pnp_scalars(*)
    real(c_double), intent(out) :: gradient_norm(*), lagrange_lambda1(*)
    integer(c_int), intent(out) :: adjust_iterations(*), converged(*)
  end function

  integer(c_int) function tg_get_state(ctx, U0, V0) bind(C, name='tg_get_state')
    import
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: U0(*), V0(*)
  end function

  ! construct_canonical_basis of the resident states on the device
  integer(c_int) function tg_canonical_basis(ctx, eps, nocc, nemp, ovacc, voveru) &
                                             bind(C, name='tg_canonical_basis')
    import
    type(c_ptr), value :: ctx
    real(c_double), value :: eps
    integer(c_int), intent(out) :: nocc(*), nemp(*)
    real(c_double), intent(out) :: ovacc(*), voveru(*)
  end function

  ! tg_canonical_basis then tg_iterate in one call (the canonical basis beside the projection)
  integer(c_int) function tg_iterate_canonical(ctx, canon_eps, gradient_eps, pnp_scalars, &
                                               gradient_norm, lagrange_lambda1,              &
                                               adjust_iterations, converged, nocc, nemp,      &
                                               ovacc, voveru) bind(C, name='tg_iterate_canonical')
    import
    type(c_ptr), value :: ctx
    real(c_double), value :: canon_eps, gradient_eps
    complex(c_double_complex), intent(out) :: pnp_scalars(*)
    real(c_double), intent(out) :: gradient_norm(*), lagrange_lambda1(*)
    integer(c_int), intent(out) :: adjust_iterations(*), converged(*), nocc(*), nemp(*)
    real(c_double), intent(out) :: ovacc(*), voveru(*)
  end function

  ! diagonalize_hsp_and_H11(0, ndim) of the resident state
  integer(c_int) function tg_diagonalize_hsp_and_H11(ctx, constraint_dim, lagrange_lambda1, &
                                                     eigen_hsp, eigen_H11, info_hsp, info_H11, &
                                                     hsp_RR, field_H11) &
                                                     bind(C, name='tg_diagonalize_hsp_and_H11')
    import
    type(c_ptr), value :: ctx
    integer(c_int), value :: constraint_dim
    real(c_double), intent(in) :: lagrange_lambda1(*)
    real(c_double), intent(out) :: eigen_hsp(*), eigen_H11(*), hsp_RR(*), field_H11(*)
    integer(c_int), intent(out) :: info_hsp, info_H11
  end function

  ! per-angle observables of project_wavefunction(1, ...) (final printing)
  integer(c_int) function tg_project_observables(ctx, nobo, obo_ops, nobos, obos_ops, &
                                                 nshell, sp_shell, per_angle, pnp_rho) &
                                                 bind(C, name='tg_project_observables')
    import
    type(c_ptr), value :: ctx
    integer(c_int), value :: nobo, nobos, nshell
    real(c_double), intent(in) :: obo_ops(*), obos_ops(*)
    integer(c_int), intent(in) :: sp_shell(*)
    complex(c_double_complex), intent(out) :: per_angle(*), pnp_rho(*)
  end function
end interface

CONTAINS

!------------------------------------------------------------------------------!
! Message of the last failing library call                                     !
!------------------------------------------------------------------------------!
function taurus_gpu_error() result(msg)
  character(len=:), allocatable :: msg
  character(kind=c_char), pointer :: cs(:)
  type(c_ptr) :: p
  integer :: i
  p = tg_last_error()
  call c_f_pointer(p, cs, [512])
  i = 0
  do while ( i < 512 )
    if ( cs(i+1) == c_null_char ) exit
    i = i + 1
  enddo
  allocate(character(len=i) :: msg)
  msg = transfer(cs(1:i), msg)
end function taurus_gpu_error

!------------------------------------------------------------------------------!
! Creates the device context from the basis and the Hamiltonian of the run     !
! (once; the TBMEs are read once into device memory).                          !
!------------------------------------------------------------------------------!
subroutine taurus_gpu_setup(ndim)
  integer, intent(in) :: ndim
  integer(c_int) :: status, flags, h2b
  character(len=8) :: env
  integer :: lenv, st
  if ( c_associated(taurus_gpu_ctx) ) return
  flags = 0
  call get_environment_variable('TAURUS_GPU_PLAIN', env, lenv, st)
  if ( st == 0 .and. lenv > 0 ) then
    if ( env(1:1) == '1' ) flags = ior(flags, TG_FLAG_PLAIN)
  endif
  call get_environment_variable('TAURUS_GPU_GRAPH', env, lenv, st)
  if ( st == 0 .and. lenv > 0 ) then
    if ( env(1:1) == '1' ) flags = ior(flags, TG_FLAG_GRAPH)
  endif
  call get_environment_variable('TAURUS_GPU_CONJ_PAIRS', env, lenv, st)
  if ( st == 0 .and. lenv > 0 ) then
    if ( env(1:1) == '1' ) flags = ior(flags, TG_FLAG_CONJ_PAIRS)
  endif
  h2b = int(storage_size(hamil_H2) / 8, c_int)
  if ( h2b == 4 ) flags = ior(flags, TG_FLAG_TBME_FP32)
  taurus_gpu_flags = flags
  status = tg_init(taurus_gpu_ctx, int(ndim,c_int), int(HOsp_2mj,c_int32_t), &
                   int(HOsp_2mt,c_int32_t), int(HOsp_l,c_int32_t), int(HOsp_tr,c_int32_t), &
                   real(hamil_H0,c_double), aimag(hamil_H0), hamil_H1, &
                   int(hamil_H2dim,c_int64_t), hamil_abcd, hamil_H2, h2b, hamil_trperm, flags)
  if ( status /= TG_OK ) then
    print*, 'tg_init failed: ', status, ' ', taurus_gpu_error()
    stop
  endif
end subroutine taurus_gpu_setup

END MODULE TaurusGPU
!==============================================================================!
! End of file                                                                  !
!==============================================================================!
