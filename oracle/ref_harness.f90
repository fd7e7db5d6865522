!==============================================================================!
! ref_harness.f90 -- test infrastructure only (oracle side).                   !
!                                                                              !
! A driver written for this repository that links the reference TAURUS_vap     !
! modules (compiled from their own sources by oracle/build_ref.sh) and dumps   !
! the intermediate quantities of the projected-gradient iteration so that the  !
! restatement in oracle/ and the GPU library can be checked stage by stage.    !
!                                                                              !
! Modes (first command-line argument):                                         !
!   iter  <dumpfile>        reads the STDIN deck like taurus_vap.f90:40-65,    !
!                           runs the iteration loop of taurus_vap.f90:77-123   !
!                           and dumps, for iteration 1, every per-angle stage  !
!                           of project_wavefunction (module_projection.f90:    !
!                           242-359) plus the per-iteration energies.          !
!   pfaff <infile> <outfile> Pfaffians of skew matrices with ZSKPFA('U','P')   !
!                           exactly as calculate_overlap calls it              !
!                           (module_projection.f90:921).                       !
!   over  <infile> <outfile> calculate_overlap on given canonical data         !
!   time  <outfile> <K>     reads the deck like iter (no initial constraint    !
!                           adjustment), then K times: project_wavefunction    !
!                           (0,...) at 1x2 and at 3x3 gauge angles and the     !
!                           gradient step (calculate_lagrange_multipliers +    !
!                           calculate_gradient); writes the seconds as text.   !
!   timefull <outfile> <K>  as time, but K full iterations at the deck's own   !
!                           gauge angles: project_wavefunction(0,...), the     !
!                           gradient step and adjust_constraints, each timed.  !
!   proj  <dumpfile> <K>    as iter for at most K iterations, dumping only     !
!                           what stays small at large bases: the state of      !
!                           iteration 1, the canonical-basis scalars, per-     !
!                           angle scalars, the gauge sums, field_H20v, the     !
!                           multipliers, the gradient and (gradient_type > 0)  !
!                           the h and H11 matrices of diagonalize_hsp_and_H11  !
!                           with their LAPACK eigenvalues; then histories.     !
!                                                                              !
! Dump format (stream, little endian): repeated records                        !
!   char(32) name | int32 kind | int32 rank | int64 dims(rank) | data          !
! kinds: 1 int32, 2 real64, 3 complex128, 4 int64, 5 int8, 6 int16             !
!==============================================================================!
module harness_dump
use Constants
implicit none
integer, parameter :: ud = 97
contains
subroutine hdr(name, kind, rank, dims)
  character(*), intent(in) :: name
  integer, intent(in) :: kind, rank
  integer(i64), intent(in) :: dims(:)
  character(32) :: nm
  nm = name
  write(ud) nm, int(kind,i32), int(rank,i32)
  if (rank > 0) write(ud) dims(1:rank)
end subroutine
subroutine dump_i0(name, v)
  character(*), intent(in) :: name
  integer, intent(in) :: v
  integer(i64) :: d(1)
  d = 0
  call hdr(name, 1, 0, d)
  write(ud) int(v,i32)
end subroutine
subroutine dump_i1(name, v, n)
  character(*), intent(in) :: name
  integer, intent(in) :: n
  integer, intent(in) :: v(n)
  call hdr(name, 1, 1, [int(n,i64)])
  write(ud) int(v,i32)
end subroutine
subroutine dump_r0(name, v)
  character(*), intent(in) :: name
  real(r64), intent(in) :: v
  integer(i64) :: d(1)
  d = 0
  call hdr(name, 2, 0, d)
  write(ud) v
end subroutine
subroutine dump_r1(name, v, n)
  character(*), intent(in) :: name
  integer(i64), intent(in) :: n
  real(r64), intent(in) :: v(n)
  call hdr(name, 2, 1, [n])
  write(ud) v
end subroutine
subroutine dump_r2(name, v, n1, n2)
  character(*), intent(in) :: name
  integer, intent(in) :: n1, n2
  real(r64), intent(in) :: v(n1,n2)
  call hdr(name, 2, 2, [int(n1,i64), int(n2,i64)])
  write(ud) v
end subroutine
subroutine dump_c0(name, v)
  character(*), intent(in) :: name
  complex(r64), intent(in) :: v
  integer(i64) :: d(1)
  d = 0
  call hdr(name, 3, 0, d)
  write(ud) v
end subroutine
subroutine dump_c2(name, v, n1, n2)
  character(*), intent(in) :: name
  integer, intent(in) :: n1, n2
  complex(r64), intent(in) :: v(n1,n2)
  call hdr(name, 3, 2, [int(n1,i64), int(n2,i64)])
  write(ud) v
end subroutine
end module harness_dump

program ref_harness

use Nucleus
use Basis
use Fields
use Operators
use Projection
use Gradient
use Initialization
use harness_dump

implicit none

character(16) :: mode
character(512) :: f1, f2

call get_command_argument(1, mode)
call get_command_argument(2, f1)
call get_command_argument(3, f2)

select case (trim(mode))
  case ('iter')
    call run_iter(trim(f1))
  case ('pfaff')
    call run_pfaff(trim(f1), trim(f2))
  case ('over')
    call run_over(trim(f1), trim(f2))
  case ('time')
    call run_time(trim(f1), trim(f2))
  case ('timefull')
    call run_timefull(trim(f1), trim(f2))
  case ('proj')
    call run_proj(trim(f1), trim(f2))
  case default
    stop 'ref_harness: mode must be iter, proj, pfaff, over, time or timefull'
end select

contains

!------------------------------------------------------------------------------!
! Mirror of taurus_vap.f90 with stage dumps at iteration 1                     !
!------------------------------------------------------------------------------!
subroutine run_iter(dumpfile)
character(*), intent(in) :: dumpfile
integer :: iter_max, iter_write, iter_print, iter, ndim, niter_done
real(r64), allocatable :: e_hist(:), gn_hist(:), z_hist(:), n_hist(:)

call read_input(iter_max,iter_write,iter_print)
call set_nucleus
call set_basis
call set_hamiltonian
call set_wavefunctions
call set_fields
call set_particle_number
call set_pairs
call set_angular_momentum
call set_multipoles
call set_radius
call set_projection
call generate_wavefunction(HOsp_dim)
call check_symmetries(HOsp_dim)
call set_constraints
call set_gradient
if ( (iter_max > 0) .or. (seed_type /= 1) ) then
  call adjust_constraints(gradient_Zi,HOsp_dim)
endif

ndim = HOsp_dim
open(ud, file=dumpfile, status='replace', access='stream', form='unformatted')

call dump_i0('ndim', ndim)
call dump_i1('sp_2mj', HOsp_2mj, ndim)
call dump_i1('sp_2mt', HOsp_2mt, ndim)
call dump_i1('sp_l', HOsp_l, ndim)
call dump_i1('sp_2j', HOsp_2j, ndim)
call dump_i1('sp_tr', HOsp_tr, ndim)
call dump_r0('valence_Z', valence_Z)
call dump_r0('valence_N', valence_N)
call dump_i0('Mphip', proj_Mphip)
call dump_i0('Mphin', proj_Mphin)
call dump_r0('facpi', proj_facpi)
call dump_i0('gradient_type', gradient_type)
call dump_r0('gradient_eta', gradient_eta)
call dump_r0('gradient_mu', gradient_mu)
call dump_i0('constraint_dim', constraint_dim)
call dump_i0('constraint_pair', constraint_pair)
call dump_r0('factor_delta', factor_delta)
call dump_c0('H0', hamil_H0)
call dump_c2('H1', hamil_H1, ndim, ndim)
call dump_r2('U0_start', bogo_U0, ndim, ndim)
call dump_r2('V0_start', bogo_V0, ndim, ndim)
if ( constraint_dim > 0 ) then
  call dump_r1('constraint_HO', constraint_HO, int(constraint_dim,i64)*ndim*ndim)
  call dump_r1('constraint_val', constraint_val, int(constraint_dim,i64))
endif

allocate(e_hist(iter_max), gn_hist(iter_max), z_hist(iter_max), n_hist(iter_max))
e_hist = zero
gn_hist = zero
z_hist = zero
n_hist = zero
niter_done = 0

do iter = 1, iter_max
  field_H20v = zero
  call calculate_densities_real(bogo_U0,bogo_V0,dens_rhoRR,dens_kappaRR, &
                                HOsp_dim)
  call update_constraints_qpbasis

  if ( iter == 1 ) then
    call dump_r2('U0', bogo_U0, ndim, ndim)
    call dump_r2('V0', bogo_V0, ndim, ndim)
    call dump_r2('rhoRR', dens_rhoRR, ndim, ndim)
    call dump_r2('kappaRR', dens_kappaRR, ndim, ndim)
    if ( constraint_dim > 0 ) then
      call dump_r1('constraint_20', constraint_20, int(constraint_dim,i64)*ndim*ndim)
    endif
    call dump_angles(ndim, .true.)
  endif

  call project_wavefunction(0,iter_print,proj_Mphip,proj_Mphin,HOsp_dim)
  call calculate_lagrange_multipliers(field_H20v,HOsp_dim)

  if ( iter == 1 ) then
    call dump_c0('pnp_over', pnp_over)
    call dump_c0('pnp_ener', pnp_ener)
    call dump_c0('pnp_prot', pnp_prot)
    call dump_c0('pnp_neut', pnp_neut)
    call dump_c0('pnp_prot2', pnp_prot2)
    call dump_c0('pnp_neut2', pnp_neut2)
    call dump_c0('pnp_pari', pnp_pari)
    call dump_i0('iter_print', iter_print)
    call dump_c2('pnp_A', pnp_A, ndim, ndim)
    call dump_c2('pnp_H20', pnp_H20, ndim, ndim)
    call dump_c2('pnp_AH20', pnp_AH20, ndim, ndim)
    call dump_r1('field_H20v', field_H20v, int(ndim,i64)*ndim)
    if ( constraint_dim > 0 ) then
      call dump_r1('lagrange_lambda1', lagrange_lambda1, int(constraint_dim,i64))
    endif
  endif

  if ( gradient_type > 0 ) call diagonalize_hsp_and_H11(0,HOsp_dim)
  call calculate_gradient(HOsp_dim)

  if ( iter == 1 ) then
    call dump_r2('field_hspRR_eig', field_hspRR, ndim, ndim)
    call dump_r1('gradient_Zi', gradient_Zi, int(ndim,i64)*ndim)
    call dump_r0('gradient_norm', gradient_norm)
  endif

  call print_iteration(iter_print,iter)
  e_hist(iter) = real(pnp_ener/pnp_over)
  z_hist(iter) = real(pnp_prot/pnp_over)
  n_hist(iter) = real(pnp_neut/pnp_over)
  gn_hist(iter) = gradient_norm
  niter_done = iter

  if ( gradient_norm <= gradient_eps ) exit
  call adjust_constraints(gradient_Zi,HOsp_dim)
  gradient_Zim1 = gradient_Zi
enddo

call dump_i0('niter', niter_done)
call dump_r1('hist_ener', e_hist, int(max(niter_done,1),i64))
call dump_r1('hist_prot', z_hist, int(max(niter_done,1),i64))
call dump_r1('hist_neut', n_hist, int(max(niter_done,1),i64))
call dump_r1('hist_gradnorm', gn_hist, int(max(niter_done,1),i64))
call dump_r2('U0_final', bogo_U0, ndim, ndim)
call dump_r2('V0_final', bogo_V0, ndim, ndim)
close(ud)

end subroutine run_iter

!------------------------------------------------------------------------------!
! Same body as the angle loop of project_wavefunction (iopt=0, iprint=0),      !
! built from the module's public routines, dumping each stage.                 !
!------------------------------------------------------------------------------!
subroutine dump_angles(ndim, full)
integer, intent(in) :: ndim
logical, intent(in) :: full
integer :: i, j, nocc0, nemp0, nangle
real(r64) :: ovac0
real(r64), allocatable :: voveru0(:)
complex(r64), dimension(ndim,ndim) :: zU0bar, zV0bar, zU0tilde, zV0tilde, ROTG

call reset_pnp(0)
bogo_zU0 = zone * bogo_U0
bogo_zV0 = zone * bogo_V0
call construct_canonical_basis(bogo_U0,bogo_V0,bogo_zU0c,bogo_zV0c,bogo_zD0, &
                               ovac0,nocc0,nemp0,ndim)
if ( ndim-nocc0-nemp0 == 0 ) ovac0 = 0.0d0
allocate(voveru0(ndim-nocc0))
voveru0 = zero
j = 1
do i = 1+nemp0, ndim-nocc0
  if ( (-1)**j == -1 ) then
    voveru0(i) = real( bogo_zV0c(i,i+1) / bogo_zU0c(i,i) )
  else
    voveru0(i) = real( bogo_zV0c(i,i-1) / bogo_zU0c(i,i) )
  endif
  j = j + 1
enddo

call dump_i0('nocc0', nocc0)
call dump_i0('nemp0', nemp0)
call dump_r0('ovac0', ovac0)
call dump_r1('voveru0', voveru0, int(ndim-nocc0,i64))
if ( full ) then
  call dump_c2('zD0', bogo_zD0, ndim, ndim)
  call dump_c2('zU0c', bogo_zU0c, ndim, ndim)
  call dump_c2('zV0c', bogo_zV0c, ndim, ndim)
endif

call assign_angles(proj_Mphip,proj_Mphin)
call dump_i0('nangles', proj_myangles)
do nangle = 1, proj_myangles
  call dump_r0('phip', proj_phip(nangle))
  call dump_r0('phin', proj_phin(nangle))
  call dump_c0('weip', proj_weip(nangle))
  call dump_c0('wein', proj_wein(nangle))
  call generate_rotation_gauge(proj_phip(nangle),proj_phin(nangle),ROT,ROTG,ndim)
  call rotate_wavefunction(ROTG,bogo_zU0,bogo_zV0,zU0bar,zV0bar,ndim)
  if ( max(proj_Mphip,proj_Mphin) > 1 ) then
    call calculate_overlap(nocc0,nemp0,ovac0,voveru0,bogo_zD0, &
                           nocc0,nemp0,ovac0,voveru0,bogo_zD0, &
                           ROTG,rot_over,ndim)
  else
    rot_over = zone
  endif
  call calculate_thouless(bogo_zU0,bogo_zV0,zU0bar,zV0bar,rot_A, &
                          zU0tilde,zV0tilde,ndim)
  call calculate_densities(bogo_zU0,bogo_zV0,conjg(zU0tilde), &
                           conjg(zV0tilde),dens_rhoLR,dens_kappaLR, &
                           dens_kappaRL,ndim)
  if ( max(proj_Mphip,proj_Mphin) > 1 ) then
    call calculate_fields(dens_rhoLR,dens_kappaLR,dens_kappaRL, &
                          field_gammaLR,field_hspLR,field_deltaLR, &
                          field_deltaRL,ndim)
  else
    call calculate_fields_diag(dens_rhoLR,dens_kappaLR, &
                               field_gammaLR,field_hspLR,field_deltaLR, &
                               field_deltaRL,ndim)
  endif
  call calculate_particle_number(0,dens_rhoLR,dens_kappaLR,dens_kappaRL, &
                                 rot_prot,rot_neut,rot_prot2,rot_neut2,ndim)
  call calculate_expectval_energy(dens_rhoLR,dens_kappaRL,field_gammaLR, &
                                  field_deltaLR,rot_ener,ndim)
  call calculate_H20(zU0tilde,zV0tilde,rot_H20,ndim)

  call dump_c0('rot_over', rot_over)
  call dump_c0('rot_ener', rot_ener)
  call dump_c0('rot_prot', rot_prot)
  call dump_c0('rot_neut', rot_neut)
  call dump_c0('rot_prot2', rot_prot2)
  call dump_c0('rot_neut2', rot_neut2)
  if ( .not. full ) then
    call sum_gauge(0,proj_weip(nangle),proj_wein(nangle))
    cycle
  endif
  call dump_c2('rot_A', rot_A, ndim, ndim)
  call dump_c2('Utilde', zU0tilde, ndim, ndim)
  call dump_c2('Vtilde', zV0tilde, ndim, ndim)
  call dump_c2('rhoLR', dens_rhoLR, ndim, ndim)
  call dump_c2('kappaLR', dens_kappaLR, ndim, ndim)
  call dump_c2('kappaRL', dens_kappaRL, ndim, ndim)
  call dump_c2('gammaLR', field_gammaLR, ndim, ndim)
  call dump_c2('hspLR', field_hspLR, ndim, ndim)
  call dump_c2('deltaLR', field_deltaLR, ndim, ndim)
  call dump_c2('deltaRL', field_deltaRL, ndim, ndim)
  call dump_c2('rot_H20', rot_H20, ndim, ndim)
  call sum_gauge(0,proj_weip(nangle),proj_wein(nangle))
enddo
deallocate(proj_phip,proj_weip,proj_phin,proj_wein)
deallocate(voveru0)

end subroutine dump_angles

!------------------------------------------------------------------------------!
! Pfaffians: infile = int32 count, then per matrix: int32 n, complex128 A(n,n) !
!------------------------------------------------------------------------------!
subroutine run_pfaff(infile, outfile)
character(*), intent(in) :: infile, outfile
integer(i32) :: cnt, k, n
integer :: info
complex(r64) :: pf
complex(r64), allocatable :: A(:,:), zwork(:)
integer, allocatable :: iwork(:)
real(r64), allocatable :: rwork(:)

open(98, file=infile, status='old', access='stream', form='unformatted')
open(ud, file=outfile, status='replace', access='stream', form='unformatted')
read(98) cnt
do k = 1, cnt
  read(98) n
  allocate(A(n,n), zwork(max(1,n*n)), iwork(max(1,n)), rwork(max(1,n)))
  read(98) A
  call zskpfa('u','p',n,A,n,pf,iwork,zwork,max(1,n*n),rwork,info)
  write(ud) pf, int(info,i32)
  deallocate(A, zwork, iwork, rwork)
enddo
close(98)
close(ud)
end subroutine run_pfaff

!------------------------------------------------------------------------------!
! Overlaps: infile = int32 count, int32 ndim, then per case: int32 nocc, nemp, !
! real64 ovac, real64 voveru(ndim-nocc), complex128 D(ndim,ndim),              !
! complex128 ROT(ndim,ndim)   ->  complex128 overlap per case                  !
!------------------------------------------------------------------------------!
subroutine run_over(infile, outfile)
character(*), intent(in) :: infile, outfile
integer(i32) :: cnt, k, n, nocc, nemp
real(r64) :: ovac
real(r64), allocatable :: voveru(:)
complex(r64), allocatable :: D(:,:), R(:,:)
complex(r64) :: over

open(98, file=infile, status='old', access='stream', form='unformatted')
open(ud, file=outfile, status='replace', access='stream', form='unformatted')
read(98) cnt, n
allocate(D(n,n), R(n,n))
do k = 1, cnt
  read(98) nocc, nemp, ovac
  allocate(voveru(n-nocc))
  read(98) voveru
  read(98) D
  read(98) R
  call calculate_overlap(nocc,nemp,ovac,voveru,D,nocc,nemp,ovac,voveru,D, &
                         R,over,n)
  write(ud) over
  deallocate(voveru)
enddo
close(98)
close(ud)
end subroutine run_over

!------------------------------------------------------------------------------!
! Wall-clock timing of the reference's projected-gradient step                 !
!------------------------------------------------------------------------------!
subroutine run_time(outfile, kstr)
character(*), intent(in) :: outfile, kstr
integer :: iter_max, iter_write, iter_print, k, kmax, isamp, mp, mn
integer(i64) :: c0, c1, c2, rate
real(r64) :: tproj, tgrad

read(kstr,*) kmax
call read_input(iter_max,iter_write,iter_print)
call set_nucleus
call set_basis
call set_hamiltonian
call set_wavefunctions
call set_fields
call set_particle_number
call set_pairs
call set_angular_momentum
call set_multipoles
call set_radius
call set_projection
call generate_wavefunction(HOsp_dim)
call check_symmetries(HOsp_dim)
call set_constraints
call set_gradient

open(ud, file=outfile, status='replace', form='formatted')
do k = 1, kmax
  do isamp = 1, 2
    ! samples of 1x2 and 3x3 gauge angles (a wide spread of the angle count
    ! keeps the per-angle slope clear of host timing noise)
    mp = 1 + 2 * (isamp - 1)
    mn = 2 + (isamp - 1)
    call system_clock(c0, rate)
    field_H20v = zero
    call calculate_densities_real(bogo_U0,bogo_V0,dens_rhoRR,dens_kappaRR,HOsp_dim)
    call update_constraints_qpbasis
    call project_wavefunction(0,0,mp,mn,HOsp_dim)
    call system_clock(c1)
    call calculate_lagrange_multipliers(field_H20v,HOsp_dim)
    if ( gradient_type > 0 ) call diagonalize_hsp_and_H11(0,HOsp_dim)
    call calculate_gradient(HOsp_dim)
    call system_clock(c2)
    tproj = real(c1-c0,r64) / real(rate,r64)
    tgrad = real(c2-c1,r64) / real(rate,r64)
    write(ud,'(i4,i6,2es16.8,es20.10)') k, mp*mn, tproj, tgrad, real(pnp_ener/pnp_over)
  enddo
enddo
close(ud)
end subroutine run_time

!------------------------------------------------------------------------------!
! Iterations of taurus_vap.f90:77-113 with dumps that stay small at large n    !
!------------------------------------------------------------------------------!
subroutine run_proj(dumpfile, kstr)
character(*), intent(in) :: dumpfile, kstr
integer :: iter_max, iter_write, iter_print, iter, ndim, niter_done, kmax, m, k, l, i, j, &
           info
real(r64) :: q11_aux, q00_aux
real(r64), allocatable :: e_hist(:), gn_hist(:), z_hist(:), n_hist(:), hmat(:,:), &
                          evals(:), work(:)
complex(r64), allocatable :: gRR(:,:), hRR(:,:), dRR(:,:)

read(kstr,*) kmax
call read_input(iter_max,iter_write,iter_print)
call set_nucleus
call set_basis
call set_hamiltonian
call set_wavefunctions
call set_fields
call set_particle_number
call set_pairs
call set_angular_momentum
call set_multipoles
call set_radius
call set_projection
call generate_wavefunction(HOsp_dim)
call check_symmetries(HOsp_dim)
call set_constraints
call set_gradient
if ( (iter_max > 0) .or. (seed_type /= 1) ) then
  call adjust_constraints(gradient_Zi,HOsp_dim)
endif
iter_max = min(iter_max, kmax)

ndim = HOsp_dim
open(ud, file=dumpfile, status='replace', access='stream', form='unformatted')
call dump_i0('ndim', ndim)
call dump_i1('sp_2mj', HOsp_2mj, ndim)
call dump_i1('sp_2mt', HOsp_2mt, ndim)
call dump_i1('sp_l', HOsp_l, ndim)
call dump_i1('sp_2j', HOsp_2j, ndim)
call dump_i1('sp_tr', HOsp_tr, ndim)
call dump_r0('valence_Z', valence_Z)
call dump_r0('valence_N', valence_N)
call dump_i0('Mphip', proj_Mphip)
call dump_i0('Mphin', proj_Mphin)
call dump_r0('facpi', proj_facpi)
call dump_i0('gradient_type', gradient_type)
call dump_r0('gradient_eta', gradient_eta)
call dump_r0('gradient_mu', gradient_mu)
call dump_i0('constraint_dim', constraint_dim)
call dump_i0('constraint_pair', constraint_pair)
call dump_r0('factor_delta', factor_delta)
call dump_c0('H0', hamil_H0)
if ( constraint_dim > 0 ) then
  call dump_r1('constraint_HO', constraint_HO, int(constraint_dim,i64)*ndim*ndim)
  call dump_r1('constraint_val', constraint_val, int(constraint_dim,i64))
endif

allocate(e_hist(max(iter_max,1)), gn_hist(max(iter_max,1)), z_hist(max(iter_max,1)), &
         n_hist(max(iter_max,1)))
allocate(hmat(ndim,ndim), evals(ndim), work(3*ndim), gRR(ndim,ndim), hRR(ndim,ndim), &
         dRR(ndim,ndim))
e_hist = zero
gn_hist = zero
z_hist = zero
n_hist = zero
niter_done = 0

do iter = 1, iter_max
  field_H20v = zero
  call calculate_densities_real(bogo_U0,bogo_V0,dens_rhoRR,dens_kappaRR, &
                                HOsp_dim)
  call update_constraints_qpbasis
  if ( iter == 1 ) then
    call dump_r2('U0', bogo_U0, ndim, ndim)
    call dump_r2('V0', bogo_V0, ndim, ndim)
    if ( constraint_dim > 0 ) then
      call dump_r1('constraint_20', constraint_20, int(constraint_dim,i64)*ndim*ndim)
    endif
    call dump_angles(ndim, .false.)
  endif

  call project_wavefunction(0,iter_print,proj_Mphip,proj_Mphin,HOsp_dim)
  call calculate_lagrange_multipliers(field_H20v,HOsp_dim)

  if ( iter == 1 ) then
    call dump_c0('pnp_over', pnp_over)
    call dump_c0('pnp_ener', pnp_ener)
    call dump_c0('pnp_prot', pnp_prot)
    call dump_c0('pnp_neut', pnp_neut)
    call dump_c0('pnp_prot2', pnp_prot2)
    call dump_c0('pnp_neut2', pnp_neut2)
    call dump_c0('pnp_pari', pnp_pari)
    call dump_i0('iter_print', iter_print)
    call dump_r1('field_H20v', field_H20v, int(ndim,i64)*ndim)
    if ( constraint_dim > 0 ) then
      call dump_r1('lagrange_lambda1', lagrange_lambda1, int(constraint_dim,i64))
    endif
  endif

  if ( gradient_type > 0 ) then
    if ( iter == 1 ) then
      ! the matrices diagonalize_hsp_and_H11(0) hands to dsyev (module_gradient.f90:
      ! 177-213), built from the same public routines in the same order, and their
      ! eigenvalues from the same LAPACK call
      if ( max(proj_Mphip,proj_Mphip) == 1 ) then
        field_hspRR = real(field_hspLR)
        field_deltaRR = real(field_deltaRL)
      else
        call calculate_fields_diag(zone*dens_rhoRR,zone*dens_kappaRR,gRR,hRR, &
                                   dRR,ndim=ndim)
        field_hspRR = real(hRR)
        field_deltaRR = real(dRR)
      endif
      call calculate_H11_real(ndim)
      do m = 1, constraint_dim-constraint_pair
        k = (m-1)*(ndim**2) + 1
        call set_operator_qpbasis('f11',bogo_U0,bogo_V0,constraint_HO(k), &
                                  constraint_11(k),ndim)
      enddo
      l = 0
      do j = 1, ndim
        do i = 1, ndim
          l = l + 1
          q11_aux = zero
          q00_aux = zero
          do m = 1, constraint_dim
            k = (m-1)*(ndim**2)
            q11_aux = q11_aux + lagrange_lambda1(m) * constraint_11(k+l)
            q00_aux = q00_aux + lagrange_lambda1(m) * constraint_HO(k+l)
          enddo
          field_H11(i,j) = field_H11(i,j) - q11_aux
          field_hspRR(i,j) = field_hspRR(i,j) - q00_aux
        enddo
      enddo
      call dump_r2('hsp_RR', field_hspRR, ndim, ndim)
      call dump_r2('H11', field_H11, ndim, ndim)
      hmat = field_hspRR
      call dsyev('v','u',ndim,hmat,ndim,evals,work,3*ndim-1,info)
      call dump_r1('eigen_hsp', evals, int(ndim,i64))
      call dump_i0('info_hsp', info)
      hmat = field_H11
      call dsyev('v','u',ndim,hmat,ndim,evals,work,3*ndim-1,info)
      call dump_r1('eigen_H11', evals, int(ndim,i64))
      call dump_i0('info_H11', info)
    endif
    call diagonalize_hsp_and_H11(0,HOsp_dim)
  endif
  call calculate_gradient(HOsp_dim)

  if ( iter == 1 ) then
    call dump_r1('gradient_Zi', gradient_Zi, int(ndim,i64)*ndim)
    call dump_r0('gradient_norm', gradient_norm)
  endif

  e_hist(iter) = real(pnp_ener/pnp_over)
  z_hist(iter) = real(pnp_prot/pnp_over)
  n_hist(iter) = real(pnp_neut/pnp_over)
  gn_hist(iter) = gradient_norm
  niter_done = iter

  if ( gradient_norm <= gradient_eps ) exit
  call adjust_constraints(gradient_Zi,HOsp_dim)
  gradient_Zim1 = gradient_Zi
enddo

call dump_i0('niter', niter_done)
call dump_r1('hist_ener', e_hist, int(max(niter_done,1),i64))
call dump_r1('hist_prot', z_hist, int(max(niter_done,1),i64))
call dump_r1('hist_neut', n_hist, int(max(niter_done,1),i64))
call dump_r1('hist_gradnorm', gn_hist, int(max(niter_done,1),i64))
close(ud)

end subroutine run_proj

!------------------------------------------------------------------------------!
! Wall-clock timing of full iterations at the deck's gauge angles              !
!------------------------------------------------------------------------------!
subroutine run_timefull(outfile, kstr)
character(*), intent(in) :: outfile, kstr
integer :: iter_max, iter_write, iter_print, k, kmax
integer(i64) :: c0, c1, c2, c3, rate
real(r64) :: tproj, tgrad, tadj

read(kstr,*) kmax
call read_input(iter_max,iter_write,iter_print)
call set_nucleus
call set_basis
call set_hamiltonian
call set_wavefunctions
call set_fields
call set_particle_number
call set_pairs
call set_angular_momentum
call set_multipoles
call set_radius
call set_projection
call generate_wavefunction(HOsp_dim)
call check_symmetries(HOsp_dim)
call set_constraints
call set_gradient

open(ud, file=outfile, status='replace', form='formatted')
do k = 1, kmax
  ! one pass of the loop body of taurus_vap.f90:77-113
  call system_clock(c0, rate)
  field_H20v = zero
  call calculate_densities_real(bogo_U0,bogo_V0,dens_rhoRR,dens_kappaRR,HOsp_dim)
  call update_constraints_qpbasis
  call project_wavefunction(0,0,proj_Mphip,proj_Mphin,HOsp_dim)
  call system_clock(c1)
  call calculate_lagrange_multipliers(field_H20v,HOsp_dim)
  if ( gradient_type > 0 ) call diagonalize_hsp_and_H11(0,HOsp_dim)
  call calculate_gradient(HOsp_dim)
  call system_clock(c2)
  call adjust_constraints(gradient_Zi,HOsp_dim)
  gradient_Zim1 = gradient_Zi
  call system_clock(c3)
  tproj = real(c1-c0,r64) / real(rate,r64)
  tgrad = real(c2-c1,r64) / real(rate,r64)
  tadj = real(c3-c2,r64) / real(rate,r64)
  write(ud,'(i4,i6,3es16.8,es20.10)') k, proj_Mphip*proj_Mphin, tproj, tgrad, tadj, &
                                      real(pnp_ener/pnp_over)
  flush(ud)
enddo
close(ud)
end subroutine run_timefull

end program ref_harness
