! mod_solver (drop-in) -- the plain-array solver procedures of the
! reference's src/modules/mod_solver.f90 (calc_grad :40, solve_gs :255,
! solve :329) with the same module name, names and arguments, computed by
! libcfdl_hip.so on the GPU context that mod_uvwp registers for the mesh.
! Arrays are copied in and out around each call, so the host arrays stay
! authoritative as in the reference.  Residual lines are printed in the
! reference's format (:6).  Build with -fdefault-real-8 like the reference.
module mod_solver
  use mod_util
  use mod_subdomains
  use cfdl_c
  implicit none

  character(len=*), parameter :: oformat = "(5x,A16,x,i5,x,15x,es9.3e2,3x,es9.3e2,3x,es9.3e2)"
  ! context of the mesh in use (set by mod_uvwp) and its sizes
  type(c_ptr), save :: cfdl_active = c_null_ptr
  integer, save :: cfdl_ne = 0, cfdl_nf = 0, cfdl_nbf = 0, cfdl_nit = -1
  integer, save :: cfdl_logged = 0
  ! set by the drop-in mod_physics: the device owns the uvwp state between
  ! the steps where the program writes output (mod_uvwp, mod_energy)
  logical, save :: cfdl_resident = .false.

contains

  subroutine cfdl_bind(ctx, ne, nf, nbf, nit)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: ne, nf, nbf, nit
    cfdl_active = ctx
    cfdl_ne = ne
    cfdl_nf = nf
    cfdl_nbf = nbf
    cfdl_nit = nit
    cfdl_logged = 0
  end subroutine

  subroutine require_ctx(ne, nf, nbf, what)
    integer, intent(in) :: ne, nf, nbf
    character(len=*), intent(in) :: what
    if (.not. c_associated(cfdl_active) .or. ne /= cfdl_ne .or. nf /= cfdl_nf &
        .or. nbf /= cfdl_nbf) then
      write(*,*) 'mod_solver: ', what, ': no GPU context for this mesh (construct uvwp first)'
      error stop 1
    end if
  end subroutine

  ! print the solver lines the device logged since the last call
  subroutine print_new_log_lines()
    integer(c_int) :: n, it
    character(kind=c_char) :: nm(8)
    real(c_double) :: ri, rf, rm
    character(len=16) :: name
    integer :: k, j
    call cfdl_check(cfdl_log_count(cfdl_active, n), 'log')
    do k = cfdl_logged, n - 1
      call cfdl_check(cfdl_log_get(cfdl_active, k, nm, it, ri, rf, rm), 'log')
      name = ' '
      do j = 1, 8
        if (nm(j) == c_null_char) exit
        name(j:j) = nm(j)
      end do
      write(*,oformat) name, it, ri, rf, rm
    end do
    cfdl_logged = n
  end subroutine

  subroutine calc_grad(phi,grad,xc,yc,zc,ef2nb_idx,ef2nb,ne,nf,nbf)
    integer :: ne, nf, nbf
    real, dimension(3,ne+nbf) :: grad
    real, dimension(ne+nbf) :: phi
    integer :: ef2nb_idx(ne+1), ef2nb(2*nf-nbf,2)
    real :: xc(ne+nbf), yc(ne+nbf), zc(ne+nbf)
    integer(c_int64_t) :: n
    call require_ctx(ne, nf, nbf, 'calc_grad')
    n = ne + nbf
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_WORK, phi, n), 'calc_grad')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_GWORK, grad, 3*n), 'calc_grad')
    call cfdl_check(cfdl_calc_grad(cfdl_active, CFDL_WORK, CFDL_GWORK), 'calc_grad')
    call cfdl_check(cfdl_get_field(cfdl_active, CFDL_GWORK, grad, 3*n), 'calc_grad')
  end subroutine

  subroutine solve_gs(cname,phi,ap,anb,b,ef2nb_idx,ef2nb,ne,nf,nbf,nit)
    character(len=*) :: cname
    integer :: ne, nf, nbf, nit
    real, dimension(ne+nbf) :: phi
    real, dimension(ne) :: ap, b
    real, dimension(2*nf-nbf) :: anb
    integer :: ef2nb_idx(ne+1), ef2nb(2*nf-nbf,2)
    call require_ctx(ne, nf, nbf, 'solve_gs')
    if (nit /= cfdl_nit) then
      write(*,*) 'mod_solver: solve_gs: nit differs from the context''s'
      error stop 1
    end if
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_WORK, phi, int(ne+nbf, c_int64_t)), 'solve_gs')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_AP, ap, int(ne, c_int64_t)), 'solve_gs')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_B, b, int(ne, c_int64_t)), 'solve_gs')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_ANB, anb, int(2*nf-nbf, c_int64_t)), 'solve_gs')
    call cfdl_check(cfdl_solve_gs(cfdl_active, cfdl_str(cname), CFDL_WORK, CFDL_B), 'solve_gs')
    call cfdl_check(cfdl_get_field(cfdl_active, CFDL_WORK, phi, int(ne+nbf, c_int64_t)), 'solve_gs')
    call print_new_log_lines()
  end subroutine

  subroutine solve(cname,subdomain,intf,geom,ap,anb,b,phi,nsubd,nit)
    integer :: nit, nsubd
    character(len=*) :: cname
    real, dimension(*) :: ap, anb, b, phi
    type(geometry_t) :: geom
    type(subdomain_t) :: subdomain(nsubd)
    type(intf_t) :: intf(nsubd,nsubd)
    integer :: ne, nf, nbf
    ne = geom%ne
    nf = geom%nf
    nbf = geom%nbf
    call require_ctx(ne, nf, nbf, 'solve')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_AP, ap, int(ne, c_int64_t)), 'solve')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_B, b, int(ne, c_int64_t)), 'solve')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_ANB, anb, int(2*nf-nbf, c_int64_t)), 'solve')
    call cfdl_check(cfdl_set_field(cfdl_active, CFDL_PHIC, phi, int(ne+nbf, c_int64_t)), 'solve')
    call cfdl_check(cfdl_solve(cfdl_active, cfdl_str(cname)), 'solve')
    call cfdl_check(cfdl_get_field(cfdl_active, CFDL_PHIC, phi, int(ne+nbf, c_int64_t)), 'solve')
    call print_new_log_lines()
  end subroutine
end module mod_solver
