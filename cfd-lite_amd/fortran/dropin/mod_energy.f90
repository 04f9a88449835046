! mod_energy (drop-in) -- the reference's src/equations/mod_energy.f90
! interface (type energy_t and construct_energy, destroy_energy,
! solve_energy, calc_coef_energy and the BC setters lid / dirichlet0) with
! the equation's computations done by libcfdl_hip.so on the GPU context that
! the drop-in mod_uvwp creates for the mesh (mod_solver's cfdl_active).
!
! As in the drop-in mod_uvwp, the host arrays stay authoritative: the BC
! setters and construct_energy work on them as the reference's do (the rest
! of the reference, update_time and output, reads them), solve_energy and
! calc_coef_energy upload the state they read and download what they write.
! Build with -fdefault-real-8 like the reference.
module mod_energy
  use mod_eqn_setup
  use cfdl_c
  implicit none

  type, extends(equation_t) :: energy_t
     real, allocatable, dimension(:) :: t,gt
     real, pointer, dimension(:) :: mip
   contains
    procedure :: calc_coef_energy
  end type

contains

  ! construct_energy, mod_energy.f90:14-48: t = 273 K, enthalpy phi = cp*t on
  ! the cells (the halos are set by the BC setters before any use), phi0 =
  ! phi, zero gradients; 'top' is bound to lid, every other section to
  ! dirichlet0
  function construct_energy(geom,mip,prop) result(eqn)
    type(energy_t), pointer :: eqn
    type(properties_t) :: prop
    real, target, dimension(:) :: mip
    type(geometry_t) :: geom
    integer :: length
    type(bc_t), pointer :: bcp

    allocate(eqn)
    length = geom%ne+geom%nbf
    eqn%name='energy'
    allocate(eqn%phi(length), eqn%phi0(length), eqn%grad(3*length))
    allocate(eqn%t(length), eqn%gt(3*length))
    eqn%mip=>mip
    eqn%t=273._c_double
    eqn%phi=0._c_double
    eqn%phi(1:geom%ne)=eqn%t(1:geom%ne)*prop%cp(1:geom%ne)
    eqn%phi0=eqn%phi
    eqn%grad=0._c_double
    eqn%gt=0._c_double
    allocate(eqn%bcs(geom%mg%nintf_c2b))
    bcp=>eqn%make_bc(geom,'top');bcp%coef=>lid
    bcp=>eqn%make_bc(geom,'west');bcp%coef=>dirichlet0
    bcp=>eqn%make_bc(geom,'east');bcp%coef=>dirichlet0
    bcp=>eqn%make_bc(geom,'south');bcp%coef=>dirichlet0
    bcp=>eqn%make_bc(geom,'north');bcp%coef=>dirichlet0
    bcp=>eqn%make_bc(geom,'bottom');bcp%coef=>dirichlet0
  end function

  subroutine destroy_energy(eqn)
    type(energy_t),pointer :: eqn
    deallocate(eqn%bcs,eqn%grad,eqn%phi0,eqn%phi,eqn%t,eqn%gt)
    deallocate(eqn)
  end subroutine

  subroutine energy_ctx(ctx, geom)
    use mod_solver, only: cfdl_active
    type(c_ptr), intent(out) :: ctx
    type(geometry_t) :: geom
    if (.not. c_associated(cfdl_active)) then
      write(*,*) 'mod_energy: no GPU context for this mesh (construct uvwp first)'
      error stop 1
    end if
    ctx = cfdl_active
  end subroutine

  ! the state the energy routines read, host -> device
  subroutine put_energy(ctx, eqn, geom, prop, dt)
    use mod_solver, only: cfdl_resident
    type(c_ptr) :: ctx
    class(energy_t) :: eqn
    type(geometry_t) :: geom
    type(properties_t) :: prop
    real :: dt
    integer(c_int64_t) :: n
    n = geom%ne + geom%nbf
    call cfdl_check(cfdl_set_thermal(ctx, prop%cp, prop%tc), 'thermal properties')
    call cfdl_check(cfdl_set_properties(ctx, prop%rho, prop%mu), 'properties')
    call cfdl_check(cfdl_set_dt(ctx, dt), 'dt')
    call cfdl_check(cfdl_set_field(ctx, CFDL_T, eqn%t, n), 'upload')
    call cfdl_check(cfdl_set_field(ctx, CFDL_ENTH, eqn%phi, n), 'upload')
    call cfdl_check(cfdl_set_field(ctx, CFDL_ENTH0, eqn%phi0, n), 'upload')
    call cfdl_check(cfdl_set_field(ctx, CFDL_GENTH, eqn%grad, 3*n), 'upload')
    call cfdl_check(cfdl_set_field(ctx, CFDL_GT, eqn%gt, 3*n), 'upload')
    ! the face fluxes are the uvwp state's: on the device already when it
    ! owns that state (the drop-in mod_physics)
    if (.not. cfdl_resident) &
      call cfdl_check(cfdl_set_field(ctx, CFDL_MIP, eqn%mip, int(geom%nf, c_int64_t)), 'upload')
  end subroutine

  subroutine get_work(ctx, geom, ap, anb, b)
    type(c_ptr) :: ctx
    type(geometry_t) :: geom
    real, dimension(*) :: ap, anb, b
    integer(c_int64_t) :: ne, l
    ne = geom%ne
    l = 2*geom%nf - geom%nbf
    call cfdl_check(cfdl_get_field(ctx, CFDL_AP, ap, ne), 'download')
    call cfdl_check(cfdl_get_field(ctx, CFDL_ANB, anb, l), 'download')
    call cfdl_check(cfdl_get_field(ctx, CFDL_B, b, ne), 'download')
  end subroutine

  ! solve_energy, mod_energy.f90:58-80, on the device
  subroutine solve_energy(eqn,prop,geom,dt,nit,ap,anb,b,phic)
    use mod_solver, only: print_new_log_lines
    type(energy_t) :: eqn
    type(properties_t) :: prop
    type(geometry_t) :: geom
    real :: dt
    real, dimension(*) :: ap,anb,b,phic
    integer :: nit
    type(c_ptr) :: ctx
    integer(c_int64_t) :: n
    call energy_ctx(ctx, geom)
    n = geom%ne + geom%nbf
    call put_energy(ctx, eqn, geom, prop, dt)
    call cfdl_check(cfdl_solve_energy(ctx), 'solve_energy')
    call cfdl_check(cfdl_get_field(ctx, CFDL_T, eqn%t, n), 'download')
    call cfdl_check(cfdl_get_field(ctx, CFDL_ENTH, eqn%phi, n), 'download')
    call cfdl_check(cfdl_get_field(ctx, CFDL_GENTH, eqn%grad, 3*n), 'download')
    call cfdl_check(cfdl_get_field(ctx, CFDL_GT, eqn%gt, 3*n), 'download')
    call get_work(ctx, geom, ap, anb, b)
    call print_new_log_lines()
  end subroutine

  ! calc_coef_energy, mod_energy.f90:82-169, on the device
  subroutine calc_coef_energy(eqn,ap,anb,b,geom,prop,dt)
    type(properties_t) :: prop
    type(geometry_t) :: geom
    class(energy_t) :: eqn
    real :: dt
    real :: ap(*),anb(*),b(*)
    type(c_ptr) :: ctx
    call energy_ctx(ctx, geom)
    call put_energy(ctx, eqn, geom, prop, dt)
    call cfdl_check(cfdl_calc_coef_energy(ctx), 'calc_coef_energy')
    call get_work(ctx, geom, ap, anb, b)
  end subroutine

  ! BC setters lid (t = 373) and dirichlet0 (t = 273), mod_energy.f90:
  ! 182-212: host halo values, phi = cp(owner)*t
  subroutine lid(bc,geom,eqn,prop)
    use mod_util
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    call set_halos(bc, geom, eqn, prop, 373._c_double)
  end subroutine

  subroutine dirichlet0(bc,geom,eqn,prop)
    use mod_util
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    call set_halos(bc, geom, eqn, prop, 273._c_double)
  end subroutine

  subroutine set_halos(bc, geom, eqn, prop, tb)
    use mod_util
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    real :: tb
    integer :: e, enb, lfnb
    bc%bc_type='dirichlet'
    select type(eqn)
    type is(energy_t)
      do e=bc%esec(1),bc%esec(2)
        call get_idx(abs(geom%mg%fine_lvl%bs(e)),0,enb,lfnb)
        eqn%t(e)=tb
        eqn%phi(e)=prop%cp(enb)*eqn%t(e)
      end do
    end select
  end subroutine

end module
