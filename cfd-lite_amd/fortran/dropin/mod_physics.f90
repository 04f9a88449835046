! mod_physics (drop-in) -- the reference's src/modules/mod_physics.f90
! interface (type phys_t, update_boundaries, construct_physics, destroy_phys,
! update_time) for the GPU build, keeping the uvwp state on the device.
!
! With the reference's own mod_physics the drop-in mod_uvwp moves the whole
! uvwp state to the device and back around every solve_uvwp, because
! update_time (mod_physics.f90:101-112) and the output read and write the
! host arrays.  This module makes the device the owner of that state instead
! (mod_solver's cfdl_resident): update_time runs on the device, solve_uvwp
! sends the state once and then only the properties each call, and the
! host arrays are refreshed at the time steps where the reference's program
! writes its output (every 10th step and the last, src/main.f90:77-89).
! The energy equation stays host-owned as in the reference (its BC setters
! set host halos); its solve, if called, takes the face fluxes from the
! device.  Build with -fdefault-real-8 like the reference.
module mod_physics
  use mod_cell
  use mod_util
  use mod_eqn_setup
  use mod_scalar
  use mod_energy
  use mod_multiphase
  use mod_uvwp
  use mod_subdomains
  use mod_solver, only: cfdl_resident
  use cfdl_c
  implicit none

  type :: phys_t
    real :: dt=0.01
    integer :: ntstep=10
    integer :: ncoef=3
    integer :: nit=100
    integer :: n_subdomains=4
    integer :: nintf_c2b
    type(scalar_t), pointer :: scalar => NULL()
    type(uvwp_t), pointer :: uvwp => NULL()
    type(energy_t), pointer :: energy => null()
    integer :: iphase
    type (phase_t) :: phase(NPHASES)
    real, allocatable, dimension(:) :: ap,anb,b,phic
    type(subdomain_t), allocatable :: subdomain(:)
    type(intf_t) , allocatable :: intf(:,:)
    type(properties_t) :: prop
    ! time steps completed (update_time calls)
    integer :: steps_done = 0
  end type
 contains

  ! update_boundaries, mod_physics.f90:38-50: the setters name each
  ! section's type (uvwp, applied by the device inside solve_uvwp) and set
  ! the energy halos on the host
  subroutine update_boundaries(phys,geom)
    type(phys_t) :: phys
    type(geometry_t) :: geom
    integer :: i
    do i=1,phys%nintf_c2b
      call phys%uvwp%bcs(i)%coef(geom,phys%uvwp,phys%prop)
      call phys%energy%bcs(i)%coef(geom,phys%energy,phys%prop)
    end do
  end subroutine

  ! construct_physics, mod_physics.f90:52-79, then the device owns the
  ! uvwp state
  subroutine construct_physics(phys,geom)
    type(phys_t) :: phys
    type(geometry_t) :: geom
    allocate(phys%ap(geom%ne)); phys%ap=0.
    allocate(phys%b(geom%ne)); phys%b=0.
    allocate(phys%anb(2*geom%nf-geom%nbf)); phys%anb=0.
    allocate(phys%phic(geom%ne+geom%nbf)); phys%phic=0.
    phys%nintf_c2b=geom%mg%nintf_c2b
    if(phys%n_subdomains>1) call construct_subdomains(phys%subdomain,phys%n_subdomains,phys%intf,geom)
    call init_properties(phys%prop,0,geom%ne)
    phys%uvwp => construct_uvwp(geom,phys%prop,phys%dt)
    phys%energy => construct_energy(geom,phys%uvwp%mip,phys%prop)
    phys%steps_done = 0
    cfdl_resident = .true.
  end subroutine

  ! destroy_phys, mod_physics.f90:81-99
  subroutine destroy_phys(phys)
    type(phys_t) :: phys
    integer :: c,cnb
    deallocate(phys%anb,phys%ap,phys%b,phys%phic)
    if(phys%n_subdomains>1) then
      do c=1,phys%n_subdomains
        deallocate(phys%subdomain(c)%ap, phys%subdomain(c)%b, phys%subdomain(c)%anb)
        deallocate(phys%subdomain(c)%phic, phys%subdomain(c)%ef2nb, phys%subdomain(c)%ef2nb_idx)
        do cnb=1,phys%n_subdomains
          deallocate(phys%intf(c,cnb)%index1)
        end do
      end do
      deallocate(phys%subdomain)
      deallocate(phys%intf)
    end if
  end subroutine

  ! update_time, mod_physics.f90:101-112: u0, v0, w0 and mip0 on the device
  ! (on the host before the first solve), phi0 of the energy on the host;
  ! then, at the steps the reference's program writes output after, the
  ! host copy of the uvwp state and of the work arrays
  subroutine update_time(phys)
    type(phys_t) :: phys
    if (c_associated(phys%uvwp%ctx)) then
      call cfdl_check(cfdl_update_time(phys%uvwp%ctx), 'update_time')
    else
      phys%uvwp%u0=phys%uvwp%u
      phys%uvwp%v0=phys%uvwp%v
      phys%uvwp%w0=phys%uvwp%w
      phys%uvwp%mip0=phys%uvwp%mip
    end if
    phys%energy%phi0=phys%energy%phi
    phys%steps_done = phys%steps_done + 1
    if (c_associated(phys%uvwp%ctx) .and. (mod(phys%steps_done,10) == 0 .or. &
        phys%steps_done >= phys%ntstep)) &
      call pull_state(phys%uvwp, phys%ap, phys%anb, phys%b, phys%phic)
  end subroutine

end module
