! mod_uvwp (drop-in) -- the reference's src/equations/mod_uvwp.f90 interface
! (type uvwp_t and the procedures construct_uvwp, destroy_uvwp, solve_uvwp,
! adjust_pc, calc_coef_uvw, calc_coef_p, update_uvwp, calc_mip and the BC
! setters) with every computation done by libcfdl_hip.so on the GPU.
!
! With the reference's mod_physics the host arrays of uvwp_t stay
! authoritative between calls, as the rest of the reference (mod_physics's
! update_time, output) reads and writes them: each entry point uploads the
! state it reads and downloads what it writes.  With the drop-in
! mod_physics (mod_solver's cfdl_resident) the device owns the state: it is
! uploaded once, solve_uvwp downloads nothing, the other entry points
! download what they write, and pull_state refreshes the host copy.  The BC setters keep the reference's host-side role of naming the
! boundary type of their section; the halo values they define are applied
! by the device inside solve_uvwp (update_boundaries runs directly before
! solve_uvwp in the reference's loop, src/main.f90:52-56).
! Build with -fdefault-real-8 like the reference.
module mod_uvwp
  use mod_eqn_setup
  use mod_util
  use cfdl_c
  implicit none

  type, extends(equation_t) :: uvwp_t
    real, allocatable, dimension(:) :: u, v, w, p
    real, allocatable, dimension(:) :: gu, gv, gw, gp, gpc
    real, allocatable, dimension(:) :: mip, mip0
    real, allocatable, dimension(:) :: u0, v0, w0
    real, allocatable, dimension(:) :: bu, bv, bw, d, dc
    ! device context of this equation, its solver cap and subdomain count
    type(c_ptr) :: ctx = c_null_ptr
    integer :: ctx_nit = -1, ctx_nsubd = -1
    ! per boundary interface: CFDL_BC_* kind named by its setter
    integer, allocatable :: bc_kind(:)
    ! the device holds the state (resident mode, after the first upload)
    logical :: on_device = .false.
  contains
    procedure :: calc_coef_uvw
  end type

contains

  ! ------------------------------------------------------------ context ----
  subroutine ensure_ctx(eqn, geom, prop, dt, nit, nsubd)
    use mod_properties
    use mod_solver, only: cfdl_bind
    class(uvwp_t) :: eqn
    type(geometry_t) :: geom
    type(properties_t) :: prop
    real :: dt
    integer :: nit, nsubd
    integer(c_int), allocatable :: kind(:), gf2g(:), g2gf(:), g2gf_idx(:)
    integer :: s, ne, nbf
    if (c_associated(eqn%ctx) .and. eqn%ctx_nit == nit .and. eqn%ctx_nsubd == nsubd) return
    if (c_associated(eqn%ctx)) call cfdl_ctx_destroy(eqn%ctx)
    ne = geom%ne
    nbf = geom%nbf
    allocate(kind(geom%mg%nsec))
    kind = CFDL_BC_WALL
    if (nsubd > 1) then
      allocate(gf2g(ne), g2gf(ne), g2gf_idx(nsubd+1))
      gf2g = geom%mg%gf2g(1)%p(1:ne)
      g2gf = geom%mg%g2gf(1)%p(1:ne)
      g2gf_idx = geom%mg%g2gf(1)%idx(1:nsubd+1)
    else
      allocate(gf2g(1), g2gf(1), g2gf_idx(2))
      gf2g = 0
      g2gf = 0
      g2gf_idx = 0
    end if
    call cfdl_check(cfdl_ctx_create(ne, geom%nf, nbf, geom%ef2nb_idx, geom%ef2nb, &
         geom%mg%fine_lvl%s2g, geom%mg%fine_lvl%bs(ne+1:ne+nbf), geom%xc, geom%yc, geom%zc, &
         geom%vol, geom%aip, geom%rip, geom%mg%nsec, geom%mg%esec, kind, nsubd, gf2g, g2gf, &
         g2gf_idx, prop%rho, prop%mu, dt, nit, eqn%ctx), 'context')
    eqn%ctx_nit = nit
    eqn%ctx_nsubd = nsubd
    eqn%on_device = .false.
    call cfdl_bind(eqn%ctx, ne, geom%nf, nbf, nit)
    do s = 1, size(eqn%bcs)
      if (eqn%bc_kind(s) >= 0) call cfdl_check(cfdl_set_bc_kind(eqn%ctx, eqn%bcs(s)%esec(1), &
           eqn%bcs(s)%esec(2), eqn%bc_kind(s)), 'bc kind')
    end do
  end subroutine

  subroutine sync_params(eqn, prop, dt)
    use mod_properties
    class(uvwp_t) :: eqn
    type(properties_t) :: prop
    real :: dt
    integer :: s
    call cfdl_check(cfdl_set_properties(eqn%ctx, prop%rho, prop%mu), 'properties')
    call cfdl_check(cfdl_set_dt(eqn%ctx, dt), 'dt')
    do s = 1, size(eqn%bcs)
      if (eqn%bc_kind(s) >= 0) call cfdl_check(cfdl_set_bc_kind(eqn%ctx, eqn%bcs(s)%esec(1), &
           eqn%bcs(s)%esec(2), eqn%bc_kind(s)), 'bc kind')
    end do
  end subroutine

  subroutine dev_put(eqn, field, a)
    class(uvwp_t) :: eqn
    integer(c_int) :: field
    real, intent(in) :: a(:)
    call cfdl_check(cfdl_set_field(eqn%ctx, field, a, int(size(a), c_int64_t)), 'upload')
  end subroutine

  subroutine dev_get(eqn, field, a)
    class(uvwp_t) :: eqn
    integer(c_int) :: field
    real, intent(inout) :: a(:)
    call cfdl_check(cfdl_get_field(eqn%ctx, field, a, int(size(a), c_int64_t)), 'download')
  end subroutine

  ! the state solve_uvwp reads.  bu, bv, bw, d, dc and gpc are not sent:
  ! solve_uvwp assigns them before any read (mod_uvwp.f90:236-238, 277-283,
  ! 131 before 385), so host changes to them (the VTU writer overwrites dc)
  ! cannot reach its results.
  subroutine put_inputs(eqn)
    use mod_solver, only: cfdl_resident
    class(uvwp_t) :: eqn
    if (cfdl_resident .and. eqn%on_device) return
    eqn%on_device = cfdl_resident
    call dev_put(eqn, CFDL_U, eqn%u); call dev_put(eqn, CFDL_V, eqn%v); call dev_put(eqn, CFDL_W, eqn%w)
    call dev_put(eqn, CFDL_P, eqn%p)
    call dev_put(eqn, CFDL_U0, eqn%u0); call dev_put(eqn, CFDL_V0, eqn%v0); call dev_put(eqn, CFDL_W0, eqn%w0)
    call dev_put(eqn, CFDL_GU, eqn%gu); call dev_put(eqn, CFDL_GV, eqn%gv); call dev_put(eqn, CFDL_GW, eqn%gw)
    call dev_put(eqn, CFDL_GP, eqn%gp)
    call dev_put(eqn, CFDL_MIP, eqn%mip); call dev_put(eqn, CFDL_MIP0, eqn%mip0)
  end subroutine

  ! the whole state, for the routines called one by one (calc_coef_p reads
  ! dc, calc_mip d, update_uvwp dc and gpc)
  subroutine put_state(eqn)
    use mod_solver, only: cfdl_resident
    class(uvwp_t) :: eqn
    if (cfdl_resident .and. eqn%on_device) return
    call put_inputs(eqn)
    call dev_put(eqn, CFDL_GPC, eqn%gpc)
    call dev_put(eqn, CFDL_BU, eqn%bu); call dev_put(eqn, CFDL_BV, eqn%bv); call dev_put(eqn, CFDL_BW, eqn%bw)
    call dev_put(eqn, CFDL_D, eqn%d); call dev_put(eqn, CFDL_DC, eqn%dc)
  end subroutine

  subroutine get_state(eqn)
    class(uvwp_t) :: eqn
    call dev_get(eqn, CFDL_U, eqn%u); call dev_get(eqn, CFDL_V, eqn%v); call dev_get(eqn, CFDL_W, eqn%w)
    call dev_get(eqn, CFDL_P, eqn%p)
    call dev_get(eqn, CFDL_GU, eqn%gu); call dev_get(eqn, CFDL_GV, eqn%gv); call dev_get(eqn, CFDL_GW, eqn%gw)
    call dev_get(eqn, CFDL_GP, eqn%gp); call dev_get(eqn, CFDL_GPC, eqn%gpc)
    call dev_get(eqn, CFDL_MIP, eqn%mip)
    call dev_get(eqn, CFDL_BU, eqn%bu); call dev_get(eqn, CFDL_BV, eqn%bv); call dev_get(eqn, CFDL_BW, eqn%bw)
    call dev_get(eqn, CFDL_D, eqn%d); call dev_get(eqn, CFDL_DC, eqn%dc)
  end subroutine

  ! resident mode: the host copy of the whole state and of the work arrays
  ! (the drop-in mod_physics's update_time, at the steps the program writes
  ! output after)
  subroutine pull_state(eqn, ap, anb, b, phic)
    class(uvwp_t) :: eqn
    real, dimension(:) :: ap, anb, b, phic
    if (.not. c_associated(eqn%ctx)) return
    call get_state(eqn)
    call dev_get(eqn, CFDL_U0, eqn%u0); call dev_get(eqn, CFDL_V0, eqn%v0)
    call dev_get(eqn, CFDL_W0, eqn%w0); call dev_get(eqn, CFDL_MIP0, eqn%mip0)
    call dev_get(eqn, CFDL_AP, ap); call dev_get(eqn, CFDL_ANB, anb)
    call dev_get(eqn, CFDL_B, b); call dev_get(eqn, CFDL_PHIC, phic)
  end subroutine

  subroutine put_work(eqn, geom, ap, anb, b, phic)
    class(uvwp_t) :: eqn
    type(geometry_t) :: geom
    real, dimension(*) :: ap, anb, b, phic
    integer :: ne, l, n
    ne = geom%ne
    l = 2*geom%nf - geom%nbf
    n = ne + geom%nbf
    call cfdl_check(cfdl_set_field(eqn%ctx, CFDL_AP, ap, int(ne, c_int64_t)), 'upload')
    call cfdl_check(cfdl_set_field(eqn%ctx, CFDL_ANB, anb, int(l, c_int64_t)), 'upload')
    call cfdl_check(cfdl_set_field(eqn%ctx, CFDL_B, b, int(ne, c_int64_t)), 'upload')
    call cfdl_check(cfdl_set_field(eqn%ctx, CFDL_PHIC, phic, int(n, c_int64_t)), 'upload')
  end subroutine

  subroutine get_work(eqn, geom, ap, anb, b, phic)
    class(uvwp_t) :: eqn
    type(geometry_t) :: geom
    real, dimension(*) :: ap, anb, b, phic
    integer :: ne, l, n
    ne = geom%ne
    l = 2*geom%nf - geom%nbf
    n = ne + geom%nbf
    call cfdl_check(cfdl_get_field(eqn%ctx, CFDL_AP, ap, int(ne, c_int64_t)), 'download')
    call cfdl_check(cfdl_get_field(eqn%ctx, CFDL_ANB, anb, int(l, c_int64_t)), 'download')
    call cfdl_check(cfdl_get_field(eqn%ctx, CFDL_B, b, int(ne, c_int64_t)), 'download')
    call cfdl_check(cfdl_get_field(eqn%ctx, CFDL_PHIC, phic, int(n, c_int64_t)), 'download')
  end subroutine

  ! ------------------------------------------------ reference interface ----
  ! construct_uvwp, mod_uvwp.f90:20-84
  function construct_uvwp(geom, prop, dt) result(eqn)
    use mod_properties
    type(uvwp_t), pointer :: eqn
    type(properties_t) :: prop
    type(geometry_t) :: geom
    real :: dt
    integer :: n
    type(bc_t), pointer :: bcp
    allocate(eqn)
    eqn%name = 'uvwp'
    n = geom%ne + geom%nbf
    allocate(eqn%u(n), eqn%v(n), eqn%w(n), eqn%p(n), eqn%u0(n), eqn%v0(n), eqn%w0(n))
    allocate(eqn%gu(3*n), eqn%gv(3*n), eqn%gw(3*n), eqn%gp(3*n), eqn%gpc(3*n))
    allocate(eqn%mip(geom%nf), eqn%mip0(geom%nf))
    allocate(eqn%bu(geom%ne), eqn%bv(geom%ne), eqn%bw(geom%ne), eqn%d(geom%ne), eqn%dc(geom%ne))
    eqn%u = 0.; eqn%v = 0.; eqn%w = 0.; eqn%p = 0.
    eqn%u0 = 0.; eqn%v0 = 0.; eqn%w0 = 0.
    eqn%gp = 0.; eqn%gu = 0.; eqn%gv = 0.; eqn%gw = 0.; eqn%gpc = 0.
    eqn%mip = 0.; eqn%mip0 = 0.
    eqn%bu = 0.; eqn%bv = 0.; eqn%bw = 0.; eqn%d = 0.; eqn%dc = 0.
    allocate(eqn%bcs(geom%mg%nintf_c2b))
    allocate(eqn%bc_kind(geom%mg%nintf_c2b))
    eqn%bc_kind = -1
    bcp => eqn%make_bc(geom, 'top');    bcp%coef => lid
    bcp => eqn%make_bc(geom, 'west');   bcp%coef => dirichlet0
    bcp => eqn%make_bc(geom, 'east');   bcp%coef => dirichlet0
    bcp => eqn%make_bc(geom, 'south');  bcp%coef => dirichlet0
    bcp => eqn%make_bc(geom, 'north');  bcp%coef => dirichlet0
    bcp => eqn%make_bc(geom, 'bottom'); bcp%coef => dirichlet0
    call calc_mip(eqn, prop, geom, .false., dt)
    eqn%mip0 = eqn%mip
  end function

  subroutine destroy_uvwp(eqn)
    type(uvwp_t), pointer :: eqn
    if (c_associated(eqn%ctx)) call cfdl_ctx_destroy(eqn%ctx)
    deallocate(eqn%bcs, eqn%u, eqn%v, eqn%w, eqn%p, eqn%u0, eqn%v0, eqn%w0, eqn%gp, eqn%gpc, &
               eqn%mip, eqn%bu, eqn%bv, eqn%bw, eqn%d, eqn%dc)
    deallocate(eqn)
  end subroutine

  ! solve_uvwp, mod_uvwp.f90:95-134: the BC setters of this outer iteration
  ! (update_boundaries) and the whole SIMPLE(C) iteration on the device
  subroutine solve_uvwp(eqn, prop, geom, dt, nit, ap, anb, b, phic, subdomain, intf, nsubd)
    use mod_properties
    use mod_subdomains
    use mod_solver, only: print_new_log_lines, cfdl_resident
    type(uvwp_t) :: eqn
    type(properties_t) :: prop
    type(geometry_t) :: geom
    real :: dt
    real, dimension(*) :: ap, anb, b, phic
    integer :: nit, nsubd
    type(subdomain_t) :: subdomain(nsubd)
    type(intf_t) :: intf(nsubd,nsubd)
    call ensure_ctx(eqn, geom, prop, dt, nit, nsubd)
    call sync_params(eqn, prop, dt)
    call put_inputs(eqn)
    call cfdl_check(cfdl_update_boundaries(eqn%ctx), 'update_boundaries')
    call cfdl_check(cfdl_solve_uvwp(eqn%ctx), 'solve_uvwp')
    if (.not. cfdl_resident) then
      call get_state(eqn)
      call get_work(eqn, geom, ap, anb, b, phic)
    end if
    call print_new_log_lines()
  end subroutine

  ! adjust_pc, mod_uvwp.f90:136-158 (pref = pc(1), as solve_uvwp passes it)
  subroutine adjust_pc(pc, pref, geom)
    type(geometry_t) :: geom
    real :: pc(geom%ne+geom%nbf), pref
    type(c_ptr) :: ctx
    if (pref /= pc(1)) then
      write(*,*) 'mod_uvwp: adjust_pc supports pref = pc(1), the reference''s only use'
      error stop 1
    end if
    call active_ctx(ctx)
    call cfdl_check(cfdl_set_field(ctx, CFDL_PHIC, pc, int(geom%ne+geom%nbf, c_int64_t)), 'adjust_pc')
    call cfdl_check(cfdl_adjust_pc(ctx), 'adjust_pc')
    call cfdl_check(cfdl_get_field(ctx, CFDL_PHIC, pc, int(geom%ne+geom%nbf, c_int64_t)), 'adjust_pc')
  end subroutine

  subroutine active_ctx(ctx)
    use mod_solver, only: cfdl_active
    type(c_ptr), intent(out) :: ctx
    if (.not. c_associated(cfdl_active)) then
      write(*,*) 'mod_uvwp: no GPU context (construct uvwp first)'
      error stop 1
    end if
    ctx = cfdl_active
  end subroutine

  ! calc_coef_uvw, mod_uvwp.f90:161-286
  subroutine calc_coef_uvw(eqn, ap, anb, geom, prop, dt)
    use mod_properties
    class(uvwp_t) :: eqn
    type(properties_t) :: prop
    type(geometry_t) :: geom
    real :: dt, ap(geom%ne), anb(geom%nf*2-geom%nbf)
    real :: bdummy(geom%ne), pdummy(geom%ne+geom%nbf)
    call ensure_ctx(eqn, geom, prop, dt, max(eqn%ctx_nit, 0), max(eqn%ctx_nsubd, 1))
    call sync_params(eqn, prop, dt)
    call put_state(eqn)
    call cfdl_check(cfdl_calc_coef_uvw(eqn%ctx), 'calc_coef_uvw')
    call get_state(eqn)
    call get_work(eqn, geom, ap, anb, bdummy, pdummy)
  end subroutine

  ! calc_coef_p, mod_uvwp.f90:289-368
  subroutine calc_coef_p(eqn, ap, anb, b, geom, prop)
    use mod_properties
    class(uvwp_t) :: eqn
    type(properties_t) :: prop
    type(geometry_t) :: geom
    real, dimension(geom%ne) :: ap, b
    real, dimension(2*geom%nf-geom%nbf) :: anb
    real :: pdummy(geom%ne+geom%nbf)
    if (.not. c_associated(eqn%ctx)) then
      write(*,*) 'mod_uvwp: calc_coef_p before any other uvwp routine'
      error stop 1
    end if
    call cfdl_check(cfdl_set_properties(eqn%ctx, prop%rho, prop%mu), 'properties')
    call put_state(eqn)
    call cfdl_check(cfdl_calc_coef_p(eqn%ctx), 'calc_coef_p')
    call get_work(eqn, geom, ap, anb, b, pdummy)
  end subroutine

  ! update_uvwp, mod_uvwp.f90:370-436
  subroutine update_uvwp(eqn, prop, geom, pc, ne, nf, nbf)
    use mod_properties
    type(properties_t) :: prop
    type(uvwp_t) :: eqn
    type(geometry_t) :: geom
    integer :: ne, nf, nbf
    real, dimension(ne+nbf) :: pc
    if (.not. c_associated(eqn%ctx)) then
      write(*,*) 'mod_uvwp: update_uvwp before any other uvwp routine'
      error stop 1
    end if
    call cfdl_check(cfdl_set_properties(eqn%ctx, prop%rho, prop%mu), 'properties')
    call put_state(eqn)
    call cfdl_check(cfdl_set_field(eqn%ctx, CFDL_PHIC, pc, int(ne+nbf, c_int64_t)), 'update_uvwp')
    call cfdl_check(cfdl_update_uvwp(eqn%ctx), 'update_uvwp')
    call get_state(eqn)
  end subroutine

  ! calc_mip, mod_uvwp.f90:438-490
  subroutine calc_mip(eqn, prop, geom, lRhieChow, dt)
    use mod_properties
    type(uvwp_t) :: eqn
    type(geometry_t) :: geom
    type(properties_t) :: prop
    real :: dt
    logical :: lRhieChow
    call ensure_ctx(eqn, geom, prop, dt, max(eqn%ctx_nit, 0), max(eqn%ctx_nsubd, 1))
    call sync_params(eqn, prop, dt)
    call put_state(eqn)
    call cfdl_check(cfdl_calc_mip(eqn%ctx, merge(1_c_int, 0_c_int, lRhieChow)), 'calc_mip')
    call dev_get(eqn, CFDL_MIP, eqn%mip)
  end subroutine

  ! ------------------------------------------- BC setters (procedure ptrs) ----
  ! dirichlet0 / lid / symmetry, mod_uvwp.f90:493-570: the setter names the
  ! section's boundary type; the device applies the values in solve_uvwp.
  subroutine set_kind(bc, eqn, bctype, kind)
    class(bc_t) :: bc
    class(equation_t) :: eqn
    character(len=*) :: bctype
    integer :: kind
    bc%bc_type = bctype
    select type(eqn)
    type is(uvwp_t)
      eqn%bc_kind(bc%idx) = kind
    end select
  end subroutine

  subroutine dirichlet0(bc, geom, eqn, prop)
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    call set_kind(bc, eqn, 'dirichlet', CFDL_BC_WALL)
  end subroutine

  subroutine symmetry(bc, geom, eqn, prop)
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    call set_kind(bc, eqn, 'zero_flux', CFDL_BC_SYMMETRY)
  end subroutine

  subroutine lid(bc, geom, eqn, prop)
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    call set_kind(bc, eqn, 'dirichlet', CFDL_BC_LID)
  end subroutine

  ! inlet / outlet (mod_uvwp.f90:572-606) are bound nowhere in the reference
  ! and its assembly has no branch for their types (stale coefficients);
  ! they stop here instead of computing something different.
  subroutine inlet(bc, geom, eqn, prop)
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    write(*,*) 'mod_uvwp: the inlet setter has no assembly in the reference'
    error stop 1
  end subroutine

  subroutine outlet(bc, geom, eqn, prop)
    class(bc_t) :: bc
    type(geometry_t) :: geom
    class(equation_t) :: eqn
    type(properties_t) :: prop
    write(*,*) 'mod_uvwp: the outlet setter has no assembly in the reference'
    error stop 1
  end subroutine
end module mod_uvwp
