! cfdl_c.f90 -- iso_c_binding interface to libcfdl_hip.so (include/cfdl.h).
!
! The thin layer a Fortran host binds: every C entry point with its exact
! C signature.  Reals are c_double (the reference's real under
! -fdefault-real-8), integers c_int, arrays are passed as the host stores
! them.  Strings go in null-terminated (cfdl_str).
module cfdl_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CFDL_OK = 0
  integer(c_int), parameter :: CFDL_BC_WALL = 0, CFDL_BC_LID = 1, CFDL_BC_SYMMETRY = 2
  integer(c_int), parameter :: CFDL_PC_REFERENCE = 0, CFDL_PC_COLORED = 1, &
                                CFDL_PC_COLORED_MIXED = 2
  integer(c_int), parameter :: CFDL_GS_REFERENCE = 0, CFDL_GS_COLORED = 1
  integer(c_int), parameter :: CFDL_U = 0, CFDL_V = 1, CFDL_W = 2, CFDL_P = 3, &
       CFDL_U0 = 4, CFDL_V0 = 5, CFDL_W0 = 6, CFDL_GU = 7, CFDL_GV = 8, CFDL_GW = 9, &
       CFDL_GP = 10, CFDL_GPC = 11, CFDL_MIP = 12, CFDL_MIP0 = 13, CFDL_BU = 14, &
       CFDL_BV = 15, CFDL_BW = 16, CFDL_D = 17, CFDL_DC = 18, CFDL_AP = 19, CFDL_B = 20, &
       CFDL_ANB = 21, CFDL_PHIC = 22, CFDL_WORK = 23, CFDL_GWORK = 24, &
       CFDL_T = 25, CFDL_GT = 26, CFDL_ENTH = 27, CFDL_ENTH0 = 28, CFDL_GENTH = 29

  interface
    function cfdl_last_error() bind(C, name="cfdl_last_error")
      import :: c_ptr
      type(c_ptr) :: cfdl_last_error
    end function
    function cfdl_version() bind(C, name="cfdl_version")
      import :: c_int
      integer(c_int) :: cfdl_version
    end function

    ! ---- host-side setup ----
    function cfdl_mesh_box(nx, ny, nz, lx, ly, lz, jitter, seed, out) bind(C, name="cfdl_mesh_box")
      import :: c_int, c_double, c_int64_t, c_ptr
      integer(c_int), value :: nx, ny, nz
      real(c_double), value :: lx, ly, lz, jitter
      integer(c_int64_t), value :: seed
      type(c_ptr), intent(out) :: out
      integer(c_int) :: cfdl_mesh_box
    end function
    function cfdl_mesh_from_elements(nvx, x, y, z, nsec, etype, esec, e2vx, out) &
        bind(C, name="cfdl_mesh_from_elements")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nvx, nsec
      real(c_double), intent(in) :: x(*), y(*), z(*)
      integer(c_int), intent(in) :: etype(*), esec(*), e2vx(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: cfdl_mesh_from_elements
    end function
    function cfdl_mesh_write_vtu(m, path, nvar, names, ncomp, data) bind(C, name="cfdl_mesh_write_vtu")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: m
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nvar
      type(c_ptr), intent(in) :: names(*), data(*)
      integer(c_int), intent(in) :: ncomp(*)
      integer(c_int) :: cfdl_mesh_write_vtu
    end function
    function cfdl_mesh_partition(m, nsubd) bind(C, name="cfdl_mesh_partition")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nsubd
      integer(c_int) :: cfdl_mesh_partition
    end function
    function cfdl_mesh_info(m, ne, nf, nbf, nsec, nsubd, nvx) bind(C, name="cfdl_mesh_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), intent(out) :: ne, nf, nbf, nsec, nsubd, nvx
      integer(c_int) :: cfdl_mesh_info
    end function
    function cfdl_mesh_get_i32(m, name, dst, n) bind(C, name="cfdl_mesh_get_i32")
      import :: c_int, c_int64_t, c_ptr, c_char
      type(c_ptr), value :: m
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: dst(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: cfdl_mesh_get_i32
    end function
    function cfdl_mesh_get_f64(m, name, dst, n) bind(C, name="cfdl_mesh_get_f64")
      import :: c_int, c_int64_t, c_ptr, c_char, c_double
      type(c_ptr), value :: m
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: dst(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: cfdl_mesh_get_f64
    end function
    subroutine cfdl_mesh_free(m) bind(C, name="cfdl_mesh_free")
      import :: c_ptr
      type(c_ptr), value :: m
    end subroutine

    ! ---- device context ----
    function cfdl_ctx_create(ne, nf, nbf, ef2nb_idx, ef2nb, s2g, bs, xc, yc, zc, vol, aip, rip, &
        nsec, esec, bckind, nsubd, gf2g, g2gf, g2gf_idx, rho, mu, dt, nit, out) &
        bind(C, name="cfdl_ctx_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ne, nf, nbf, nsec, nsubd, nit
      integer(c_int), intent(in) :: ef2nb_idx(*), ef2nb(*), s2g(*), bs(*), esec(*), bckind(*)
      integer(c_int), intent(in) :: gf2g(*), g2gf(*), g2gf_idx(*)
      real(c_double), intent(in) :: xc(*), yc(*), zc(*), vol(*), aip(*), rip(*), rho(*), mu(*)
      real(c_double), value :: dt
      type(c_ptr), intent(out) :: out
      integer(c_int) :: cfdl_ctx_create
    end function
    function cfdl_ctx_create_cavity(m, rho, mu, dt, nit, out) bind(C, name="cfdl_ctx_create_cavity")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: m
      real(c_double), value :: rho, mu, dt
      integer(c_int), value :: nit
      type(c_ptr), intent(out) :: out
      integer(c_int) :: cfdl_ctx_create_cavity
    end function
    subroutine cfdl_ctx_destroy(c) bind(C, name="cfdl_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: c
    end subroutine
    function cfdl_set_relaxation(c, au, ap) bind(C, name="cfdl_set_relaxation")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: c
      real(c_double), value :: au, ap
      integer(c_int) :: cfdl_set_relaxation
    end function
    function cfdl_set_pc_order(c, order) bind(C, name="cfdl_set_pc_order")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: order
      integer(c_int) :: cfdl_set_pc_order
    end function
    function cfdl_set_gs_order(c, order) bind(C, name="cfdl_set_gs_order")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: order
      integer(c_int) :: cfdl_set_gs_order
    end function
    function cfdl_set_properties(c, rho, mu) bind(C, name="cfdl_set_properties")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: c
      real(c_double), intent(in) :: rho(*), mu(*)
      integer(c_int) :: cfdl_set_properties
    end function
    function cfdl_set_dt(c, dt) bind(C, name="cfdl_set_dt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: c
      real(c_double), value :: dt
      integer(c_int) :: cfdl_set_dt
    end function
    function cfdl_set_bc_kind(c, first, last, kind) bind(C, name="cfdl_set_bc_kind")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: first, last, kind
      integer(c_int) :: cfdl_set_bc_kind
    end function
    function cfdl_set_field(c, field, src, n) bind(C, name="cfdl_set_field")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: field
      real(c_double), intent(in) :: src(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: cfdl_set_field
    end function
    function cfdl_get_field(c, field, dst, n) bind(C, name="cfdl_get_field")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: field
      real(c_double), intent(out) :: dst(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: cfdl_get_field
    end function
    function cfdl_synchronize(c) bind(C, name="cfdl_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_synchronize
    end function

    ! ---- one entry point per reference routine ----
    function cfdl_update_boundaries(c) bind(C, name="cfdl_update_boundaries")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_update_boundaries
    end function
    function cfdl_calc_coef_uvw(c) bind(C, name="cfdl_calc_coef_uvw")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_calc_coef_uvw
    end function
    function cfdl_solve_gs(c, name, field, rhs) bind(C, name="cfdl_solve_gs")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), value :: field, rhs
      integer(c_int) :: cfdl_solve_gs
    end function
    function cfdl_solve_gs_uvw(c) bind(C, name="cfdl_solve_gs_uvw")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_solve_gs_uvw
    end function
    function cfdl_calc_grad(c, src, dst) bind(C, name="cfdl_calc_grad")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: src, dst
      integer(c_int) :: cfdl_calc_grad
    end function
    function cfdl_calc_mip(c, rc) bind(C, name="cfdl_calc_mip")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: rc
      integer(c_int) :: cfdl_calc_mip
    end function
    function cfdl_calc_coef_p(c) bind(C, name="cfdl_calc_coef_p")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_calc_coef_p
    end function
    function cfdl_set_a_0(c, field) bind(C, name="cfdl_set_a_0")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: field
      integer(c_int) :: cfdl_set_a_0
    end function
    function cfdl_solve(c, name) bind(C, name="cfdl_solve")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int) :: cfdl_solve
    end function
    function cfdl_adjust_pc(c) bind(C, name="cfdl_adjust_pc")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_adjust_pc
    end function
    function cfdl_update_uvwp(c) bind(C, name="cfdl_update_uvwp")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_update_uvwp
    end function
    function cfdl_solve_uvwp(c) bind(C, name="cfdl_solve_uvwp")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_solve_uvwp
    end function
    function cfdl_update_time(c) bind(C, name="cfdl_update_time")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_update_time
    end function
    function cfdl_calc_mass_imbalance(c, dst) bind(C, name="cfdl_calc_mass_imbalance")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: dst
      integer(c_int) :: cfdl_calc_mass_imbalance
    end function

    ! ---- energy equation (src/equations/mod_energy.f90) ----
    function cfdl_set_thermal(c, cp, tc) bind(C, name="cfdl_set_thermal")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: c
      real(c_double), intent(in) :: cp(*), tc(*)
      integer(c_int) :: cfdl_set_thermal
    end function
    function cfdl_construct_energy(c) bind(C, name="cfdl_construct_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_construct_energy
    end function
    function cfdl_update_boundaries_energy(c) bind(C, name="cfdl_update_boundaries_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_update_boundaries_energy
    end function
    function cfdl_calc_coef_energy(c) bind(C, name="cfdl_calc_coef_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_calc_coef_energy
    end function
    function cfdl_solve_energy(c) bind(C, name="cfdl_solve_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_solve_energy
    end function
    function cfdl_run_energy(c, ntstep, ncoef) bind(C, name="cfdl_run_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: ntstep, ncoef
      integer(c_int) :: cfdl_run_energy
    end function
    function cfdl_run(c, ntstep, ncoef) bind(C, name="cfdl_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: ntstep, ncoef
      integer(c_int) :: cfdl_run
    end function
    function cfdl_log_count(c, n) bind(C, name="cfdl_log_count")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), intent(out) :: n
      integer(c_int) :: cfdl_log_count
    end function
    function cfdl_log_get(c, k, name, it, res_i, res_f, res_max) bind(C, name="cfdl_log_get")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: c
      integer(c_int), value :: k
      character(kind=c_char), intent(out) :: name(8)
      integer(c_int), intent(out) :: it
      real(c_double), intent(out) :: res_i, res_f, res_max
      integer(c_int) :: cfdl_log_get
    end function
    function cfdl_log_clear(c) bind(C, name="cfdl_log_clear")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_log_clear
    end function
    function cfdl_set_profile(c, every) bind(C, name="cfdl_set_profile")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: every
      integer(c_int) :: cfdl_set_profile
    end function
    function cfdl_stats(c, stats, n) bind(C, name="cfdl_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: stats(*)
      integer(c_int), value :: n
      integer(c_int) :: cfdl_stats
    end function
    function cfdl_stats_reset(c) bind(C, name="cfdl_stats_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: cfdl_stats_reset
    end function
  end interface

contains

  function cfdl_str(s) result(cs)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: cs
    cs = trim(s)//c_null_char
  end function

  ! The reference stops the program on errors (mod_eqn_setup.f90:63-66);
  ! so does the binding, with the library's message.
  subroutine cfdl_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == CFDL_OK) return
    call c_f_pointer(cfdl_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*,'(A,A,A,I0,A,512A)') 'cfdl: ', what, ' failed (', rc, '): ', msg(1:n)
    error stop 1
  end subroutine
end module cfdl_c
