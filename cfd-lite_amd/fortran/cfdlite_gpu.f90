! cfdlite_gpu -- Fortran host program of the GPU build.
!
! The loop of the reference's main program (src/main.f90:50-63) over a
! generated hexahedral cavity: per outer iteration the boundary setters and
! one SIMPLE(C) iteration, per pseudo-time step the time-level shift, all
! through the iso_c_binding layer (cfdl_c) into libcfdl_hip.so.  Prints the
! solver lines in the reference's format and can dump the final fields.
!
! usage: cfdlite_gpu N nsubd rho mu dt ntstep ncoef [alpha_u alpha_p [dump.bin]]
program cfdlite_gpu
  use cfdl_c
  implicit none
  character(len=*), parameter :: oformat = "(5x,A16,x,i5,x,15x,es9.3e2,3x,es9.3e2,3x,es9.3e2)"
  type(c_ptr) :: mesh, ctx
  character(len=256) :: arg, dumpfile
  integer(c_int) :: n, nsubd, ntstep, ncoef, ne, nf, nbf, nsec, nsd, nvx, it, nlog, k, j
  real(c_double) :: rho, mu, dt, au, apr, ri, rf, rm
  real(c_double), allocatable :: a(:)
  character(kind=c_char) :: nm(8)
  character(len=16) :: name
  integer :: tstep, icoef, u, f, logged
  integer(c_int), parameter :: dump_fields(7) = [CFDL_U, CFDL_V, CFDL_W, CFDL_P, CFDL_MIP, &
                                                 CFDL_GP, CFDL_PHIC]

  call get_command_argument(1, arg); read(arg,*) n
  call get_command_argument(2, arg); read(arg,*) nsubd
  call get_command_argument(3, arg); read(arg,*) rho
  call get_command_argument(4, arg); read(arg,*) mu
  call get_command_argument(5, arg); read(arg,*) dt
  call get_command_argument(6, arg); read(arg,*) ntstep
  call get_command_argument(7, arg); read(arg,*) ncoef
  au = 1.0_c_double
  apr = 1.0_c_double
  dumpfile = ' '
  if (command_argument_count() >= 9) then
    call get_command_argument(8, arg); read(arg,*) au
    call get_command_argument(9, arg); read(arg,*) apr
  end if
  if (command_argument_count() >= 10) call get_command_argument(10, dumpfile)

  call cfdl_check(cfdl_mesh_box(n, n, n, 1.0_c_double, 1.0_c_double, 1.0_c_double, 0.0_c_double, 20261202_c_int64_t, mesh), 'mesh')
  call cfdl_check(cfdl_mesh_partition(mesh, nsubd), 'partition')
  call cfdl_check(cfdl_mesh_info(mesh, ne, nf, nbf, nsec, nsd, nvx), 'info')
  call cfdl_check(cfdl_ctx_create_cavity(mesh, rho, mu, dt, 100_c_int, ctx), 'context')
  call cfdl_check(cfdl_set_relaxation(ctx, au, apr), 'relaxation')
  write(*,'(A,I0,A,I0,A,I0)') 'cfdlite_gpu: cells ', ne, ' faces ', nf, ' subdomains ', nsubd
  write(*,'(5x,A16,x,A5,x,A15,A9,3x,A9,3x,A9)') 'solve eqn       ','nit','residual(rms)','initial','final','max'
  logged = 0
  do tstep = 1, ntstep
    do icoef = 1, ncoef
      call cfdl_check(cfdl_update_boundaries(ctx), 'update_boundaries')
      call cfdl_check(cfdl_solve_uvwp(ctx), 'solve_uvwp')
      call cfdl_check(cfdl_log_count(ctx, nlog), 'log')
      do k = logged, nlog - 1
        call cfdl_check(cfdl_log_get(ctx, k, nm, it, ri, rf, rm), 'log')
        name = ' '
        do j = 1, 8
          if (nm(j) == c_null_char) exit
          name(j:j) = nm(j)
        end do
        write(*,oformat) name, it, ri, rf, rm
      end do
      logged = nlog
    end do
    call cfdl_check(cfdl_update_time(ctx), 'update_time')
    write(*,'(A,i5,A)') '------------------------------------time step(', tstep, ')'
  end do
  if (len_trim(dumpfile) > 0) then
    open(newunit=u, file=trim(dumpfile), access='stream', form='unformatted', status='replace')
    do f = 1, size(dump_fields)
      select case (dump_fields(f))
      case (CFDL_MIP)
        allocate(a(nf))
      case (CFDL_GP)
        allocate(a(3*(ne+nbf)))
      case default
        allocate(a(ne+nbf))
      end select
      call cfdl_check(cfdl_get_field(ctx, dump_fields(f), a, int(size(a), c_int64_t)), 'download')
      write(u) a
      deallocate(a)
    end do
    close(u)
  end if
  call cfdl_ctx_destroy(ctx)
  call cfdl_mesh_free(mesh)
end program
