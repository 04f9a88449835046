! drv_main.f90 -- driver program for the reference build used as an oracle.
!
! TEST INFRASTRUCTURE ONLY.  Stands where src/main.f90 stands: it sets the
! physics parameters a case needs (the reference hard-codes rho=5, mu=0.01,
! src/modules/mod_properties.f90:86-87), runs the same loop nesting as
! src/main.f90:50-63 through the reference's own mod_physics / mod_uvwp,
! times the outer iterations and dumps setup arrays and final fields.
!
! usage: refsim mesh.bin out.bin nsubd rho mu dt ntstep ncoef nit [project/]
! With the optional project path the run ends like src/main.f90:89, with the
! reference's write_vtubin of the uvwp fields into project/VTK/output-N/.
! With REFSIM_PROPS=<file> the per-cell density and viscosity come from that
! file (ne float64 values of rho, then ne of mu) instead of the constants.
! With REFSIM_SYMMETRY='<name> <name> ...' the uvwp setters of those boundary
! sections are rebound to the reference's symmetry (mod_uvwp.f90:517-546),
! which src/equations/mod_uvwp.f90:73-78 never binds.
! With REFSIM_ENERGY=1 in the environment every outer iteration also calls
! the reference's solve_energy (src/equations/mod_energy.f90:58-80), which
! src/main.f90 never calls, and the energy fields are appended to the dump.
program refsim
  use mod_cell
  use mod_physics
  use mod_solver
  use mod_vtu_output
  use mod_uvwp, only: symmetry
  implicit none
  type(geometry_t) :: geom
  type(phys_t) :: phys
  character(len=512) :: meshfile, outfile, arg
  character(len=180) :: projpath
  integer :: nsubd, ntstep, ncoef, nit, tstep, icoef, u, c, cnb, ne, nbf
  real :: rho, mu, dt
  integer(8) :: t0, t1, rate
  character(len=8) :: envval
  character(len=512) :: propfile
  integer :: plen, up, ib
  character(len=512) :: symlist
  logical :: energy

  call get_command_argument(1, meshfile)
  call get_command_argument(2, outfile)
  call get_command_argument(3, arg); read(arg,*) nsubd
  call get_command_argument(4, arg); read(arg,*) rho
  call get_command_argument(5, arg); read(arg,*) mu
  call get_command_argument(6, arg); read(arg,*) dt
  call get_command_argument(7, arg); read(arg,*) ntstep
  call get_command_argument(8, arg); read(arg,*) ncoef
  call get_command_argument(9, arg); read(arg,*) nit

  call get_environment_variable('REFSIM_ENERGY', envval)
  energy = trim(envval) == '1'
  phys%n_subdomains = nsubd
  phys%dt = dt
  phys%ntstep = ntstep
  phys%ncoef = ncoef
  phys%nit = nit
  call drv_input(geom, meshfile, phys%n_subdomains)
  call construct_physics(phys, geom)
  phys%prop%rho = rho
  phys%prop%mu = mu
  ne = geom%ne
  nbf = geom%nbf
  call get_environment_variable('REFSIM_PROPS', propfile, plen)
  if (plen > 0) then
    open(newunit=up, file=propfile(1:plen), access='stream', form='unformatted', status='old')
    read(up) phys%prop%rho(1:ne)
    read(up) phys%prop%mu(1:ne)
    close(up)
  end if
  call get_environment_variable('REFSIM_SYMMETRY', symlist)
  do ib=1,size(phys%uvwp%bcs)
    if (index(' '//trim(symlist)//' ', ' '//trim(phys%uvwp%bcs(ib)%name)//' ') > 0) &
      phys%uvwp%bcs(ib)%coef => symmetry
  end do

  open(newunit=u, file=outfile, access='stream', form='unformatted', status='replace')
  write(u) geom%ne, geom%nf, geom%nbf, phys%n_subdomains
  write(u) geom%ef2nb_idx, geom%ef2nb(:,1), geom%ef2nb(:,2), geom%mg%fine_lvl%s2g(1:geom%nf)
  write(u) geom%mg%fine_lvl%bs(ne+1:ne+nbf)
  write(u) geom%xc, geom%yc, geom%zc, geom%vol, geom%aip, geom%rip
  if (nsubd > 1) then
    write(u) geom%mg%gf2g(1)%p(1:ne), geom%mg%g2gf(1)%p(1:ne), geom%mg%g2gf(1)%idx(1:nsubd+1)
    do c=1,nsubd
      write(u) phys%subdomain(c)%ne, phys%subdomain(c)%nf, phys%subdomain(c)%nbf
      write(u) phys%subdomain(c)%ef2nb_idx, phys%subdomain(c)%ef2nb
      do cnb=1,nsubd
        write(u) phys%intf(c,cnb)%ncs
        write(u) phys%intf(c,cnb)%index1
      end do
    end do
  end if

  write(*,'(A)') 'refsim: start'
  call system_clock(t0, rate)
  do tstep=1,phys%ntstep
    do icoef=1,phys%ncoef
      call update_boundaries(phys,geom)
      call solve_uvwp(phys%uvwp,phys%prop,geom,phys%dt,phys%nit,phys%ap,phys%anb,phys%b,phys%phic, &
                      phys%subdomain,phys%intf,phys%n_subdomains)
      if (energy) call solve_energy(phys%energy,phys%prop,geom,phys%dt,phys%nit,phys%ap,phys%anb, &
                                    phys%b,phys%phic)
    end do
    call update_time(phys)
  end do
  call system_clock(t1)
  write(*,'(A,ES24.16)') 'refsim: seconds ', real(t1-t0)/real(rate)

  associate(q => phys%uvwp)
    write(u) q%u, q%v, q%w, q%p, q%u0, q%v0, q%w0
    write(u) q%gu, q%gv, q%gw, q%gp, q%gpc(1:3*ne)
    write(u) q%mip, q%mip0, q%bu, q%bv, q%bw, q%d, q%dc
  end associate
  write(u) phys%ap, phys%b, phys%anb, phys%phic
  if (energy) then
    associate(q => phys%energy)
      write(u) q%t, q%gt(1:3*ne), q%phi, q%phi0(1:ne), q%grad(1:3*ne)
    end associate
  end if
  close(u)
  if (command_argument_count() >= 10) then
    call get_command_argument(10, projpath)
    call write_vtubin(phys%uvwp, geom, projpath, tstep)
  end if
end program
