! drv_cell.f90 -- mesh input for the reference build used as an oracle.
!
! TEST INFRASTRUCTURE ONLY.  The reference reads its mesh through CGNS
! (src/setup/cell_input.f90:10-165, src/modules/mod_cgns.f90), a library this
! image does not have.  This module takes the place of cell_input.f90 as a
! whole: it defines module mod_cell (which the reference's own modules `use`)
! and a routine drv_input that reads the element lists and coordinates from a
! plain binary file and then calls the reference's own setup routines in the
! order cell_input.f90:69-163 calls them (up to the first-level partition
! transformation, see below).  No CGNS routine is emulated.
!
! File layout (stream, little endian): int32 nvx, nsec; per section
! char(32) name, int32 etype, int32 first, int32 last; real(8) x(nvx),
! y(nvx), z(nvx); int32 e2vx(8, nelem) (unused slots zero).
module mod_cell
  use mod_mg_lvl
  implicit none
contains
  subroutine drv_input(geom, fname, n_subdomains)
    use mod_calc_vol_cv_centers
    use mod_calc_aip_xyzip
    use mod_agglomeration
    use mod_util
    implicit none
    type(geometry_t) :: geom
    character(len=*) :: fname
    integer :: n_subdomains
    integer :: u, nvx, nsec, sec, et, e1, e2, ne, nf, nbf, nelem, ne2vxmax, vx2e_size
    integer :: cvs(3), p, npmax, mem_size(2), aiplength, length, vlength, nbndry
    integer, pointer :: nl(:)
    integer, allocatable :: tmp(:)
    character(len=32) :: secname

    open(newunit=u, file=fname, access='stream', form='unformatted', status='old')
    read(u) nvx, nsec
    associate(mg_lvl => geom%mg)
      allocate(mg_lvl%esec(2,nsec))
      allocate(mg_lvl%etype(nsec))
      allocate(mg_lvl%ne2vx(nsec))
      allocate(mg_lvl%sectionName(nsec))
      cvs = 0
      ne2vxmax = 0
      vx2e_size = 0
      do sec=1,nsec
        read(u) secname, et, e1, e2
        mg_lvl%sectionName(sec) = secname
        mg_lvl%etype(sec) = et
        mg_lvl%esec(1,sec) = e1
        mg_lvl%esec(2,sec) = e2
        if(et>=10 .and. et<=20) then
          cvs(1) = cvs(1) + e2-e1+1
          cvs(2) = cvs(2) + element_nface(et)*(e2-e1+1)
        else
          cvs(3) = cvs(3) + e2-e1+1
        endif
        mg_lvl%ne2vx(sec) = element_nvx(et)
        ne2vxmax = max(ne2vxmax, element_nvx(et))
        vx2e_size = vx2e_size + element_nvx(et)*(e2-e1+1)
      end do
      cvs(2) = (cvs(2) + cvs(3))/2
      ne = cvs(1); nf = cvs(2); nbf = cvs(3)
      nelem = ne+nbf
      nbndry = nbf
      allocate(mg_lvl%e2vx(ne2vxmax*nelem))
      mg_lvl%nsec = nsec
      mg_lvl%ne2vx_max = ne2vxmax
      mg_lvl%nvx = nvx
      mg_lvl%nelem = nelem
      mg_lvl%nbndry = nbndry
      mg_lvl%nfaces = nf
      allocate(geom%x(nvx), geom%y(nvx), geom%z(nvx))
      read(u) geom%x
      read(u) geom%y
      read(u) geom%z
      allocate(tmp(8*nelem))
      read(u) tmp
      close(u)
      if (ne2vxmax /= 8) stop 'drv_input: expects hexahedral cells'
      mg_lvl%e2vx = tmp
      deallocate(tmp)

      call mg_lvl%add_meshds(nsec,geom%x,geom%y,geom%z,vx2e_size)

      aiplength = nf*3
      allocate(geom%rip(aiplength))
      allocate(geom%aip(aiplength))
      call calc_aip_xyzip_uns(geom%x,geom%y,geom%z,geom%rip,geom%aip,mg_lvl%e2vx,mg_lvl%etype,nelem,nbndry,nvx,&
                              ne2vxmax,mg_lvl%fine_lvl%s2g,nsec,mg_lvl%esec,nf)
      length = nelem
      allocate(geom%xc(length), geom%yc(length), geom%zc(length))
      vlength = nelem-nbndry
      allocate(geom%vol(vlength))
      call calc_vol_cv_centers_uns(geom%xc,geom%yc,geom%zc,geom%vol,geom%x,geom%y,geom%z,mg_lvl%e2vx,&
             mg_lvl%fine_lvl%gs2nb,mg_lvl%fine_lvl%gs2nb_idx,mg_lvl%etype,nvx,nelem,nbndry,nf,ne2vxmax,nsec,&
             mg_lvl%esec,geom%rip,geom%aip)

      npmax = 5
      allocate(nl(npmax))
      nl = 0
      do p=1,npmax
        nl(p) = int((nelem-nbndry)/(8**p))
        if(nl(p)==0) exit
      enddo
      npmax = p-1
      if(n_subdomains>1) then
        npmax = 2
        nl = 0
        nl(1) = n_subdomains
        nl(2) = 1
      endif
      call mg_lvl%generate_seeds(nl,npmax,ne,geom%vol,geom%xc,geom%yc,geom%zc)
      ! cell_input.f90:144-151 loops over npmax levels, each followed by a
      ! coarse add_meshds.  The coarse add_meshds assigns ns elements to an
      ! array of ns+1 (mod_mg_lvl_uns.f90:219), which flang's runtime rejects;
      ! no hot-path routine reads the coarse levels, so only the first-level
      ! transformation (g2gf(1), the subdomain-local order) is built here.
      mem_size = mg_lvl%add_transformation_bt()
    end associate

    geom%ef2nb => geom%mg%fine_lvl%gs2nb
    geom%ef2nb_idx => geom%mg%fine_lvl%gs2nb_idx
    geom%ne = geom%mg%fine_lvl%ng
    geom%nf = geom%mg%fine_lvl%ns
    geom%nbf = geom%mg%fine_lvl%nbs
    geom%nvx = geom%mg%nvx
  end subroutine
end module
