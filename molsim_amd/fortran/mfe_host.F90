! mfe_host.F90 -- a thin Fortran host that drives the engine the way
! MOLSIM's stages do: PotTwoBodyTab1 (tables), EwaldSetup/IONList (setup),
! SetVList, UTotal, then MDDriver steps.  It reads one system from a raw
! stream file and writes UTotal's forceo, energies and the positions after
! the MD steps.
!
!   mfe_host <input> <output>
!
! input  (little-endian stream): np, npt, ewald kind, order, mesh, ncut,
!        nsteps, precision (int32); boxlen, rcut, alpha, drnlist, tstep
!        (real64); iptpn(np) (int32); zat, sigat, epsat, massat (npt each);
!        ro(3,np), rod(3,np) (real64)
! output: forceo(3,np), tot, twob0, rec, virial, ro(3,np) after nsteps
program mfe_host

   use, intrinsic :: iso_c_binding
   use mfe_mod
   implicit none

   integer(c_int32_t) :: np, npt, kind, order, mesh, ncut, nsteps, prec
   real(c_double) :: boxlen, rcut, alpha, drnlist, tstep
   integer(c_int32_t), allocatable :: iptpn(:), iubuflow(:), nugrid(:)
   real(c_double), allocatable, target :: zat(:), sigat(:), epsat(:), massat(:)
   real(c_double), allocatable :: az(:), massipt(:), ro(:,:), rod(:,:), rodd(:,:), forceo(:,:)
   real(c_double), allocatable :: ubuf(:), r2umin(:), r2atat(:)
   type(mfe_pot_1_6_12) :: spec
   type(mfe_config) :: cfg
   type(mfe_energies) :: en
   type(c_ptr) :: eng
   integer(c_int) :: rc
   integer(c_int32_t) :: nbuf, nptpt, meshv(3)
   integer :: u, i
   character(len=512) :: fin, fout
   real(c_double), parameter :: Pi = 3.1415926535897932d0
   real(c_double) :: efp

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read(u) np, npt, kind, order, mesh, ncut, nsteps, prec
   read(u) boxlen, rcut, alpha, drnlist, tstep
   allocate(iptpn(np), zat(npt), sigat(npt), epsat(npt), massat(npt), az(np), massipt(npt))
   allocate(ro(3,np), rod(3,np), rodd(3,np), forceo(3,np))
   read(u) iptpn
   read(u) zat, sigat, epsat, massat
   read(u) ro, rod
   close(u)

   ! Epsi0FourPi / relpermitt (Src/molsim.F90:805, Src/potential.F90:415)
   efp = ((1.602176487d-19**2)/((4.0d0*Pi*8.854187817d-12)*1.0d-10))*6.02214179d23/1.0d3/78.4d0

   ! PotTwoBodyTab1 for the (1,6,12) potential
   rc = mfe_pot_defaults(spec)
   spec%npt = npt
   spec%zat = c_loc(zat)
   spec%sigat = c_loc(sigat)
   spec%epsat = c_loc(epsat)
   spec%epsi_four_pi = efp
   spec%rcut = rcut
   spec%alpha = alpha
   spec%lewald = merge(1, 0, kind /= MFE_EWALD_NONE)
   spec%lljcut = 1
   nptpt = npt*(npt+1)/2
   allocate(ubuf(200000), iubuflow(nptpt), nugrid(nptpt), r2umin(nptpt), r2atat(nptpt))
   r2atat = 0.0d0
   eng = c_null_ptr
   call mfe_check(eng, mfe_build_tables_1_6_12(spec, ubuf, 200000, nbuf, iubuflow, nugrid, r2umin), &
                  'mfe_build_tables_1_6_12')

   cfg = mfe_config(prec, 0, np, 0)
   call mfe_check(eng, mfe_create(cfg, eng), 'mfe_create')
   call mfe_check(eng, mfe_set_box(eng, [boxlen, boxlen, boxlen]), 'mfe_set_box')
   az = zat(iptpn)
   massipt = 1.0d0/massat
   call mfe_check(eng, mfe_set_particles(eng, np, npt, iptpn, az, massipt), 'mfe_set_particles')
   call mfe_check(eng, mfe_set_tables(eng, ubuf, nbuf, iubuflow, r2atat, r2umin, rcut), 'mfe_set_tables')
   meshv = mesh
   call mfe_check(eng, mfe_set_ewald(eng, kind, alpha, order, meshv, ncut, 0, 0, efp), 'mfe_set_ewald')
   call mfe_check(eng, mfe_set_neighbor(eng, MFE_NL_VLIST, drnlist, 0), 'mfe_set_neighbor')
   call mfe_check(eng, mfe_upload_positions(eng, ro), 'mfe_upload_positions')
   rodd = 0.0d0
   call mfe_check(eng, mfe_upload_velocities(eng, rod, rodd), 'mfe_upload_velocities')

   ! SetVList + UTotal
   call mfe_check(eng, mfe_build_list(eng), 'mfe_build_list')
   call mfe_check(eng, mfe_compute(eng, ior(MFE_FORCES, ior(MFE_ENERGY, MFE_VIRIAL)), forceo, en), &
                  'mfe_compute')
   ! MDDriver(iSimulationStep) x nsteps
   call mfe_check(eng, mfe_step_md(eng, nsteps, tstep, ior(MFE_FORCES, MFE_ENERGY)), 'mfe_step_md')
   call mfe_check(eng, mfe_download_state(eng, ro, rod, rodd), 'mfe_download_state')

   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u) forceo
   write(u) en%tot, en%twob0, en%rec, en%virial
   write(u) ro
   close(u)
   write(*,'(a,i8,a,es22.14)') 'mfe_host: np =', np, '  u%tot =', en%tot
   do i = 1, 0
   end do
   rc = mfe_destroy(eng)

end program mfe_host
