! mfe_mod.F90 -- ISO_C_BINDING interface of include/mfe.h for a Fortran
! host such as MOLSIM.  MOLSIM's own arrays pass straight through:
! ro(3,np), force(3,na), iptpn(np), az(na), ubuf(nbuf), iubuflow(natat)
! (Src/mol.F90:1087-1135) have the layout the C ABI expects.
module mfe_mod

   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: MFE_OK = 0
   integer(c_int), parameter :: MFE_PREC_FP64 = 0, MFE_PREC_MIXED = 1
   integer(c_int), parameter :: MFE_EWALD_NONE = 0, MFE_EWALD_STD = 1, MFE_EWALD_SPME = 2
   integer(c_int), parameter :: MFE_NL_VLIST = 0, MFE_NL_CELL = 1
   integer(c_int), parameter :: MFE_FORCES = 1, MFE_ENERGY = 2, MFE_VIRIAL = 4
   integer, parameter :: MFE_MAX_PAIRS = 36

   type, bind(c) :: mfe_config
      integer(c_int32_t) :: precision, device, max_np, reserved
   end type mfe_config

   type, bind(c) :: mfe_ewald_opt   ! EwaldSetup's iewaldopt inputs and outputs
      integer(c_int32_t) :: iewaldopt, spme, lq2sum, ncut
      real(c_double) :: q2sum, boxlenshort, epsi_four_pi
      real(c_double) :: uewaldtol, ualphared, ualpha, rcut
   end type mfe_ewald_opt

   type, bind(c) :: mfe_energies   ! struct mfe_energy
      real(c_double) :: tot, twob0, rec, bond, virial
      integer(c_int32_t) :: nptpt, rebuilds
      real(c_double) :: twob(MFE_MAX_PAIRS)
   end type mfe_energies

   type, bind(c) :: mfe_pot_1_6_12
      integer(c_int32_t) :: npt
      type(c_ptr) :: zat, sigat, epsat
      real(c_double) :: epsi_four_pi, rcut, alpha
      integer(c_int32_t) :: lewald, lljcut
      real(c_double) :: ljrcut, ljushift
      real(c_double) :: r2uminin, utoltab, ftoltab, umaxtab, fmaxtab, rumaxfac
   end type mfe_pot_1_6_12

   interface

      integer(c_int) function mfe_abi_version() bind(c, name='mfe_abi_version')
         import :: c_int
      end function

      integer(c_int) function mfe_pot_defaults(spec) bind(c, name='mfe_pot_defaults')
         import :: c_int, mfe_pot_1_6_12
         type(mfe_pot_1_6_12), intent(inout) :: spec
      end function

      integer(c_int) function mfe_build_tables_1_6_12(spec, ubuf, cap, nbuf_used, iubuflow, &
            nugrid, r2umin) bind(c, name='mfe_build_tables_1_6_12')
         import :: c_int, c_int32_t, c_double, mfe_pot_1_6_12
         type(mfe_pot_1_6_12), intent(in) :: spec
         real(c_double), intent(out) :: ubuf(*)
         integer(c_int32_t), value :: cap
         integer(c_int32_t), intent(out) :: nbuf_used
         integer(c_int32_t), intent(out) :: iubuflow(*), nugrid(*)
         real(c_double), intent(out) :: r2umin(*)
      end function

      integer(c_int) function mfe_create(cfg, eng) bind(c, name='mfe_create')
         import :: c_int, c_ptr, mfe_config
         type(mfe_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: eng
      end function

      integer(c_int) function mfe_destroy(eng) bind(c, name='mfe_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
      end function

      type(c_ptr) function mfe_last_error(eng) bind(c, name='mfe_last_error')
         import :: c_ptr
         type(c_ptr), value :: eng
      end function

      integer(c_int) function mfe_set_box(eng, boxlen) bind(c, name='mfe_set_box')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: boxlen(3)
      end function

      integer(c_int) function mfe_set_particles(eng, np, npt, iptpn, az, massipt) &
            bind(c, name='mfe_set_particles')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: np, npt
         integer(c_int32_t), intent(in) :: iptpn(*)
         real(c_double), intent(in) :: az(*), massipt(*)
      end function

      integer(c_int) function mfe_set_tables(eng, ubuf, nbuf, iubuflow, r2atat, r2umin, rcut) &
            bind(c, name='mfe_set_tables')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: ubuf(*)
         integer(c_int32_t), value :: nbuf
         integer(c_int32_t), intent(in) :: iubuflow(*)
         real(c_double), intent(in) :: r2atat(*), r2umin(*)
         real(c_double), value :: rcut
      end function

      integer(c_int) function mfe_set_ewald(eng, kind, alpha, order, mesh, ncut, ncut_cube, &
            lsurf, epsi_four_pi) bind(c, name='mfe_set_ewald')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: kind
         real(c_double), value :: alpha
         integer(c_int32_t), value :: order
         integer(c_int32_t), intent(in) :: mesh(3)
         integer(c_int32_t), value :: ncut, ncut_cube, lsurf
         real(c_double), value :: epsi_four_pi
      end function

      integer(c_int) function mfe_set_neighbor(eng, kind, drnlist, maxnneigh) &
            bind(c, name='mfe_set_neighbor')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: kind
         real(c_double), value :: drnlist
         integer(c_int32_t), value :: maxnneigh
      end function

      integer(c_int) function mfe_set_chains(eng, nchain, chain_first, chain_len, k, p, eq) &
            bind(c, name='mfe_set_chains')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: nchain
         integer(c_int32_t), intent(in) :: chain_first(*), chain_len(*)
         real(c_double), value :: k
         integer(c_int32_t), value :: p
         real(c_double), value :: eq
      end function

      integer(c_int) function mfe_set_angles(eng, k, p, eq) bind(c, name='mfe_set_angles')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), value :: k
         integer(c_int32_t), value :: p
         real(c_double), value :: eq
      end function

      integer(c_int) function mfe_get_angle_energy(eng, u_angle) &
            bind(c, name='mfe_get_angle_energy')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: u_angle
      end function

      integer(c_int) function mfe_upload_positions(eng, ro) bind(c, name='mfe_upload_positions')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: ro(3,*)
      end function

      integer(c_int) function mfe_upload_velocities(eng, rod, rodd) &
            bind(c, name='mfe_upload_velocities')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: rod(3,*), rodd(3,*)
      end function

      integer(c_int) function mfe_download_state(eng, ro, rod, rodd) &
            bind(c, name='mfe_download_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: ro(3,*), rod(3,*), rodd(3,*)
      end function

      integer(c_int) function mfe_build_list(eng) bind(c, name='mfe_build_list')
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
      end function

      integer(c_int) function mfe_compute(eng, flags, forceo, energy) bind(c, name='mfe_compute')
         import :: c_int, c_int32_t, c_ptr, c_double, mfe_energies
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: flags
         real(c_double), intent(out) :: forceo(3,*)
         type(mfe_energies), intent(out) :: energy
      end function

      integer(c_int) function mfe_step_md(eng, nsteps, tstep, flags) bind(c, name='mfe_step_md')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: nsteps
         real(c_double), value :: tstep
         integer(c_int32_t), value :: flags
      end function

      integer(c_int) function mfe_step_bd(eng, nsteps, fac1, fac2, noise, flags) &
            bind(c, name='mfe_step_bd')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: nsteps
         real(c_double), intent(in) :: fac1(*), fac2(*), noise(*)
         integer(c_int32_t), value :: flags
      end function

      integer(c_int) function mfe_get_energy(eng, energy) bind(c, name='mfe_get_energy')
         import :: c_int, c_ptr, mfe_energies
         type(c_ptr), value :: eng
         type(mfe_energies), intent(out) :: energy
      end function

      integer(c_int) function mfe_get_pairs(eng, nneighpn, jpnlist, maxnneigh) &
            bind(c, name='mfe_get_pairs')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: eng
         integer(c_int32_t), intent(out) :: nneighpn(*), jpnlist(*)
         integer(c_int32_t), value :: maxnneigh
      end function

      integer(c_int) function mfe_time_stage(eng, stage, nrep, ms) bind(c, name='mfe_time_stage')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: stage, nrep
         real(c_double), intent(out) :: ms
      end function

      ! BD with the device noise stream: pass c_null_ptr for fac1/fac2
      ! (factors of mfe_set_bd) and for noise
      integer(c_int) function mfe_step_bd_ptr(eng, nsteps, fac1, fac2, noise, flags) &
            bind(c, name='mfe_step_bd')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: eng, fac1, fac2, noise
         integer(c_int32_t), value :: nsteps
         integer(c_int32_t), value :: flags
      end function

      integer(c_int) function mfe_set_bd(eng, fac1, fac2, seed) bind(c, name='mfe_set_bd')
         import :: c_int, c_int64_t, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: fac1(*), fac2(*)
         integer(c_int64_t), value :: seed
      end function

      integer(c_int) function mfe_download_forces(eng, forceo) bind(c, name='mfe_download_forces')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: forceo(*)
      end function

      integer(c_int) function mfe_set_pair_params(eng, spec) bind(c, name='mfe_set_pair_params')
         import :: c_int, c_ptr, mfe_pot_1_6_12
         type(c_ptr), value :: eng
         type(mfe_pot_1_6_12), intent(in) :: spec
      end function

      integer(c_int) function mfe_set_knots(eng, nknot, r_lo, r_hi, c) bind(c, name='mfe_set_knots')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: nknot
         real(c_double), value :: r_lo, r_hi
         real(c_double), intent(in) :: c(*)
      end function

      integer(c_int) function mfe_adjoint(eng, dq, dzat, dsigat, depsat, dknot, dro, energy) &
            bind(c, name='mfe_adjoint')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: dq(*), dzat(*), dsigat(*), depsat(*), dknot(*), dro(*)
         real(c_double), intent(out) :: energy
      end function

      ! force matching: model forces at the current positions, frames
      ! ro(3,np,nframes) with targets ft(3,np,nframes), loss and gradient
      integer(c_int) function mfe_model_forces(eng, force) bind(c, name='mfe_model_forces')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: force(*)
      end function

      integer(c_int) function mfe_fm_set_frames(eng, nframes, ro, ft) &
            bind(c, name='mfe_fm_set_frames')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: nframes
         real(c_double), intent(in) :: ro(*), ft(*)
      end function

      integer(c_int) function mfe_fm_loss_grad(eng, loss, dzat, dsigat, depsat, dknot) &
            bind(c, name='mfe_fm_loss_grad')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: loss, dzat(*), dsigat(*), depsat(*), dknot(*)
      end function

      ! opt-in neutralising-background term of a net charge
      integer(c_int) function mfe_set_net_charge(eng, on) bind(c, name='mfe_set_net_charge')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: on
      end function

      ! DUTwoBodyEwaldRecSPM for a trial move of natm particles ianatm
      ! (1-based) to rtm(3,natm): du = change of u%rec
      integer(c_int) function mfe_mc_delta_rec(eng, natm, ianatm, rtm, reuse, du) &
            bind(c, name='mfe_mc_delta_rec')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: natm
         integer(c_int32_t), intent(in) :: ianatm(*)
         real(c_double), intent(in) :: rtm(*)
         integer(c_int32_t), value :: reuse
         real(c_double), intent(out) :: du
      end function

      ! EwaldSetup's iewaldopt parameter choice
      integer(c_int) function mfe_ewald_parameters(p) bind(c, name='mfe_ewald_parameters')
         import :: c_int, mfe_ewald_opt
         type(mfe_ewald_opt), intent(inout) :: p
      end function

      ! UTwoBodyANew - UTwoBodyAOld for the same trial move: du = change of
      ! the two-body energy, overlap = 1 for lhsoverlap
      integer(c_int) function mfe_mc_delta_pair(eng, natm, ianatm, rtm, du, overlap) &
            bind(c, name='mfe_mc_delta_pair')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: natm
         integer(c_int32_t), intent(in) :: ianatm(*)
         real(c_double), intent(in) :: rtm(*)
         real(c_double), intent(out) :: du
         integer(c_int32_t), intent(out) :: overlap
      end function

      ! accepted move: EwaldUpdateArray ('spm') + ro(:,ianatm) = rtm + NList
      integer(c_int) function mfe_mc_accept(eng, natm, ianatm, rtm) bind(c, name='mfe_mc_accept')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int32_t), value :: natm
         integer(c_int32_t), intent(in) :: ianatm(*)
         real(c_double), intent(in) :: rtm(*)
      end function

      ! batched replicas: engines(nrep) are the handles of set-up engines
      integer(c_int) function mfe_batch_create(engines, nrep, batch) bind(c, name='mfe_batch_create')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), intent(in) :: engines(*)
         integer(c_int32_t), value :: nrep
         type(c_ptr), intent(out) :: batch
      end function

      integer(c_int) function mfe_batch_destroy(batch) bind(c, name='mfe_batch_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function

      type(c_ptr) function mfe_batch_last_error(batch) bind(c, name='mfe_batch_last_error')
         import :: c_ptr
         type(c_ptr), value :: batch
      end function

      integer(c_int32_t) function mfe_batch_failed_replica(batch) &
            bind(c, name='mfe_batch_failed_replica')
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: batch
      end function

      integer(c_int) function mfe_batch_build_list(batch) bind(c, name='mfe_batch_build_list')
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function

      integer(c_int) function mfe_batch_compute(batch, flags) bind(c, name='mfe_batch_compute')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: flags
      end function

      integer(c_int) function mfe_batch_step_md(batch, nsteps, tstep, flags) &
            bind(c, name='mfe_batch_step_md')
         import :: c_int, c_int32_t, c_ptr, c_double
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: nsteps
         real(c_double), value :: tstep
         integer(c_int32_t), value :: flags
      end function

      integer(c_int) function mfe_batch_step_bd(batch, nsteps, flags) &
            bind(c, name='mfe_batch_step_bd')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: nsteps, flags
      end function

      integer(c_int) function mfe_batch_set_persistent(batch, on) &
            bind(c, name='mfe_batch_set_persistent')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: on
      end function

      integer(c_int32_t) function mfe_batch_uses_persistent(batch) &
            bind(c, name='mfe_batch_uses_persistent')
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: batch
      end function

      integer(c_int) function mfe_batch_get_energies(batch, energies) &
            bind(c, name='mfe_batch_get_energies')
         import :: c_int, c_ptr, mfe_energies
         type(c_ptr), value :: batch
         type(mfe_energies), intent(out) :: energies(*)
      end function

   end interface

contains

   ! MOLSIM's Stop for a failed engine call: prints the engine's message
   subroutine mfe_check(eng, rc, what)
      type(c_ptr), intent(in) :: eng
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(len=512), pointer :: msg
      integer :: n
      if (rc == MFE_OK) return
      if (c_associated(eng)) then
         call c_f_pointer(mfe_last_error(eng), msg)
         n = index(msg, c_null_char) - 1
         if (n < 0) n = len(msg)
         write(*,'(a,a,i3,a,a)') trim(what), ': error ', rc, ': ', msg(1:n)
      else
         write(*,'(a,a,i3)') trim(what), ': error ', rc
      end if
      error stop 1
   end subroutine mfe_check

end module mfe_mod
