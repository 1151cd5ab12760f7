! ref_probe.F90 -- test-infrastructure driver for the amdflang build of the
! MOLSIM reference (oracle/_ref).  It is linked against the reference's own
! objects, built from the sources where they lie by oracle/ref.mk, and is
! never part of the product.
!
! It runs the reference's setup stages exactly as MolsimDriver does
! (Src/molsim.F90:55-79) for the input <project>.in, then writes, in raw
! little-endian stream files:
!   <project>.probe_tab   : nbuf_used, nptpt, then ubuf(1:nbuf_used),
!                           iubuflow(1:nptpt), nugrid(1:nptpt),
!                           r2umin(1:nptpt), rcut, ualpha, EpsiFourPi
!   <project>.probe_f     : (only if <project>.probe_pos exists: np*3 reals
!                           giving ro, already inside the box) the state
!                           after SetVList (or SetCellList) + UTotal there:
!                           np, maxnneigh, force(3,np), forceo(3,np),
!                           u%tot, u%twob(0:nptpt), u%rec, u%bond, virial,
!                           nneighpn(1:np), jpnlist(1:maxnneigh,1:np)
!   <project>.probe_bdout : (only if <project>.probe_pos and .probe_bd exist,
!                           the latter one integer nbd, and the input is a
!                           BD run) np, nbd, iseed, then for each of nbd
!                           BDSteps forceo, ro before and drostep, ro after
!   <project>.probe_math  : (only if <project>.probe_args exists: n reals)
!                           Erf2(x) for each x, then for orders 3..8
!                           CardinalBSpline(order,x) values and derivs.
!   <project>.probe_pot   : (only if <project>.probe_r exists: n reals) for
!                           each r: CalcEwaldUfac(r, ualpha) c0,c1,c2 and
!                           Pot_1_6_12('calc') u0,u1,u2 for every atom-type
!                           pair iat <= jat.

program RefProbe

   use MolModule
   use NListModule, only: maxnneigh
   use PotentialModule, only: nugrid
   use CardinalBSplineModule
   use CellListModule, only: SetCellList
   implicit none

   integer(4) :: ios, ufile, nused, narg, i, iord, kmax, iat, jat, idum
   real(8) :: c0, c1, c2, u0, u1, u2
   external Pot_1_6_12
   real(8), allocatable :: args(:), vals(:), ders(:)
   real(8) :: Erf2
   real(8) :: GauRandom, Random
   integer(4) :: nrng, iseedg, idumr, nbd, ibd
   real(8), allocatable :: rv(:)
   logical :: lexist
   character(200) :: proj

   nproc  = 1
   myid   = 0
   master = .true.
   slave  = .false.

   call IOMolsim(iReadInput)
   call IOSystem(iReadInput)
   call IOScale(iReadInput)
   call Particle(iReadInput)
   call PotentialDriver(iReadInput)
   call Coordinate(iReadInput)
   call ProbeSub(iReadInput)

   call IOMolsim(iWriteInput)
   call IOSystem(iWriteInput)
   call IOScale(iWriteInput)
   call Particle(iWriteInput)
   call PotentialDriver(iWriteInput)
   call IOCnf('open')
   call Coordinate(iWriteInput)
   call ProbeSub(iWriteInput)

   call get_command_argument(1, proj)

! ... potential table

   nused = iubuflow(nptpt) + 12*nugrid(nptpt) - 1
   open(newunit=ufile, file=trim(proj)//'.probe_tab', access='stream', form='unformatted', status='replace')
   write(ufile) nused, nptpt
   write(ufile) ubuf(1:nused)
   write(ufile) iubuflow(1:nptpt), nugrid(1:nptpt)
   write(ufile) r2umin(1:nptpt)
   write(ufile) rcut, ualpha, EpsiFourPi
   close(ufile)

! ... Ewald parameters as EwaldSetup chose them (iewaldopt)

   open(newunit=ufile, file=trim(proj)//'.probe_ewald', access='stream', form='unformatted', status='replace')
   write(ufile) iewaldopt, lq2sum, ncut
   write(ufile) uewaldtol, ualphared, ualpha, rcut, q2sum, boxlenshort, EpsiFourPi
   close(ufile)

! ... forces at given positions

   inquire(file=trim(proj)//'.probe_pos', exist=lexist)
   if (lexist) then
      open(newunit=ufile, file=trim(proj)//'.probe_pos', access='stream', form='unformatted', status='old')
      read(ufile) ro(1:3,1:np)
      close(ufile)
      call SetAtomProp(1, np, .false.)
      if (lvlist) call SetVList
      if (lCellList) call SetCellList
      call UTotal(iSimulationStep)
      open(newunit=ufile, file=trim(proj)//'.probe_f', access='stream', form='unformatted', status='replace')
      write(ufile) np, maxnneigh
      write(ufile) force(1:3,1:np), forceo(1:3,1:np)
      write(ufile) u%tot, u%twob(0:nptpt), u%rec, u%bond, virial
      write(ufile) nneighpn(1:np)
      write(ufile) jpnlist(1:maxnneigh,1:np)
      close(ufile)

! ... Brownian dynamics steps from there (only if <project>.probe_bd exists:
!     one integer, the number of steps): iseed before the first step, then
!     per step forceo(3,np) and ro(3,np) before BDStep, drostep(3,np) and
!     ro(3,np) after it, as BDDriver runs BDStep and UTotal
!     (Src/bd.F90:54-58); the list is rebuilt before each UTotal

      inquire(file=trim(proj)//'.probe_bd', exist=lexist)
      if (lexist .and. lbd) then
         open(newunit=ufile, file=trim(proj)//'.probe_bd', access='stream', form='unformatted', status='old')
         read(ufile) nbd
         close(ufile)
         open(newunit=ufile, file=trim(proj)//'.probe_bdout', access='stream', form='unformatted', status='replace')
         write(ufile) np, nbd, iseed
         do ibd = 1, nbd
            write(ufile) forceo(1:3,1:np), ro(1:3,1:np)
            call BDStep
            write(ufile) drostep(1:3,1:np), ro(1:3,1:np)
            if (lvlist) call SetVList
            call UTotal(iSimulationStep)
         end do
         close(ufile)
      end if
   end if

! ... math routines

   inquire(file=trim(proj)//'.probe_args', exist=lexist, size=narg)
   if (lexist) then
      narg = narg/8
      allocate(args(narg))
      open(newunit=ufile, file=trim(proj)//'.probe_args', access='stream', form='unformatted', status='old')
      read(ufile, iostat=ios) args
      close(ufile)
      open(newunit=ufile, file=trim(proj)//'.probe_math', access='stream', form='unformatted', status='replace')
      write(ufile) (Erf2(args(i)), i = 1, narg)
      do iord = 3, 8
         allocate(vals(0:iord-1), ders(0:iord-1))
         do i = 1, narg
            call CardinalBSpline(iord, args(i), kmax, vals, ders)
            write(ufile) kmax, vals, ders
         end do
         deallocate(vals, ders)
      end do
      close(ufile)
   end if

! ... pair potential

   inquire(file=trim(proj)//'.probe_r', exist=lexist, size=narg)
   if (lexist) then
      narg = narg/8
      if (allocated(args)) deallocate(args)
      allocate(args(narg))
      open(newunit=ufile, file=trim(proj)//'.probe_r', access='stream', form='unformatted', status='old')
      read(ufile, iostat=ios) args
      close(ufile)
      open(newunit=ufile, file=trim(proj)//'.probe_pot', access='stream', form='unformatted', status='replace')
      do i = 1, narg
         call CalcEwaldUfac(args(i), ualpha, c0, c1, c2)
         write(ufile) c0, c1, c2
         do iat = 1, nat
            do jat = iat, nat
               call Pot_1_6_12('calc', idum, idum, iat, jat, args(i), u0, u1, u2)
               write(ufile) u0, u1, u2
            end do
         end do
      end do
      close(ufile)
   end if

! ... random number streams: GauRandom and Random (Src/mollib.F90:2689-2766)
!     from given seeds, 2 nrng and nrng values

   inquire(file=trim(proj)//'.probe_rng', exist=lexist)
   if (lexist) then
      open(newunit=ufile, file=trim(proj)//'.probe_rng', access='stream', form='unformatted', status='old')
      read(ufile) nrng, iseedg, idumr
      close(ufile)
      allocate(rv(2*nrng))
      do i = 1, 2*nrng
         rv(i) = GauRandom(iseedg)
      end do
      open(newunit=ufile, file=trim(proj)//'.probe_rngout', access='stream', form='unformatted', status='replace')
      write(ufile) rv(1:2*nrng)
      do i = 1, nrng
         rv(i) = Random(idumr)
      end do
      write(ufile) rv(1:nrng)
      close(ufile)
   end if

contains

subroutine ProbeSub(iStage)
   integer(4), intent(in) :: iStage
   if (lmd) call MDDriver(iStage)
   if (lbd) call BDDriver(iStage)
   call NListDriver(iStage)
end subroutine ProbeSub

end program RefProbe
