! mcref_harness.f90 - drives the UNMODIFIED reference (MOCASSIN 2.02 modules,
! compiled from /root/reference/source by oracle/Makefile) to write golden
! vectors for the transport and balance path.  Runs only in the build
! container; its outputs are turned into small committed fixtures by
! tests/golden/make_fixtures.py.
!
! The setup replays mocassin.f90:66-104 (readInput .. dustDriver) and the
! pre-transport part of iterateMC (iteration_mod.f90:81-400: opacity and
! re-emission tables); transport calls the public energyPacketDriver one
! packet at a time with the random stream pinned per packet by
! mcref_set_stream (oracle/ref/mcref_rng.c); balance calls the public
! updateCell for every cell in the reference's own loop order
! (iteration_mod.f90:826-837).
!
! usage (cwd holds input/input.in):
!   mcref_harness tables    OUTDIR
!   mcref_harness transport OUTDIR SEED NPACKETS
!   mcref_harness walk      OUTDIR SEED NPACKETS NUBIN ORIGINFILE
!   mcref_harness balance   OUTDIR SEED NPACKETS
!   mcref_harness opacity   OUTDIR SEED NPACKETS
!   mcref_harness iterate   OUTDIR SEED          (under mpiexec: any rank count)
!   mcref_harness sweep     OUTDIR SEED NPACKETS (under mpiexec: any rank count)
!
! iterate runs mocassin.f90's MCIterationDriver to the deck's convergence
! criterion with the random stream of rank r pinned to (SEED, 0, r); the
! state files writeGrid leaves under output/ are the converged state.
! sweep times one energyPacketDriver call of NPACKETS/ranks packets per rank
! (the transport phase of iterateMC, iteration_mod.f90:474-496) on the
! tables of the initial state, with MPI_Wtime, and prints packets/s.
program mcref_harness
    use common_mod
    use constants_mod
    use vector_mod
    use interpolation_mod
    use set_input_mod
    use elements_mod
    use xSec_mod
    use continuum_mod
    use ionization_mod
    use grid_mod
    use dust_mod
    use emission_mod
    use photon_mod
    use update_mod
    use readdata_mod
    use iteration_mod
    use iso_c_binding
    implicit none
    include 'mpif.h'

    interface
       subroutine mcref_set_stream(seed, stream, packet) bind(C, name="mcref_set_stream")
         import :: c_int64_t, c_int32_t
         integer(c_int64_t), intent(in) :: seed
         integer(c_int32_t), intent(in) :: stream
         integer(c_int64_t), intent(in) :: packet
       end subroutine mcref_set_stream
    end interface

    type(grid_type) :: grid3D(maxGrids)
    character(len=512) :: mode, outdir, arg, originFile
    integer(c_int64_t) :: seed, p, npk
    integer(c_int32_t) :: stream
    integer :: iGrid, i, j, k, ic, nu, nOrig, io, nuBin, ios
    real :: orig(3), junk, w(3)
    integer :: fold
    type(vector) :: dirv

    call get_command_argument(1, mode)
    call get_command_argument(2, outdir)
    seed = 0; npk = 0
    if (command_argument_count() >= 3) then
       call get_command_argument(3, arg); read(arg, *) seed
    end if
    if (command_argument_count() >= 4) then
       call get_command_argument(4, arg); read(arg, *) npk
    end if

    call mpi_init(ierr)
    call mpi_comm_rank(MPI_COMM_WORLD, taskid, ierr)
    call mpi_comm_size(MPI_COMM_WORLD, numtasks, ierr)

    ! the warm start of mocassinWarm.f90:51: the state files a previous run
    ! wrote under output/, read back by resetGrid
    if (trim(mode) == "resetgrid") then
       call resetGrid(grid3D)
       call dump_reset()
       call mpi_finalize(ierr)
       stop
    end if

    ! --- mocassin.f90:66-104
    call readInput()
    do iGrid = 1, nGrids
       call initCartesianGrid(grid3D(iGrid), nxIn(iGrid), nyIn(iGrid), nzIn(iGrid))
    end do
    call initXSecArray()
    call setContinuum()
    call fillGrid(grid3D(1:nGrids))
    call setStarPosition(grid3D(1)%xAxis, grid3D(1)%yAxis, grid3D(1)%zAxis, grid3D(1:nGrids))
    if (lgGas) then
       call makeElements()
       call readData()
    end if
    if (lgDust) then
       do iGrid = 1, nGrids
          call dustDriver(grid3D(iGrid))
       end do
    end if

    if (trim(mode) == "iterate") then
       ! mocassin.f90:106-124 with the random stream pinned per rank
       stream = 0
       p = taskid
       call mcref_set_stream(seed, stream, p)
       call MCIterationDriver(grid3D(1:nGrids))
       call mpi_finalize(ierr)
       stop
    end if

    call build_tables()
    if (trim(mode) == "sweep") then
       call run_sweep()
       call mpi_finalize(ierr)
       stop
    end if
    if (lgGas) call dump_atomic()
    ! optional uniform opacity scaling (stress fixtures with many absorptions)
    call get_environment_variable("MCREF_OPACITY_SCALE", arg, status=ios)
    if (ios == 0) then
       read(arg, *) junk
       grid3D(1)%opacity = grid3D(1)%opacity * junk
    end if
    call dump_problem()

    select case (trim(mode))
    case ("tables")
    case ("transport")
       call run_transport()
    case ("balance")
       call run_transport()
       call run_balance()
    case ("opacity")
       call run_transport()
       call run_balance()
       call run_lines()
       call run_emission()
       if (lgGas .and. .not. lgDust) call run_cold_emission()
       call run_opacity()
       ! the state files of writeGrid (grid_mod.f90:2646-2872) under output/
       call writeGrid(grid3D(1:nGrids))
    case ("walk")
       call get_command_argument(5, arg); read(arg, *) nuBin
       call get_command_argument(6, originFile)
       call run_walk()
    case default
       print*, "mcref_harness: unknown mode ", trim(mode)
       stop 2
    end select

    call mpi_finalize(ierr)

contains

    ! iteration_mod.f90:81-400 for one grid and one rank
    subroutine build_tables()
       integer :: iG
       iG = 1
       grid3D(iG)%lgConverged(0:grid3D(iG)%nCells) = 0
       grid3D(iG)%lgBlack(0:grid3D(iG)%nCells) = 0
       if (lgGas) then
          grid3D(iG)%recPDF(0:grid3D(iG)%nCells, 1:nbins) = 0.
          grid3D(iG)%totalLines(0:grid3D(iG)%nCells) = 0.
          BjumpTemp = 0.
          Bjump = 0.
       end if
       grid3D(iG)%opacity(0:grid3D(iG)%nCells, 1:nbins) = 0.
       if (lgGas) then
          do i = 1, grid3D(iG)%nx
             do j = 1, grid3D(iG)%ny
                do k = 1, grid3D(iG)%nz
                   call ionizationDriver(grid3D(iG), i, j, k)
                end do
             end do
          end do
       end if
       if (lgDust) then
          grid3D(iG)%scaOpac(0:grid3D(iG)%nCells, 1:nbins) = 0.
          grid3D(iG)%absOpac(0:grid3D(iG)%nCells, 1:nbins) = 0.
          do i = 1, grid3D(iG)%nx
             do j = 1, grid3D(iG)%ny
                do k = 1, grid3D(iG)%nz
                   ic = grid3D(iG)%active(i,j,k)
                   if (ic > 0) then
                      call add_dust(iG, ic)
                   end if
                end do
             end do
          end do
       end if
       do i = 1, grid3D(iG)%nz
          do j = 1, grid3D(iG)%ny
             do k = 1, grid3D(iG)%nx
                call emissionDriver(grid3D, k, j, i, iG)
             end do
          end do
       end do
    end subroutine build_tables

    ! dust opacity add, iteration_mod.f90:184-222 (single dust chemistry)
    subroutine add_dust(iG, ic)
       integer, intent(in) :: iG, ic
       integer :: nS, ai, freq, dcp, nsp
       dcp = dustComPoint(1)
       nsp = 1
       do nS = 1, nSpeciesPart(nsp)
          do ai = 1, nSizes
             if (grid3D(iG)%Tdust(nS,ai,ic) < TdustSublime(dcp-1+nS)) then
                do freq = 1, nbins
                   grid3D(iG)%scaOpac(ic,freq) = grid3D(iG)%scaOpac(ic,freq) + &
                        & grainAbun(nsp,nS)*grainWeight(ai)*grid3D(iG)%Ndust(ic)* &
                        & xSecArray(dustScaXsecP(nS+dcp-1,ai)+freq-1)
                   grid3D(iG)%absOpac(ic,freq) = grid3D(iG)%absOpac(ic,freq) + &
                        & grainAbun(nsp,nS)*grainWeight(ai)*grid3D(iG)%Ndust(ic)* &
                        & xSecArray(dustAbsXsecP(nS+dcp-1,ai)+freq-1)
                end do
             end if
          end do
       end do
       do freq = 1, nbins
          grid3D(iG)%opacity(ic,freq) = grid3D(iG)%opacity(ic,freq) + &
               & (grid3D(iG)%scaOpac(ic,freq) + grid3D(iG)%absOpac(ic,freq))
       end do
    end subroutine add_dust

    ! Atomic inputs of updateCell (update_mod.f90:15-2027): photoionisation
    ! channels with their cross-section segments and frequency ranges as
    ! updateCell and thermBalance index them (update_mod.f90:184-215,
    ! 1150-1185), outer-shell weights (getOuterShell), dielectronic fits
    ! (direc_coeffs, aldropequi_coeffs) and the CEL model ions that
    ! forLines/equilibrium use (atomic_data_array).
    subroutine dump_atomic()
       integer :: elem, ion, nElec, outShell, g0, g1, IPnuP, hiR, hiH, j, nch, ncel
       integer :: ii, jj, nl, nt, xs
       integer, allocatable :: meta(:)
       real, allocatable :: sig(:)
       integer :: dri(25*3), api(167*2), hdr(8)
       real :: drr(25*5), apr(167*4)
       open(unit=81, file=trim(outdir)//'/ch_meta.bin', access='stream', form='unformatted', status='replace')
       open(unit=82, file=trim(outdir)//'/ch_sigma.bin', access='stream', form='unformatted', status='replace')
       allocate(sig(nbins))
       nch = 0
       do elem = 1, nElements
          do ion = 1, min(elem, nStages-1)
             if (.not.lgElementOn(elem)) exit
             g0 = 0; g1 = 0; outShell = 0
             if (elem > 2) then
                nElec = elem - ion + 1
                call getOuterShell(elem, nElec, outShell, g0, g1)
                IPnuP = elementP(elem, ion, outShell, 1)
                hiR = elementP(elem, ion, outShell, 2) - 1
                hiH = elementP(elem, ion, outShell, 2)
                xs = elementP(elem, ion, outShell, 3)
             else
                if (elem == 1) then
                   IPnuP = HlevNuP(1); xs = HlevXSecP(1)
                else if (ion == 1) then
                   IPnuP = HeIlevNuP(1); xs = HeISingXSecP(1)
                else
                   IPnuP = HeIIlevNuP(1); xs = HeIIXSecP(1)
                end if
                hiR = nbins; hiH = nbins
                nElec = elem - ion + 1
                call getOuterShell(elem, nElec, outShell, g0, g1)
             end if
             sig = 0.
             do j = IPnuP, min(hiH, nbins)
                sig(j) = xSecArray(j + xs - IPnuP)
             end do
             write(81) elem, ion, g0, g1, IPnuP, hiR, hiH, outShell
             write(82) sig
             nch = nch + 1
          end do
       end do
       close(81); close(82)
       do j = 1, 25
          dri(3*j-2) = direc_coeffs(j)%elem; dri(3*j-1) = direc_coeffs(j)%n; dri(3*j) = direc_coeffs(j)%g
          drr(5*j-4) = direc_coeffs(j)%a; drr(5*j-3) = direc_coeffs(j)%b; drr(5*j-2) = direc_coeffs(j)%c
          drr(5*j-1) = direc_coeffs(j)%d; drr(5*j) = direc_coeffs(j)%f
       end do
       do j = 1, 167
          api(2*j-1) = aldropequi_coeffs(j)%elem; api(2*j) = aldropequi_coeffs(j)%n
          apr(4*j-3) = aldropequi_coeffs(j)%a; apr(4*j-2) = aldropequi_coeffs(j)%b
          apr(4*j-1) = aldropequi_coeffs(j)%t0; apr(4*j) = aldropequi_coeffs(j)%t1
       end do
       call dump_i4('dr_int', dri, 75); call dump_r4('dr_real', drr, 125)
       call dump_i4('ap_int', api, 334); call dump_r4('ap_real', apr, 668)
       ! CEL model ions in forLines order (update_mod.f90:1270-1310)
       open(unit=83, file=trim(outdir)//'/cel_meta.bin', access='stream', form='unformatted', status='replace')
       open(unit=84, file=trim(outdir)//'/cel_dbl.bin', access='stream', form='unformatted', status='replace')
       open(unit=85, file=trim(outdir)//'/cel_flt.bin', access='stream', form='unformatted', status='replace')
       ncel = 0
       do elem = 3, nElements
          do ion = 1, min(elem+1, nstages)
             if (.not.lgElementOn(elem)) exit
             if (.not.lgDataAvailable(elem, ion)) cycle
             ! the entry equilibrium() finds for dataFile(elem, ion) (emission_mod.f90:1960-1968)
             do ii = 3, nElements
                do jj = 1, 10
                   if (trim(atomic_data_array(ii,jj)%ion) == trim(dataFile(elem, ion))) goto 77
                end do
             end do
77           continue
             nl = atomic_data_array(ii,jj)%nlevs
             nt = atomic_data_array(ii,jj)%ntemps
             write(83) elem, ion, ii, jj, nl, nt
             write(84) atomic_data_array(ii,jj)%e(1:nl), atomic_data_array(ii,jj)%a(1:nl,1:nl), &
                  & atomic_data_array(ii,jj)%qom(1:nt,1:nl,1:nl)
             write(85) atomic_data_array(ii,jj)%logtemp(1:nt), real(atomic_data_array(ii,jj)%g(1:nl)), &
                  & atomic_data_array(ii,jj)%br, atomic_data_array(ii,jj)%z, &
                  & atomic_data_array(ii,jj)%a_r(1:4), atomic_data_array(ii,jj)%a_d(1:5)
             ncel = ncel + 1
          end do
       end do
       close(83); close(84); close(85)
       hdr = 0
       hdr(1) = nstages; hdr(2) = nElementsUsed; hdr(3) = nAbComponents; hdr(4) = nch; hdr(5) = ncel
       call dump_i4('at_ints', hdr, 8)
       call dump_i4('at_elem_on', merge(1, 0, lgElementOn), nElements)
       call dump_i4('at_xref', elementXref, nElements)
       call dump_r4('at_abun', grid3D(1)%elemAbun(1, 1:nElements), nElements)
       call dump_i4('at_abfile', grid3D(1)%abFileIndex, grid3D(1)%nx*grid3D(1)%ny*grid3D(1)%nz)
       deallocate(sig)
    end subroutine dump_atomic

    subroutine dump_r4(name, a, n)
       character(*), intent(in) :: name
       integer, intent(in) :: n
       real, intent(in) :: a(n)
       open(unit=77, file=trim(outdir)//'/'//name//'.bin', access='stream', &
            & form='unformatted', status='replace')
       write(77) a
       close(77)
    end subroutine dump_r4

    subroutine dump_i4(name, a, n)
       character(*), intent(in) :: name
       integer, intent(in) :: n
       integer, intent(in) :: a(n)
       open(unit=77, file=trim(outdir)//'/'//name//'.bin', access='stream', &
            & form='unformatted', status='replace')
       write(77) a
       close(77)
    end subroutine dump_i4

    subroutine dump_problem()
       integer :: nc, nfl(8)
       real :: sc(16)
       nc = grid3D(1)%nCells
       call dump_r4('xAxis', grid3D(1)%xAxis, grid3D(1)%nx)
       call dump_r4('yAxis', grid3D(1)%yAxis, grid3D(1)%ny)
       call dump_r4('zAxis', grid3D(1)%zAxis, grid3D(1)%nz)
       call dump_i4('active', grid3D(1)%active, grid3D(1)%nx*grid3D(1)%ny*grid3D(1)%nz)
       call dump_r4('nuArray', nuArray, nbins)
       call dump_r4('widFlx', widFlx, nbins)
       call dump_r4('opacity', grid3D(1)%opacity, (nc+1)*nbins)
       if (lgGas) then
          call dump_r4('recPDF', grid3D(1)%recPDF, (nc+1)*nbins)
          call dump_r4('totalLines', grid3D(1)%totalLines, nc+1)
          call dump_r4('Te', grid3D(1)%Te, nc+1)
          call dump_r4('Ne', grid3D(1)%Ne, nc+1)
          call dump_r4('Hden', grid3D(1)%Hden, nc+1)
          call dump_r4('ionDen', grid3D(1)%ionDen, (nc+1)*nElementsUsed*nStages)
          call dump_i4('elementXref', elementXref, nElements)
       end if
       if (lgDust) then
          call dump_r4('scaOpac', grid3D(1)%scaOpac, (nc+1)*nbins)
          call dump_r4('absOpac', grid3D(1)%absOpac, (nc+1)*nbins)
          if (allocated(gSca)) call dump_r4('gSca', gSca, nbins)
       end if
       call dump_r4('specCDF', inSpectrumProbDen(1,1:nbins), nbins)
       sc = 0.
       sc(1) = R_out; sc(2) = grid3D(1)%geoCorrX; sc(3) = grid3D(1)%geoCorrY
       sc(4) = grid3D(1)%geoCorrZ; sc(5) = ionEdge(1); sc(6) = deltaE(1)
       sc(7) = starPosition(1)%x; sc(8) = starPosition(1)%y; sc(9) = starPosition(1)%z
       sc(10) = Lstar(1)
       call dump_r4('scalars', sc, 16)
       nfl = 0
       nfl(1) = grid3D(1)%nx; nfl(2) = grid3D(1)%ny; nfl(3) = grid3D(1)%nz
       nfl(4) = nc; nfl(5) = nbins
       nfl(6) = starIndeces(1,1)*10000*10000 + starIndeces(1,2)*10000 + starIndeces(1,3)
       nfl(7) = 0
       if (lgSymmetricXYZ) nfl(7) = nfl(7) + 1
       if (lgDust) nfl(7) = nfl(7) + 2
       if (lgGas) nfl(7) = nfl(7) + 4
       if (lgIsotropic) nfl(7) = nfl(7) + 8
       if (lgMultiStars) nfl(7) = nfl(7) + 16
       nfl(8) = 0
       if (lgGas) nfl(8) = nElementsUsed*100 + nStages
       call dump_i4('ints', nfl, 8)
    end subroutine dump_problem

    ! iteration_mod.f90:458-496 with one packet per energyPacketDriver call
    subroutine run_transport()
       real :: q(4)
       grid3D(1)%Jste(0:grid3D(1)%nCells, 1:nbins) = 0.
       grid3D(1)%escapedPackets(0:grid3D(1)%nCells, 0:nbins, 0:nAngleBins) = 0.
       absInt = 0.; scaInt = 0.
       stream = 1
       q = 0.
       do p = 0, npk - 1
          call mcref_set_stream(seed, stream, p)
          call energyPacketDriver(1, 1, grid3D(1:nGrids))
          q(1) = q(1) + Qphot
       end do
       q(2) = absInt; q(3) = scaInt
       call dump_r4('Jste_raw', grid3D(1)%Jste, (grid3D(1)%nCells+1)*nbins)
       call dump_r4('esc_raw', grid3D(1)%escapedPackets(:,:,0), (grid3D(1)%nCells+1)*(nbins+1))
       call dump_r4('transport_counters', q, 4)
       if (nAngleBins > 0) then
          ! every viewing-angle plane, and the tables initCartesianGrid set
          call dump_r4('esc_all', grid3D(1)%escapedPackets, (grid3D(1)%nCells+1)*(nbins+1)*(nAngleBins+1))
          call dump_i4('view_ints', (/ nAngleBins, totAngleBinsTheta, totAngleBinsPhi /), 3)
          call dump_r4('view_real', (/ dTheta, dPhi /), 2)
          call dump_i4('view_Ptheta', viewPointPtheta, totAngleBinsTheta+1)
          call dump_i4('view_Pphi', viewPointPphi, totAngleBinsPhi+1)
          call dump_r4('view_theta', viewPointTheta, nAngleBins+1)
          call dump_r4('view_phi', viewPointPhi, nAngleBins+1)
       end if
    end subroutine run_transport

    ! the transport phase of iterateMC (iteration_mod.f90:458-496) timed on
    ! every rank: NPACKETS/ranks packets each, rank r on stream (SEED, 2, r)
    subroutine run_sweep()
       double precision :: t0, t1, tmax
       integer :: load
       grid3D(1)%Jste(0:grid3D(1)%nCells, 1:nbins) = 0.
       grid3D(1)%escapedPackets(0:grid3D(1)%nCells, 0:nbins, 0:nAngleBins) = 0.
       load = int(npk / numtasks)
       stream = 2
       p = taskid
       call mcref_set_stream(seed, stream, p)
       call mpi_barrier(MPI_COMM_WORLD, ierr)
       t0 = MPI_Wtime()
       call energyPacketDriver(1, load, grid3D(1:nGrids))
       t1 = MPI_Wtime() - t0
       call mpi_allreduce(t1, tmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)
       if (taskid == 0) then
          open(unit=77, file=trim(outdir)//'/sweep.txt', status='replace')
          write(77, '(A,I6,A,I12,A,ES14.6,A,ES14.6)') 'ranks', numtasks, ' packets', load*numtasks, &
               & ' seconds', tmax, ' packets_per_s', dble(load)*numtasks/tmax
          close(77)
       end if
    end subroutine run_sweep

    ! iteration_mod.f90:700-722 scaling, then 744-909 (balance, one rank)
    subroutine run_balance()
       integer :: nc
       nc = grid3D(1)%nCells
       do i = 0, nc
          grid3D(1)%Jste(i,:) = grid3D(1)%Jste(i,:) * 1.e-9
          if (lgSymmetricXYZ) grid3D(1)%Jste(i,:) = grid3D(1)%Jste(i,:)/8.
       end do
       call dump_r4('Jste_scaled', grid3D(1)%Jste, (nc+1)*nbins)
       allocate(lgConvergedTemp(0:nc), lgBlackTemp(0:nc))
       allocate(NeTemp(0:nc), TeTemp(0:nc), ionDenTemp(0:nc, nElementsUsed, nStages))
       if (lgDust) then
          ! iteration_mod.f90:797-803
          allocate(TdustTemp(0:nSpeciesMax, 0:nSizes, 0:nc))
          TdustTemp = 0.
          call dump_dust()
       end if
       NeTemp = 0.; TeTemp = 0.; ionDenTemp = 0.
       grid3D(1)%lgConverged(0:nc) = 0
       grid3D(1)%lgBlack(0:nc) = 0
       grid3D(1)%noHit = 0.; grid3D(1)%noIonBal = 0.; grid3D(1)%noTeBal = 0.
       do i = 1, grid3D(1)%nx
          do j = 1, grid3D(1)%ny
             do k = 1, grid3D(1)%nz
                call updateCell(grid3D(1), i, j, k)
             end do
          end do
       end do
       call dump_r4('bal_Te', TeTemp, nc+1)
       call dump_r4('bal_Ne', NeTemp, nc+1)
       call dump_r4('bal_ionDen', ionDenTemp, (nc+1)*nElementsUsed*nStages)
       call dump_i4('bal_conv', grid3D(1)%lgConverged, nc+1)
       call dump_i4('bal_black', grid3D(1)%lgBlack, nc+1)
       if (lgDust) call dump_r4('bal_Tdust', TdustTemp, (nSpeciesMax+1)*(nSizes+1)*(nc+1))
    end subroutine run_balance

    ! The opacity tables after the balance pass (iteration_mod.f90:117-228
    ! on the updated Te, Ne, ionDen, Tdust): ionizationDriver per cell in
    ! iterateMC's loop order, then the dust add; and the cross-section
    ! segments addOpacity adds, in its order: (elem, ion, shell, nuLowP,
    ! nuHighP, xSecP) with the whole xSecArray and bremsXSecP.
    subroutine run_opacity()
       integer :: nc, elem, ion, sh, nseg
       integer, allocatable :: meta(:)
       nc = grid3D(1)%nCells
       call dump_r4('op_Te', grid3D(1)%Te, nc+1)
       call dump_r4('op_Ne', grid3D(1)%Ne, nc+1)
       call dump_r4('op_ionDen', grid3D(1)%ionDen, (nc+1)*nElementsUsed*nStages)
       if (lgDust) call dump_r4('op_Tdust', grid3D(1)%Tdust, (nSpeciesMax+1)*(nSizes+1)*(nc+1))
       grid3D(1)%opacity(0:nc, 1:nbins) = 0.
       do i = 1, grid3D(1)%nx
          do j = 1, grid3D(1)%ny
             do k = 1, grid3D(1)%nz
                call ionizationDriver(grid3D(1), i, j, k)
             end do
          end do
       end do
       call dump_r4('op_gas', grid3D(1)%opacity, (nc+1)*nbins)
       if (lgDust) then
          grid3D(1)%scaOpac(0:nc, 1:nbins) = 0.
          grid3D(1)%absOpac(0:nc, 1:nbins) = 0.
          do i = 1, grid3D(1)%nx
             do j = 1, grid3D(1)%ny
                do k = 1, grid3D(1)%nz
                   ic = grid3D(1)%active(i,j,k)
                   if (ic > 0) call add_dust(1, ic)
                end do
             end do
          end do
          call dump_r4('op_total', grid3D(1)%opacity, (nc+1)*nbins)
          call dump_r4('op_sca', grid3D(1)%scaOpac, (nc+1)*nbins)
          call dump_dust_sca()
       end if
       allocate(meta(6 * 4000))
       nseg = 0
       call seg(1, 1, 0, HlevNuP(1), nbins, HlevXSecP(1))
       call seg(2, 1, 0, HeIlevNuP(1), nbins, HeISingXSecP(1))
       call seg(2, 2, 0, HeIIlevNuP(1), nbins, HeIIXSecP(1))
       do elem = 3, nElements
          if (.not.lgElementOn(elem)) cycle
          do ion = 1, min(elem, nstages)
             do sh = 1, nShells(elem, ion)
                call seg(elem, ion, sh, elementP(elem, ion, sh, 1), elementP(elem, ion, sh, 2), &
                         elementP(elem, ion, sh, 3))
             end do
          end do
       end do
       call dump_i4('op_seg', meta, 6 * nseg)
       call dump_r4('op_xsec', xSecArray, size(xSecArray))
       meta(1) = bremsXSecP; meta(2) = nseg; meta(3) = size(xSecArray)
       call dump_i4('op_ints', meta, 3)
    contains
       subroutine seg(e, n, s, lo, hi, xp)
          integer, intent(in) :: e, n, s, lo, hi, xp
          meta(6*nseg+1) = e; meta(6*nseg+2) = n; meta(6*nseg+3) = s
          meta(6*nseg+4) = lo; meta(6*nseg+5) = hi; meta(6*nseg+6) = xp
          nseg = nseg + 1
       end subroutine seg
    end subroutine run_opacity

    ! outputGas's analytic line luminosities for a 3-D grid (output_mod.f90:
    ! 302-520) from the state after the balance pass: per active cell in its
    ! loop order, forLines (1946-2008: equilibrium of every model ion, times
    ! elemAbun * ionDen * 1.9865e9) and RecLinesEmission's Hbeta (2065-2079:
    ! 2530/Te**0.833 * Ne * ionDen(H,2) * elemAbun(H); HIRecLines(4,2) is
    ! (fh/hb) * Hbeta with fh = hb), each times Hden * getVolume and summed
    ! in the REAL forbVol and HbetaVol (before their division by HbetaVol).
    subroutine run_lines()
       integer :: nc, elem, ion, ncel, ic2, xr
       double precision :: fl(nForLevels, nForLevels)
       real, allocatable :: L(:, :, :)
       real :: up, x, hbc, hb(1), dV, Hd
       nc = grid3D(1)%nCells
       ncel = 0
       do elem = 3, nElements
          do ion = 1, min(elem+1, nstages)
             if (.not.lgElementOn(elem)) exit
             if (lgDataAvailable(elem, ion)) ncel = ncel + 1
          end do
       end do
       allocate(L(nForLevels, nForLevels, max(ncel, 1)))
       L = 0.
       hb = 0.
       do i = 1, grid3D(1)%nx
          do j = 1, grid3D(1)%ny
             do k = 1, grid3D(1)%nz
                ic2 = grid3D(1)%active(i,j,k)
                if (ic2 <= 0) cycle
                dV = getVolume(grid3D(1), i, j, k)
                Hd = grid3D(1)%Hden(ic2)
                ncel = 0
                do elem = 3, nElements
                   do ion = 1, min(elem+1, nstages)
                      if (.not.lgElementOn(elem)) exit
                      if (.not.lgDataAvailable(elem, ion)) cycle
                      ncel = ncel + 1
                      xr = elementXref(elem)
                      x = grid3D(1)%ionDen(ic2, xr, ion)
                      if (ion < nstages) then
                         up = grid3D(1)%ionDen(ic2, xr, ion+1) / x
                      else
                         up = 0.
                      end if
                      fl = 0.d0
                      call equilibrium(dataFile(elem, ion), up, grid3D(1)%Te(ic2), grid3D(1)%Ne(ic2), fl)
                      fl = fl * grid3D(1)%elemAbun(1, elem) * x
                      fl = fl * 1.9865e9
                      L(:, :, ncel) = L(:, :, ncel) + real(fl * Hd * dV)
                   end do
                end do
                hbc = 2530./(grid3D(1)%Te(ic2)**0.833)
                hbc = hbc * grid3D(1)%Ne(ic2) * grid3D(1)%ionDen(ic2, elementXref(1), 2) * grid3D(1)%elemAbun(1, 1)
                hb(1) = hb(1) + real(dble(hbc) * Hd * dV)
             end do
          end do
       end do
       open(unit=86, file=trim(outdir)//'/lines_L.bin', access='stream', form='unformatted', status='replace')
       write(86) L
       close(86)
       open(unit=86, file=trim(outdir)//'/lines_hb.bin', access='stream', form='unformatted', status='replace')
       write(86) hb
       close(86)
       deallocate(L)
    end subroutine run_lines

    ! The re-emission tables of the state after the balance pass: the module
    ! tables emissionDriver reads (ph_mod/hydro_mod/readData setup), then
    ! emissionDriver for every cell in iterateMC's order (iteration_mod.f90:
    ! 229-240) and the resulting recPDF and totalLines.
    subroutine run_emission()
       integer :: nc, elem, ion, nElec, outShell, g0, g1, hiP, ipP, xsP, nfb, i0HeI, i0HeII, ii, n
       integer :: ints(16), lyP(NHeIILyman)
       integer, allocatable :: fbm(:, :)
       real, allocatable :: fbg(:), fbs(:), sHI(:), sHeI(:), sHeII(:)
       nc = grid3D(1)%nCells
       i0HeI = max(HlevNuP(1), HeINuEdgeP(nlimGammaHeI))
       i0HeII = max(i0HeI, HeIINuEdgeP(nlimGammaHeII))
       ! heavy-ion free-bound channels in fb_ff's loop order (emission_mod.f90:473-521)
       allocate(fbm(4, 200), fbg(200), fbs(200*nbins))
       nfb = 0
       n = 0
       do elem = 3, 19
          do ion = 1, nstages-1
             if (.not.lgElementOn(elem)) exit
             if (ion > elem) exit
             nElec = elem - ion + 1
             call getOuterShell(elem, nElec, outShell, g0, g1)
             hiP = elementP(elem, ion, outShell, 2)
             ipP = elementP(elem, ion, outShell, 1)
             xsP = elementP(elem, ion, outShell, 3)
             nfb = nfb + 1
             fbm(:, nfb) = (/elem, ion, ipP, hiP/)
             fbg(nfb) = real(g0)/real(g1)
             do ii = ipP, hiP-1
                n = n + 1
                fbs(n) = xSecArray(ii+xsP-ipP+1-1)
             end do
          end do
       end do
       allocate(sHI(nbins-HlevNuP(1)+1), sHeI(nbins-i0HeI+1), sHeII(nbins-i0HeII+1))
       do ii = HlevNuP(1), nbins
          sHI(ii-HlevNuP(1)+1) = xSecArray(ii-HlevNuP(1)+1+HlevXSecP(1)-1)
       end do
       do ii = i0HeI, nbins
          sHeI(ii-i0HeI+1) = xSecArray(ii-HeIlevNuP(1)+1+HeISingXSecP(1)-1)
       end do
       do ii = i0HeII, nbins
          sHeII(ii-i0HeII+1) = xSecArray(ii-HeIIlevNuP(1)+1+HeIIXSecP(1)-1)
       end do
       ints = 0
       ints(1) = ntkGamma; ints(2) = nlimGammaHI; ints(3) = nlimGammaHeI; ints(4) = nlimGammaHeII
       ints(5) = HlevNuP(1); ints(6) = HeIlevNuP(1); ints(7) = HeIIlevNuP(1)
       ints(8) = i0HeI; ints(9) = i0HeII; ints(10) = nfb; ints(11) = NHeIILyman
       call locate(nuArray, 1.45673, ints(12))
       do ii = 1, NHeIILyman
          call locate(nuArray, HeIILymanNu(ii), lyP(ii))
       end do
       call dump_i4('em_ints', ints, 16)
       call dump_i4('em_lyP', lyP, NHeIILyman)
       call dump_r4('em_tk', tkGamma, ntkGamma)
       call dump_r4('em_gHI', logGammaHI, ntkGamma*nlimGammaHI)
       call dump_r4('em_gHeI', logGammaHeI, ntkGamma*nlimGammaHeI)
       call dump_r4('em_gHeII', logGammaHeII, ntkGamma*nlimGammaHeII)
       call dump_i4('em_eHI', HINuEdgeP, nlimGammaHI)
       call dump_i4('em_eHeI', HeINuEdgeP, nlimGammaHeI)
       call dump_i4('em_eHeII', HeIINuEdgeP, nlimGammaHeII)
       call dump_r4('em_gff', gauntFF, nbins)
       call dump_r4('em_gffHeII', gauntFFHeII, nbins)
       call dump_r4('em_sHI', sHI, size(sHI))
       call dump_r4('em_sHeI', sHeI, size(sHeI))
       call dump_r4('em_sHeII', sHeII, size(sHeII))
       call dump_i4('em_fbm', fbm, 4*nfb)
       call dump_r4('em_fbg', fbg, nfb)
       call dump_r4('em_fbs', fbs, max(n, 1))
       call dump_r4('em_ydat', y_dat, 41)
       call dump_r4('em_Aydat', Ay_dat, 41)
       call dump_r4('em_lyData', HeIILymanData, NHeIILyman)
       call dump_r4('em_lyNu', HeIILymanNu, NHeIILyman)
       call dump_r4('em_HeIc', HeIrecLineCoeff, 34*3*4)
       open(unit=86, file=trim(outdir)//'/em_HIrec.bin', access='stream', form='unformatted', status='replace')
       write(86) HIRecLineData
       close(86)
       open(unit=86, file=trim(outdir)//'/em_HeIIrec.bin', access='stream', form='unformatted', status='replace')
       write(86) HeIIRecLineData
       close(86)
       grid3D(1)%recPDF(0:nc, 1:nbins) = 0.
       grid3D(1)%totalLines(0:nc) = 0.
       do i = 1, grid3D(1)%nz
          do j = 1, grid3D(1)%ny
             do k = 1, grid3D(1)%nx
                call emissionDriver(grid3D, k, j, i, 1)
             end do
          end do
       end do
       call dump_r4('em_recPDF', grid3D(1)%recPDF, (nc+1)*nbins)
       call dump_r4('em_totalLines', grid3D(1)%totalLines, nc+1)
       deallocate(fbm, fbg, fbs, sHI, sHeI, sHeII)
    end subroutine run_emission

    ! emissionDriver of four cells put at temperatures the reference's
    ! iterateT reaches in barely lit cells (its 1 K floor, and 5, 19.554 and
    ! 60 K; each cell keeps its balanced Ne and ionDen): the raw recPDF rows
    ! and totalLines, NaN entries included; the balanced state is restored
    subroutine run_cold_emission()
       integer, parameter :: ncold = 4
       real, parameter :: tcold(ncold) = (/ 1., 5., 19.554296, 60. /)
       integer :: cells(ncold), ic, ix, iy, iz
       real :: tsave(ncold), rows(ncold, nbins), tls(ncold)
       do ic = 1, ncold
          cells(ic) = grid3D(1)%nCells - 7 * (ic - 1)
       end do
       do ic = 1, ncold
          tsave(ic) = grid3D(1)%Te(cells(ic))
          grid3D(1)%Te(cells(ic)) = tcold(ic)
          do ix = 1, grid3D(1)%nx
             do iy = 1, grid3D(1)%ny
                do iz = 1, grid3D(1)%nz
                   if (grid3D(1)%active(ix,iy,iz) == cells(ic)) call emissionDriver(grid3D, ix, iy, iz, 1)
                end do
             end do
          end do
          rows(ic, :) = grid3D(1)%recPDF(cells(ic), 1:nbins)
          tls(ic) = grid3D(1)%totalLines(cells(ic))
       end do
       do ic = 1, ncold
          grid3D(1)%Te(cells(ic)) = tsave(ic)
       end do
       call dump_i4('cold_cells', cells, ncold)
       call dump_r4('cold_Te', tcold, ncold)
       call dump_r4('cold_recPDF', rows, ncold*nbins)
       call dump_r4('cold_totalLines', tls, ncold)
    end subroutine run_cold_emission

    ! the state resetGrid read back (grid_mod.f90:2967-3568)
    subroutine dump_reset()
       integer :: nc, ints(8)
       nc = grid3D(1)%nCells
       ints = 0
       ints(1) = grid3D(1)%nx; ints(2) = grid3D(1)%ny; ints(3) = grid3D(1)%nz; ints(4) = nc
       ints(5) = nElementsUsed; ints(6) = nstages
       if (lgDust) then
          ints(7) = nSpeciesMax; ints(8) = nSizes
       end if
       call dump_i4('rg_ints', ints, 8)
       call dump_r4('rg_xAxis', grid3D(1)%xAxis, grid3D(1)%nx)
       call dump_r4('rg_yAxis', grid3D(1)%yAxis, grid3D(1)%ny)
       call dump_r4('rg_zAxis', grid3D(1)%zAxis, grid3D(1)%nz)
       call dump_i4('rg_active', grid3D(1)%active, grid3D(1)%nx*grid3D(1)%ny*grid3D(1)%nz)
       call dump_i4('rg_conv', grid3D(1)%lgConverged, nc+1)
       call dump_i4('rg_black', grid3D(1)%lgBlack, nc+1)
       if (lgGas) then
          call dump_r4('rg_Te', grid3D(1)%Te, nc+1)
          call dump_r4('rg_Ne', grid3D(1)%Ne, nc+1)
          call dump_r4('rg_Hden', grid3D(1)%Hden, nc+1)
          call dump_r4('rg_ionDen', grid3D(1)%ionDen, (nc+1)*nElementsUsed*nstages)
       end if
       if (lgDust) then
          call dump_r4('rg_Ndust', grid3D(1)%Ndust, nc+1)
          call dump_r4('rg_Tdust', grid3D(1)%Tdust, (nc+1)*(nSpeciesMax+1)*(nSizes+1))
       end if
    end subroutine dump_reset

    ! scattering cross-sections of every species and size on nuArray and the
    ! sublimation temperatures (the dust add of iteration_mod.f90:184-222)
    subroutine dump_dust_sca()
       integer :: nS, ai, nsp
       real, allocatable :: xs(:)
       nsp = 1
       allocate(xs(nbins * nSizes * nSpeciesPart(nsp)))
       do nS = 1, nSpeciesPart(nsp)
          do ai = 1, nSizes
             do i = 1, nbins
                xs(i + nbins * ((ai - 1) + nSizes * (nS - 1))) = xSecArray(dustScaXsecP(nS,ai) + i - 1)
             end do
          end do
       end do
       call dump_r4('dust_sca', xs, size(xs))
       call dump_r4('dust_tsub', TdustSublime, nSpecies)
    end subroutine dump_dust_sca

    ! Inputs of getDustT (update_mod.f90:1836-1945) for one dust chemistry:
    ! absorption cross sections of every species and size on nuArray, size
    ! weights, species abundances, the emission integrals dustEmIntegral
    ! (dust_mod.f90:141-183) and the grid's initial Tdust.
    subroutine dump_dust()
       integer :: hdr(5), nS, ai, nsp, nc, elem, ist
       integer :: ipc(nElements, nstages)
       real, allocatable :: xs(:)
       nsp = 1
       nc = grid3D(1)%nCells
       hdr(1) = nSpeciesPart(nsp); hdr(2) = nSizes; hdr(3) = nTemps; hdr(4) = nSpeciesMax
       hdr(5) = nSpecies
       call dump_i4('dust_ints', hdr, 5)
       allocate(xs(nbins * nSizes * nSpeciesPart(nsp)))
       do nS = 1, nSpeciesPart(nsp)
          do ai = 1, nSizes
             do i = 1, nbins
                xs(i + nbins * ((ai - 1) + nSizes * (nS - 1))) = xSecArray(dustAbsXsecP(nS,ai) + i - 1)
             end do
          end do
       end do
       call dump_r4('dust_abs', xs, size(xs))
       call dump_r4('dust_weight', grainWeight, nSizes)
       call dump_r4('dust_abun', grainAbun(nsp, 1:nSpeciesPart(nsp)), nSpeciesPart(nsp))
       call dump_r4('dust_emint', dustEmIntegral, nSpecies * nSizes * nTemps)
       call dump_r4('Tdust0', grid3D(1)%Tdust, (nSpeciesMax+1)*(nSizes+1)*(nc+1))
       ! grain charge, photoelectric and gas-grain collision inputs
       ! (update_mod.f90:527-932): flag, grainVn, MsurfAtom, grainRadius,
       ! Ndust, aWeight and the nuArray index of the ionisation potential
       ! setDustGasCollHeatCool uses for (elem, istage >= 2)
       hdr = 0
       hdr(1) = merge(1, 0, lgPhotoelectric)
       call dump_i4('dust_pe', hdr, 1)
       call dump_r4('dust_vn', grainVn, nSpecies)
       call dump_r4('dust_msurf', MsurfAtom, nSpecies)
       call dump_r4('dust_radius', grainRadius, nSizes)
       call dump_r4('dust_ndust', grid3D(1)%Ndust, nc+1)
       call dump_r4('dust_aweight', aWeight, nElements)
       ipc = 0
       do elem = 1, nElements
          if (.not.lgElementOn(elem)) cycle
          do ist = 2, min(elem+1, nstages)
             if (elem == 1) then
                ipc(elem, ist) = HlevNuP(1)
             else if (elem == 2 .and. ist == 2) then
                ipc(elem, ist) = HeIlevNuP(1)
             else if (elem == 2 .and. ist == 3) then
                ipc(elem, ist) = HeIIlevNuP(1)
             else
                ipc(elem, ist) = elementP(elem, ist-1, nShells(elem, ist-1), 1)
             end if
          end do
       end do
       call dump_i4('dust_ipcoll', ipc, nElements*nstages)
    end subroutine dump_dust

    ! Geometry fixtures: zero opacity, source bin fixed, deltaE = 1.
    ! For each origin (x y z fold) and packet: record the direction the
    ! reference's own randomUnitVector returns on that packet's stream
    ! (draw 0 is getNu2's, photon_mod.f90:727), then run the packet and
    ! dump every non-zero Jste entry (= sum of dS*1/dV over its visits).
    subroutine run_walk()
       integer :: nc, xs, ys, zs, irep, cnt
       real :: u1
       nc = grid3D(1)%nCells
       grid3D(1)%opacity = 0.
       inSpectrumProbDen(1,:) = 0.
       inSpectrumProbDen(1,nuBin:nbins) = 1.
       deltaE(1) = 1.
       stream = 7
       open(unit=31, file=trim(originFile), status='old')
       open(unit=32, file=trim(outdir)//'/walk_dirs.txt', status='replace')
       open(unit=33, file=trim(outdir)//'/walk_J.txt', status='replace')
       nOrig = 0
       do
          read(31, *, iostat=ios) orig(1), orig(2), orig(3), fold
          if (ios /= 0) exit
          nOrig = nOrig + 1
          starPosition(1)%x = orig(1); starPosition(1)%y = orig(2); starPosition(1)%z = orig(3)
          ! the cell lookup of setStarPosition (grid_mod.f90:3590-3603)
          call locate(grid3D(1)%xAxis, orig(1), xs)
          if (xs < grid3D(1)%nx) then
             if (orig(1) > (grid3D(1)%xAxis(xs)+grid3D(1)%xAxis(xs+1))/2.) xs = xs+1
          end if
          call locate(grid3D(1)%yAxis, orig(2), ys)
          if (ys < grid3D(1)%ny) then
             if (orig(2) > (grid3D(1)%yAxis(ys)+grid3D(1)%yAxis(ys+1))/2.) ys = ys+1
          end if
          call locate(grid3D(1)%zAxis, orig(3), zs)
          if (zs < grid3D(1)%nz) then
             if (orig(3) > (grid3D(1)%zAxis(zs)+grid3D(1)%zAxis(zs+1))/2.) zs = zs+1
          end if
          starIndeces(1,1) = xs; starIndeces(1,2) = ys; starIndeces(1,3) = zs; starIndeces(1,4) = 1
          lgMultiStars = (fold == 0)
          do p = 0, npk - 1
             ! direction as initPhotonPacket builds it (photon_mod.f90:650-675)
             call mcref_set_stream(seed, stream, int(nOrig, c_int64_t)*100000 + p)
             call random_number(u1)
             if (u1 == 0. .or. u1 == 1. .or. u1 == 0.9999999) cycle
             do irep = 1, 1000000
                dirv = randomUnitVector()
                if (dirv%x /= 0. .and. dirv%y /= 0. .and. dirv%z /= 0.) exit
             end do
             if (lgSymmetricXYZ .and. .not.lgMultiStars) then
                if (dirv%x < 0.) dirv%x = -dirv%x
                if (dirv%y < 0.) dirv%y = -dirv%y
                if (dirv%z < 0.) dirv%z = -dirv%z
             end if
             grid3D(1)%Jste = 0.
             grid3D(1)%escapedPackets = 0.
             call mcref_set_stream(seed, stream, int(nOrig, c_int64_t)*100000 + p)
             call energyPacketDriver(1, 1, grid3D(1:nGrids))
             cnt = 0
             do ic = 0, nc
                if (grid3D(1)%Jste(ic, nuBin) /= 0.) cnt = cnt + 1
             end do
             write(32, '(I8,1X,I8,3(1X,Z8.8),3(1X,Z8.8),3(1X,I5),1X,I6,1X,Z8.8)') nOrig, p, &
                  & transfer(orig(1), 1), transfer(orig(2), 1), transfer(orig(3), 1), &
                  & transfer(dirv%x, 1), transfer(dirv%y, 1), transfer(dirv%z, 1), &
                  & xs, ys, zs, cnt, transfer(sum(grid3D(1)%escapedPackets(:,nuBin,0)), 1)
             do ic = 0, nc
                if (grid3D(1)%Jste(ic, nuBin) /= 0.) then
                   write(33, '(I8,1X,I8,1X,I8,1X,Z8.8)') nOrig, p, ic, &
                        & transfer(grid3D(1)%Jste(ic, nuBin), 1)
                end if
             end do
          end do
       end do
       close(31); close(32); close(33)
    end subroutine run_walk

end program mcref_harness
