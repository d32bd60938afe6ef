! mc_fortran_check.f90 - a Fortran host exercising the C ABI the way the
! reference's iterateMC would: create a context from the grid, upload the
! sweep's tables, run energyPacketDriver for every star, scale J
! (x1e-9, /8 when symmetric; iteration_mod.f90:700-722) and run the
! updateCell loop, all through mocassin_amd_mod.
!
! usage:
!   mc_fortran_check selftest         argument checks only (no GPU needed)
!   mc_fortran_check run DIR          DIR/problem.bin -> DIR/J.bin, DIR/esc.bin,
!                                     DIR/balance.bin
!   mc_fortran_check run DIR short    the same, then a sweep into a J one row
!                                     short, which the module must stop
!   mc_fortran_check grains DIR       no sweep: DIR/grains.bin gives J and the
!                                     grain tables; the updateCell loop with
!                                     grains -> DIR/grains_out.bin
! problem.bin (stream, native endian): int64 header(24) = nx ny nz nCells
! nbins flags nStars npackets seed have_balance nstages nElemUsed nch ncel
! n_cel_dbl n_cel_flt device, then the grid, the tables and the balance
! inputs in the order read below.
program mc_fortran_check
    use, intrinsic :: iso_c_binding
    use mocassin_amd_mod
    implicit none
    character(len=512) :: mode, dir
    integer :: nargs
    ! grains mode: the scaled J and the grain tables of grains.bin
    integer(c_int64_t) :: gh(8)
    real(c_float), allocatable, target :: gAbs(:), gWt(:), gAbn(:), gEm(:), gRad(:), gVn(:), gMs(:), gAw(:), gNd(:)
    real, allocatable :: gTd(:, :, :)
    real(c_float) :: gXhil
    real(c_double), allocatable, target :: gJ(:, :)

    nargs = command_argument_count()
    if (nargs < 1) then
        print *, 'usage: mc_fortran_check selftest | run DIR'
        stop 2
    end if
    call get_command_argument(1, mode)
    if (trim(mode) == 'selftest') then
        call selftest()
    else if (trim(mode) == 'run' .and. nargs >= 2) then
        call get_command_argument(2, dir)
        mode = ' '
        if (nargs >= 3) call get_command_argument(3, mode)
        call run(trim(dir), trim(mode) == 'short', .false.)
    else if (trim(mode) == 'grains' .and. nargs >= 2) then
        call get_command_argument(2, dir)
        call run(trim(dir), .false., .true.)
    else
        print *, 'usage: mc_fortran_check selftest | run DIR [short] | grains DIR'
        stop 2
    end if

contains

    subroutine selftest()
        type(mc_desc) :: d
        type(c_ptr) :: ctx
        integer(c_int) :: rc
        character(len=512) :: msg
        d%abi_version = 0
        rc = mc_ctx_create(d, ctx)
        msg = mc_error_text(c_null_ptr)
        if (rc /= MC_EINVAL .or. len_trim(msg) == 0 .or. c_associated(ctx)) then
            print *, 'selftest FAILED: rc=', rc, ' msg=', trim(msg)
            stop 1
        end if
        print *, 'selftest ok: ', trim(msg)
    end subroutine selftest

    subroutine run(dir, short, grains)
        character(len=*), intent(in) :: dir
        logical, intent(in) :: short, grains
        integer(c_int64_t) :: h(24)
        integer :: u, nx, ny, nz, nc, nb, ns_, nEl, nch, ncel, nst, iStar
        integer(c_int64_t) :: npk, per, off
        real(c_float), allocatable, target :: xA(:), yA(:), zA(:), nu(:), sPos(:), dE(:)
        integer(c_int32_t), allocatable, target :: act(:), sIdx(:)
        real(c_float), target :: sc(5)
        real(c_float), allocatable, target :: opac(:, :), rec(:, :), tl(:), spec(:, :)
        real(c_double), allocatable, target :: J(:, :), esc(:, :), Jshort(:, :)
        type(mc_desc) :: d
        type(mc_counters) :: cnt
        type(c_ptr) :: ctx
        real(c_double) :: nseg

        open (newunit=u, file=dir//'/problem.bin', access='stream', form='unformatted', status='old')
        read (u) h
        nx = int(h(1)); ny = int(h(2)); nz = int(h(3)); nc = int(h(4)); nb = int(h(5))
        nst = int(h(7)); npk = h(8)
        allocate (xA(nx), yA(ny), zA(nz), act(nx * ny * nz), nu(nb), sPos(3 * nst), sIdx(3 * nst), dE(nst))
        allocate (opac(0:nc, nb), rec(0:nc, nb), tl(0:nc), spec(nb, nst))
        read (u) xA, yA, zA, act, sc, nu, sPos, sIdx, dE, opac, rec, tl, spec

        d%abi_version = MC_ABI_VERSION
        d%nx = nx; d%ny = ny; d%nz = nz
        d%xAxis = c_loc(xA); d%yAxis = c_loc(yA); d%zAxis = c_loc(zA); d%active = c_loc(act)
        d%nCells = nc; d%nbins = nb; d%flags = int(h(6), c_int32_t)
        d%R_out = sc(1); d%geoCorr = sc(2:4); d%ionEdge1 = sc(5)
        d%nuArray = c_loc(nu); d%nStars = nst
        d%starPos = c_loc(sPos); d%starIdx = c_loc(sIdx); d%deltaE = c_loc(dE)
        d%seed = h(9); d%device = int(h(17), c_int32_t)
        call mc_check(mc_ctx_create(d, ctx), c_null_ptr, 'mc_ctx_create')
        allocate (J(0:nc, nb), esc(0:nc, 0:nb))
        per = npk / nst
        off = 0
        if (grains) then
            ! the estimator the grain balance reads, already scaled
            call read_grains(dir, nc, nb)
            call mc_check(mc_set_estimators(ctx, c_loc(gJ), c_null_ptr), ctx, 'mc_set_estimators')
        else
            call mc_check(mc_upload_tables(ctx, c_loc(opac), c_null_ptr, c_loc(rec), c_loc(tl), &
                                           c_loc(spec), c_null_ptr), ctx, 'mc_upload_tables')
            call mc_check(mc_zero_estimators(ctx), ctx, 'mc_zero_estimators')

            ! iterateMC: each star's packets; packet indices continue across
            ! stars so that every packet has its own stream
            nseg = 0.d0
            do iStar = 1, nst
                call gpu_energy_packet_driver(ctx, iStar, per, off, J, esc, cnt)
                off = off + per
                nseg = nseg + cnt%n_seg
            end do
            call write_dbl(dir//'/J.bin', J, size(J))
            call write_dbl(dir//'/esc.bin', esc, size(esc))
            print *, 'transport ok: packets', off, ' segments', nseg
        end if

        if (h(10) /= 0) then
            ns_ = int(h(11)); nEl = int(h(12)); nch = int(h(13)); ncel = int(h(14))
            call balance(u, ctx, nc, nb, ns_, nEl, nch, ncel, int(h(15)), int(h(16)), &
                         iand(int(h(6)), int(MC_SYMMETRIC)) /= 0, dir, grains)
        end if
        close (u)
        if (short) then
            allocate (Jshort(0:nc - 1, nb))
            print *, 'shape check: a J of ', nc, ' rows'
            call gpu_energy_packet_driver(ctx, 1, per, off, Jshort, esc, cnt)
            print *, 'shape check FAILED: the short J was accepted'
            stop 1
        end if
        call mc_ctx_destroy(ctx)
    end subroutine run

    subroutine balance(u, ctx, nc, nb, ns_, nEl, nch, ncel, ndbl, nflt, sym, dir, grains)
        integer, intent(in) :: u, nc, nb, ns_, nEl, nch, ncel, ndbl, nflt
        type(c_ptr), intent(in) :: ctx
        logical, intent(in) :: sym, grains
        character(len=*), intent(in) :: dir
        integer(c_int32_t), allocatable, target :: elemOn(:), xref(:), chm(:), dri(:), api(:), celm(:)
        real(c_float), allocatable, target :: abun(:), sig(:), rrec(:), rnew(:), fe(:), drr(:), apr(:), celf(:)
        real(c_double), allocatable, target :: celd(:)
        real(c_float) :: xhil
        real, allocatable :: Hden(:), Te(:), Ne(:), ion(:, :, :)
        integer, allocatable :: conv(:), black(:)
        integer :: noHit, noIonBal, noTeBal, w
        type(mc_atomic) :: A
        type(mc_dust_tables) :: D
        real(c_double) :: jscale

        allocate (elemOn(30), xref(30), abun(30), chm(8 * nch), sig(nch * nb), rrec(1800), rnew(3600), fe(30))
        allocate (dri(75), drr(125), api(334), apr(668), celm(6 * ncel), celd(ndbl), celf(nflt))
        read (u) elemOn, xref, abun, chm, sig, rrec, rnew, fe, dri, drr, api, apr, celm, celd, celf, xhil
        A%nstages = ns_; A%nElemUsed = nEl
        A%elemOn = c_loc(elemOn); A%xref = c_loc(xref); A%abun = c_loc(abun)
        A%nch = nch; A%ch_meta = c_loc(chm); A%ch_sigma = c_loc(sig)
        A%rrec = c_loc(rrec); A%rnew = c_loc(rnew); A%fe = c_loc(fe)
        A%dr_int = c_loc(dri); A%dr_real = c_loc(drr); A%ap_int = c_loc(api); A%ap_real = c_loc(apr)
        A%ncel = ncel; A%cel_meta = c_loc(celm); A%cel_dbl = c_loc(celd); A%cel_flt = c_loc(celf)
        A%XHILimit = xhil
        call mc_check(mc_balance_setup(ctx, A), ctx, 'mc_balance_setup')

        allocate (Hden(0:nc), Te(0:nc), Ne(0:nc), ion(0:nc, nEl, ns_), conv(0:nc), black(0:nc))
        read (u) Hden, Te, Ne, ion
        conv = 0
        if (grains) then
            ! dustDriver's tables, then updateCell with getDustT; grains.bin's
            ! J is already scaled
            D%nSpecies = int(gh(1), c_int32_t); D%nSizes = int(gh(2), c_int32_t)
            D%nTemps = int(gh(3), c_int32_t); D%nSpeciesMax = int(gh(4), c_int32_t)
            D%photoelectric = int(gh(5), c_int32_t)
            D%absXsec = c_loc(gAbs); D%grainWeight = c_loc(gWt); D%grainAbun = c_loc(gAbn)
            D%emIntegral = c_loc(gEm); D%grainRadius = c_loc(gRad); D%grainVn = c_loc(gVn)
            D%MsurfAtom = c_loc(gMs); D%aWeight = c_loc(gAw); D%Ndust = c_loc(gNd)
            D%scaXsec = c_null_ptr; D%TdustSublime = c_null_ptr
            D%XHILimit = gXhil
            call mc_check(mc_dust_setup(ctx, D), ctx, 'mc_dust_setup')
            call gpu_update_cells_dust(ctx, 1.d0, .true., Hden, Te, Ne, ion, gTd, conv, black, &
                                       noHit, noIonBal, noTeBal)
            open (newunit=w, file=dir//'/grains_out.bin', access='stream', form='unformatted', status='replace')
            write (w) gTd, Te, Ne, ion, conv, black, noHit, noIonBal, noTeBal
            close (w)
            print *, 'grains ok: noHit', noHit, ' noIonBal', noIonBal, ' noTeBal', noTeBal
            return
        end if
        ! Jste x 1e-9, / 8 in symmetric mode (iteration_mod.f90:706, 717-722)
        jscale = 1.d-9
        if (sym) jscale = jscale / 8.d0
        call gpu_update_cells(ctx, jscale, Hden, Te, Ne, ion, conv, black, noHit, noIonBal, noTeBal)
        open (newunit=w, file=dir//'/balance.bin', access='stream', form='unformatted', status='replace')
        write (w) Te, Ne, ion, conv, black, noHit, noIonBal, noTeBal
        close (w)
        print *, 'balance ok: noHit', noHit, ' noIonBal', noIonBal, ' noTeBal', noTeBal
    end subroutine balance

    ! grains.bin (stream, native endian): int64 header(8) = nSpecies nSizes
    ! nTemps nSpeciesMax photoelectric; then REAL absXsec [nSpecies][nSizes]
    ! [nbins], grainWeight(nSizes), grainAbun(nSpecies), dustEmIntegral(
    ! nSpecies, nSizes, nTemps), grainRadius(nSizes), grainVn(nSpecies),
    ! MsurfAtom(nSpecies), aWeight(30), Ndust(0:nCells), XHILimit,
    ! Tdust(0:nSpeciesMax, 0:nSizes, 0:nCells); DOUBLE J(0:nCells, nbins)
    subroutine read_grains(dir, nc, nb)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nc, nb
        integer :: g, nS, nSz, nT, nsm
        open (newunit=g, file=dir//'/grains.bin', access='stream', form='unformatted', status='old')
        read (g) gh
        nS = int(gh(1)); nSz = int(gh(2)); nT = int(gh(3)); nsm = int(gh(4))
        allocate (gAbs(nS * nSz * nb), gWt(nSz), gAbn(nS), gEm(nS * nSz * nT), gRad(nSz), gVn(nS), gMs(nS))
        allocate (gAw(30), gNd(0:nc), gTd(0:nsm, 0:nSz, 0:nc), gJ(0:nc, nb))
        read (g) gAbs, gWt, gAbn, gEm, gRad, gVn, gMs, gAw, gNd, gXhil, gTd, gJ
        close (g)
    end subroutine read_grains

    subroutine write_dbl(path, a, n)
        character(len=*), intent(in) :: path
        integer, intent(in) :: n
        real(c_double), intent(in) :: a(n)
        integer :: w
        open (newunit=w, file=path, access='stream', form='unformatted', status='replace')
        write (w) a
        close (w)
    end subroutine write_dbl

end program mc_fortran_check
