! mocassin_amd_mod.f90 - Fortran host interface to libmocassin_amd.so
! (include/mocassin_amd.h) through ISO_C_BINDING.
!
! Two layers:
!   * the raw bind(C) interfaces and interoperable types, one per C entry
!     point, for hosts that manage their own c_ptr/c_loc plumbing;
!   * host-side wrappers shaped like the reference's call sites, which take
!     Fortran arrays in the reference's own layouts and turn a non-zero
!     status into the reference's fail-stop (print + stop):
!       gpu_energy_packet_driver  ~ energyPacketDriver(iStar, n, grid)
!                                    (photon_mod.f90:26-284, called from
!                                    iterateMC, iteration_mod.f90:488-493)
!       gpu_update_cells          ~ the per-cell updateCell loop of iterateMC
!                                    (iteration_mod.f90:826-837) with its
!                                    lgConverged/lgBlack bookkeeping
!                                    (iteration_mod.f90:848-876)
! REAL arrays of the host are passed to the C side as they are; the balance
! state is widened to double on the way in and narrowed back on the way out.
module mocassin_amd_mod
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: MC_ABI_VERSION = 3
    integer(c_int32_t), parameter :: MC_SYMMETRIC = 1, MC_DUST = 2, MC_GAS = 4, &
                                     MC_ISOTROPIC = 8, MC_MULTISTARS = 16
    integer(c_int), parameter :: MC_OK = 0, MC_EINVAL = 1, MC_ENODEV = 2, MC_ENOMEM = 3, &
                                 MC_ESTATE = 4, MC_ETABLE = 5, MC_EPACKET = 6, MC_EUNSUP = 7

    type, bind(C) :: mc_desc
        integer(c_int32_t) :: abi_version, nx, ny, nz
        type(c_ptr) :: xAxis, yAxis, zAxis, active
        integer(c_int32_t) :: nCells, nbins, flags
        real(c_float) :: R_out, geoCorr(3), ionEdge1
        type(c_ptr) :: nuArray
        integer(c_int32_t) :: nStars
        type(c_ptr) :: starPos, starIdx, deltaE
        integer(c_int64_t) :: seed          ! uint64 on the C side (same bits)
        integer(c_int32_t) :: device
    end type mc_desc

    type, bind(C) :: mc_counters
        real(c_double) :: n_seg, n_abs, n_sca, n_esc, n_lost, n_line, n_trapped, n_flights, &
                          qphot, kernel_ms, n_walk_passes, n_walk_lanes, n_emit_passes, n_emit_lanes
    end type mc_counters

    type, bind(C) :: mc_tuning
        integer(c_int32_t) :: blocks_per_cu, threads, emit_min, walk_steps, kernel
    end type mc_tuning
    integer(c_int32_t), parameter :: MC_KERNEL_AUTO = 0, MC_KERNEL_ONE = 1, MC_KERNEL_SPARE = 2, &
                                     MC_KERNEL_POOL = 3

    integer(c_int32_t), parameter :: MC_SOLVER_REFERENCE = 0, MC_SOLVER_SECANT = 1

    type, bind(C) :: mc_balance_opts
        integer(c_int32_t) :: solver, max_iter
        real(c_double) :: tol, ion_tol
        integer(c_int32_t) :: warm_start
        real(c_double) :: first_step, first_slope
        integer(c_int32_t) :: warm_iter
    end type mc_balance_opts

    type, bind(C) :: mc_batch_tuning
        integer(c_int32_t) :: waves_per_cu, emit_min, walk_steps
        integer(c_int64_t) :: chunk
    end type mc_batch_tuning

    type, bind(C) :: mc_atomic
        integer(c_int32_t) :: nstages, nElemUsed
        type(c_ptr) :: elemOn, xref, abun
        integer(c_int32_t) :: nch
        type(c_ptr) :: ch_meta, ch_sigma, rrec, rnew, fe, dr_int, dr_real, ap_int, ap_real
        integer(c_int32_t) :: ncel
        type(c_ptr) :: cel_meta, cel_dbl, cel_flt
        real(c_float) :: XHILimit
    end type mc_atomic

    type, bind(C) :: mc_dust_tables
        integer(c_int32_t) :: nSpecies, nSpeciesMax, nSizes, nTemps, photoelectric
        type(c_ptr) :: absXsec, grainWeight, grainAbun, emIntegral
        type(c_ptr) :: grainRadius, grainVn, MsurfAtom, aWeight, Ndust, scaXsec, TdustSublime
        real(c_float) :: XHILimit
    end type mc_dust_tables

    type, bind(C) :: mc_opacity_tables
        integer(c_int32_t) :: nseg
        type(c_ptr) :: seg_meta, seg_sigma
        real(c_float) :: brems1
    end type mc_opacity_tables

    type, bind(C) :: mc_emission_tables
        integer(c_int32_t) :: ntk, nHI, nHeI, nHeII
        type(c_ptr) :: tkGamma, logGammaHI, logGammaHeI, logGammaHeII
        type(c_ptr) :: HINuEdgeP, HeINuEdgeP, HeIINuEdgeP
        type(c_ptr) :: gauntFF, gauntFFHeII
        integer(c_int32_t) :: HIPnuP, HeIPnuP, HeIIPnuP
        type(c_ptr) :: xSecHI, xSecHeI, xSecHeII
        integer(c_int32_t) :: nfb
        type(c_ptr) :: fb_meta, fb_gratio, fb_sigma
        type(c_ptr) :: y_dat, Ay_dat
        type(c_ptr) :: HIRecLineData, HeIIRecLineData
        type(c_ptr) :: HeIrecLineCoeff
        integer(c_int32_t) :: nHeIILyman
        type(c_ptr) :: HeIILymanData, HeIILymanNu
        type(c_ptr) :: widFlx
    end type mc_emission_tables

    ! the viewing-angle tables initCartesianGrid sets (grid_mod.f90:416-467)
    type, bind(C) :: mc_view_angles
        integer(c_int32_t) :: nAngleBins, totAngleBinsTheta, totAngleBinsPhi
        real(c_float) :: dTheta, dPhi
        type(c_ptr) :: viewPointPtheta, viewPointPphi, viewPointTheta, viewPointPhi
    end type mc_view_angles

    type, bind(C) :: mc_grid_state
        type(c_ptr) :: Te, Ne, Hden, ionDen, lgConverged, lgBlack, Ndust, Tdust
        integer(c_int32_t) :: nSpeciesMax, nSizes
        real(c_float) :: totalDustMass
    end type mc_grid_state

    interface
        integer(c_int) function mc_ctx_create(desc, ctx) bind(C, name="mc_ctx_create")
            import :: c_int, c_ptr, mc_desc
            type(mc_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function mc_ctx_create

        integer(c_int) function mc_upload_tables(ctx, opacity, scaOpac, recPDF, totalLines, &
                                                 specCDF, gSca) bind(C, name="mc_upload_tables")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, opacity, scaOpac, recPDF, totalLines, specCDF, gSca
        end function mc_upload_tables

        integer(c_int) function mc_zero_estimators(ctx) bind(C, name="mc_zero_estimators")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function mc_zero_estimators

        integer(c_int) function mc_transport_run(ctx, istar, npackets, packet_offset, J_out, &
                                                 esc_out, cnt) bind(C, name="mc_transport_run")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, mc_counters
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: istar
            integer(c_int64_t), value :: npackets, packet_offset
            type(c_ptr), value :: J_out, esc_out
            type(mc_counters), intent(out) :: cnt
        end function mc_transport_run

        integer(c_int) function mc_fetch_estimators(ctx, J_out, esc_out) &
                bind(C, name="mc_fetch_estimators")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, J_out, esc_out
        end function mc_fetch_estimators

        integer(c_int) function mc_set_estimators(ctx, J, esc) bind(C, name="mc_set_estimators")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, J, esc
        end function mc_set_estimators

        integer(c_int) function mc_set_view_angles(ctx, v) bind(C, name="mc_set_view_angles")
            import :: c_int, c_ptr, mc_view_angles
            type(c_ptr), value :: ctx
            type(mc_view_angles), intent(in) :: v
        end function mc_set_view_angles

        integer(c_int) function mc_esc_planes(ctx) bind(C, name="mc_esc_planes")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function mc_esc_planes

        integer(c_int) function mc_ctx_shape(ctx, nCells, nbins) bind(C, name="mc_ctx_shape")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: nCells, nbins
        end function mc_ctx_shape

        integer(c_int) function mc_probe_atomics(device, words, requests, lanes, fp64, kernel_ms) &
                bind(C, name="mc_probe_atomics")
            import :: c_int, c_int32_t, c_int64_t, c_double
            integer(c_int32_t), value :: device
            integer(c_int64_t), value :: words, requests
            integer(c_int32_t), value :: lanes, fp64
            real(c_double), intent(out) :: kernel_ms
        end function mc_probe_atomics

        integer(c_int) function mc_probe_atomic_lines(device, words, requests, group, spread, kernel_ms) &
                bind(C, name="mc_probe_atomic_lines")
            import :: c_int, c_int32_t, c_int64_t, c_double
            integer(c_int32_t), value :: device
            integer(c_int64_t), value :: words, requests
            integer(c_int32_t), value :: group, spread
            real(c_double), intent(out) :: kernel_ms
        end function mc_probe_atomic_lines

        integer(c_int) function mc_walk_replay(ctx, n, pos, dir, cell, cell_seq, ds_seq, &
                                               seq_len, status, max_len) bind(C, name="mc_walk_replay")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            type(c_ptr), value :: pos, dir, cell, cell_seq, ds_seq, seq_len, status
            integer(c_int32_t), value :: max_len
        end function mc_walk_replay

        integer(c_int) function mc_balance_setup(ctx, atomic) bind(C, name="mc_balance_setup")
            import :: c_int, c_ptr, mc_atomic
            type(c_ptr), value :: ctx
            type(mc_atomic), intent(in) :: atomic
        end function mc_balance_setup

        integer(c_int) function mc_balance_run(ctx, jscale, opts, Hden, Te, Ne, ionDen, flags, &
                                               nIter, kernel_ms) bind(C, name="mc_balance_run")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: jscale
            type(c_ptr), value :: opts, Hden, Te, Ne, ionDen, flags, nIter, kernel_ms
        end function mc_balance_run

        integer(c_int) function mc_dust_setup(ctx, dust) bind(C, name="mc_dust_setup")
            import :: c_int, c_ptr, mc_dust_tables
            type(c_ptr), value :: ctx
            type(mc_dust_tables), intent(in) :: dust
        end function mc_dust_setup

        integer(c_int) function mc_update_cells(ctx, jscale, opts, Hden, Te, Ne, ionDen, Tdust, flags, &
                                                nIter, kernel_ms) bind(C, name="mc_update_cells")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: jscale
            type(c_ptr), value :: opts, Hden, Te, Ne, ionDen, Tdust, flags, nIter, kernel_ms
        end function mc_update_cells

        integer(c_int) function mc_opacity_setup(ctx, tables) bind(C, name="mc_opacity_setup")
            import :: c_int, c_ptr, mc_opacity_tables
            type(c_ptr), value :: ctx
            type(mc_opacity_tables), intent(in) :: tables
        end function mc_opacity_setup

        integer(c_int) function mc_opacity_build(ctx, Hden, Te, Ne, ionDen, Tdust, opacity, scaOpac, &
                                                 kernel_ms) bind(C, name="mc_opacity_build")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, Hden, Te, Ne, ionDen, Tdust, opacity, scaOpac, kernel_ms
        end function mc_opacity_build

        integer(c_int) function mc_write_grid(desc, atomic, dir, state) bind(C, name="mc_write_grid")
            import :: c_int, c_char, mc_desc, mc_atomic, mc_grid_state
            type(mc_desc), intent(in) :: desc
            type(mc_atomic), intent(in) :: atomic
            character(kind=c_char), dimension(*), intent(in) :: dir
            type(mc_grid_state), intent(in) :: state
        end function mc_write_grid

        integer(c_int) function mc_read_grid(desc, atomic, dir, state) bind(C, name="mc_read_grid")
            import :: c_int, c_char, mc_desc, mc_atomic, mc_grid_state
            type(mc_desc), intent(in) :: desc
            type(mc_atomic), intent(in) :: atomic
            character(kind=c_char), dimension(*), intent(in) :: dir
            type(mc_grid_state), intent(in) :: state
        end function mc_read_grid

        integer(c_int) function mc_emission_setup(ctx, tables) bind(C, name="mc_emission_setup")
            import :: c_int, c_ptr, mc_emission_tables
            type(c_ptr), value :: ctx
            type(mc_emission_tables), intent(in) :: tables
        end function mc_emission_setup

        integer(c_int) function mc_emission_build(ctx, Hden, Te, Ne, ionDen, Tdust, recPDF, totalLines, &
                                                  kernel_ms) bind(C, name="mc_emission_build")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, Hden, Te, Ne, ionDen, Tdust, recPDF, totalLines, kernel_ms
        end function mc_emission_build

        integer(c_int) function mc_line_luminosities(ctx, Hden, Te, Ne, ionDen, lgBlack, forbVol, HbetaVol, &
                                                     kernel_ms) bind(C, name="mc_line_luminosities")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, Hden, Te, Ne, ionDen, lgBlack, forbVol, HbetaVol, kernel_ms
        end function mc_line_luminosities

        integer(c_int) function mc_mlp_set(ctx, params, scale) bind(C, name="mc_mlp_set")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, params, scale
        end function mc_mlp_set

        integer(c_int) function mc_mlp_get(ctx, params, scale) bind(C, name="mc_mlp_get")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, params, scale
        end function mc_mlp_get

        integer(c_int) function mc_mlp_features(ctx, jscale, Hden, X, hit) bind(C, name="mc_mlp_features")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: jscale
            type(c_ptr), value :: Hden, X, hit
        end function mc_mlp_features

        integer(c_int) function mc_mlp_predict(ctx, X, n, Y) bind(C, name="mc_mlp_predict")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx, X
            integer(c_int64_t), value :: n
            type(c_ptr), value :: Y
        end function mc_mlp_predict

        integer(c_int) function mc_mlp_loss_grad(ctx, X, T, M, n, grad, loss) bind(C, name="mc_mlp_loss_grad")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx, X, T, M
            integer(c_int64_t), value :: n
            type(c_ptr), value :: grad, loss
        end function mc_mlp_loss_grad

        integer(c_int) function mc_mlp_train(ctx, X, T, M, n, order, steps, batch, lr, loss, kernel_ms) &
                bind(C, name="mc_mlp_train")
            import :: c_int, c_int32_t, c_int64_t, c_float, c_ptr
            type(c_ptr), value :: ctx, X, T, M
            integer(c_int64_t), value :: n
            type(c_ptr), value :: order
            integer(c_int32_t), value :: steps, batch
            real(c_float), value :: lr
            type(c_ptr), value :: loss, kernel_ms
        end function mc_mlp_train

        integer(c_int) function mc_transport_batch(ctx, n, istar, npackets, packet_offset, tuning, cnt, &
                                                   kernel_ms) bind(C, name="mc_transport_batch")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx           ! array of n context handles
            integer(c_int32_t), value :: n
            type(c_ptr), value :: istar, npackets, packet_offset, tuning, cnt, kernel_ms
        end function mc_transport_batch

        integer(c_int) function mc_transport_batch_noalloc(ctx, n, istar, npackets, packet_offset, tuning, cnt, &
                                                           kernel_ms) bind(C, name="mc_transport_batch_noalloc")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx           ! array of n context handles
            integer(c_int32_t), value :: n
            type(c_ptr), value :: istar, npackets, packet_offset, tuning, cnt, kernel_ms
        end function mc_transport_batch_noalloc

        integer(c_int) function mc_count_crossings(ctx, enable) bind(C, name="mc_count_crossings")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: enable
        end function mc_count_crossings

        integer(c_int) function mc_fetch_crossings(ctx, crossings) bind(C, name="mc_fetch_crossings")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, crossings
        end function mc_fetch_crossings

        integer(c_int) function mc_set_tuning(ctx, t) bind(C, name="mc_set_tuning")
            import :: c_int, c_ptr, mc_tuning
            type(c_ptr), value :: ctx
            type(mc_tuning), intent(in) :: t
        end function mc_set_tuning

        integer(c_int) function mc_get_tuning(ctx, t) bind(C, name="mc_get_tuning")
            import :: c_int, c_ptr, mc_tuning
            type(c_ptr), value :: ctx
            type(mc_tuning), intent(out) :: t
        end function mc_get_tuning

        type(c_ptr) function mc_ctx_last_error(ctx) bind(C, name="mc_ctx_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function mc_ctx_last_error

        subroutine mc_ctx_destroy(ctx) bind(C, name="mc_ctx_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine mc_ctx_destroy

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

    ! energyPacketDriver by the rank of esc: plane 0 only, or with viewing angles
    interface gpu_energy_packet_driver
        module procedure gpu_energy_packet_driver_plane, gpu_energy_packet_driver_views
    end interface gpu_energy_packet_driver

contains

    ! message of the last failure on ctx (c_null_ptr: of the last failed create)
    function mc_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        integer :: i, n
        msg = ' '
        p = mc_ctx_last_error(ctx)
        if (.not. c_associated(p)) return
        n = int(min(c_strlen(p), int(len(msg), c_size_t)))
        call c_f_pointer(p, s, [max(n, 1)])
        do i = 1, n
            msg(i:i) = s(i)
        end do
    end function mc_error_text

    ! the reference's fail-stop: print the message and stop
    subroutine mc_check(rc, ctx, where)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: where
        if (rc /= MC_OK) then
            print *, '! ', where, ': ', trim(mc_error_text(ctx)), ' (code ', rc, ')'
            stop
        end if
    end subroutine mc_check

    ! J must be (0:nCells, 1:nbins) and esc (0:nCells, 0:nbins, planes) for
    ! the library to copy into them; anything else stops before the copy
    subroutine check_estimator_shape(ctx, sJ, sE)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: sJ(2), sE(3)
        integer(c_int32_t) :: nc, nb
        call mc_check(mc_ctx_shape(ctx, nc, nb), ctx, 'mc_ctx_shape')
        if (sJ(1) /= nc + 1 .or. sJ(2) /= nb) then
            print *, '! energyPacketDriver: J is ', sJ, '; the context needs ', nc + 1, nb
            stop
        end if
        if (sE(1) /= nc + 1 .or. sE(2) /= nb + 1) then
            print *, '! energyPacketDriver: esc is ', sE(1:2), '; the context needs ', nc + 1, nb + 1
            stop
        end if
    end subroutine check_estimator_shape

    ! energyPacketDriver(iStar, n, grid): transport n packets of star iStar
    ! (packet indices offset+1 .. offset+n), then copy the running Jste
    ! (0:nCells,1:nbins) and escapedPackets(0:nCells,0:nbins) totals, in
    ! double, to J and esc.  Without viewing angles only: with them, pass
    ! esc(0:nCells,0:nbins,0:nAngleBins) (gpu_energy_packet_driver_views,
    ! the same generic name).
    subroutine gpu_energy_packet_driver_plane(ctx, iStar, n, offset, J, esc, cnt)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: iStar
        integer(c_int64_t), intent(in) :: n, offset
        real(c_double), intent(inout), contiguous, target :: J(:, :), esc(:, :)
        type(mc_counters), intent(out) :: cnt
        if (mc_esc_planes(ctx) /= 1) then
            print *, '! energyPacketDriver: viewing angles are on; esc needs its (0:nAngleBins) planes'
            stop
        end if
        call check_estimator_shape(ctx, shape(J), [size(esc, 1), size(esc, 2), 1])
        call mc_check(mc_transport_run(ctx, int(iStar, c_int32_t), n, offset, c_loc(J), &
                                       c_loc(esc), cnt), ctx, 'energyPacketDriver')
    end subroutine gpu_energy_packet_driver_plane

    subroutine gpu_energy_packet_driver_views(ctx, iStar, n, offset, J, esc, cnt)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: iStar
        integer(c_int64_t), intent(in) :: n, offset
        real(c_double), intent(inout), contiguous, target :: J(:, :), esc(:, :, :)
        type(mc_counters), intent(out) :: cnt
        if (size(esc, 3) /= mc_esc_planes(ctx)) then
            print *, '! energyPacketDriver: esc has ', size(esc, 3), ' planes; the context ', mc_esc_planes(ctx)
            stop
        end if
        call check_estimator_shape(ctx, shape(J), shape(esc))
        call mc_check(mc_transport_run(ctx, int(iStar, c_int32_t), n, offset, c_loc(J), &
                                       c_loc(esc), cnt), ctx, 'energyPacketDriver')
    end subroutine gpu_energy_packet_driver_views

    ! The updateCell loop over every active cell: Te, Ne (0:nCells) and
    ! ionDen(0:nCells, nElementsUsed, nstages) are REAL as the host keeps
    ! them.  lgBlack(0:nCells) receives 1/0 per cell; lgConverged keeps its
    ! value in cells no packet reached, as updateCell leaves it there
    ! (update_mod.f90:100-154, 510-515); noHit, noIonBal, noTeBal count the
    ! cells the reference counts (iteration_mod.f90:848-855).
    subroutine gpu_update_cells(ctx, jscale, Hden, Te, Ne, ionDen, lgConverged, lgBlack, &
                                noHit, noIonBal, noTeBal)
        type(c_ptr), intent(in) :: ctx
        real(c_double), intent(in) :: jscale
        real, intent(in) :: Hden(0:)
        real, intent(inout) :: Te(0:), Ne(0:), ionDen(0:, :, :)
        integer, intent(inout) :: lgConverged(0:)
        integer, intent(out) :: lgBlack(0:)
        integer, intent(out) :: noHit, noIonBal, noTeBal
        real(c_double), allocatable, target :: dH(:), dT(:), dN(:), dI(:, :, :)
        integer(c_int32_t), allocatable, target :: fl(:), nIt(:)
        integer :: nc1
        nc1 = size(Te)
        allocate (dH(0:nc1 - 1), dT(0:nc1 - 1), dN(0:nc1 - 1), fl(0:nc1 - 1), nIt(0:nc1 - 1))
        allocate (dI(0:nc1 - 1, size(ionDen, 2), size(ionDen, 3)))
        dH = real(Hden, c_double)
        dT = real(Te, c_double)
        dN = real(Ne, c_double)
        dI = real(ionDen, c_double)
        call mc_check(mc_balance_run(ctx, jscale, c_null_ptr, c_loc(dH), c_loc(dT), c_loc(dN), &
                                     c_loc(dI), c_loc(fl), c_loc(nIt), c_null_ptr), ctx, 'updateCell')
        Te = real(dT)
        Ne = real(dN)
        ionDen = real(dI)
        where (iand(fl, 16) == 0) lgConverged = iand(fl, 1)
        lgBlack = ishft(iand(fl, 2), -1)
        lgBlack(0) = 0
        noHit = count(iand(fl(1:), 16) /= 0)
        noIonBal = count(iand(fl(1:), 8) /= 0)
        noTeBal = count(iand(fl(1:), 4) /= 0)
        deallocate (dH, dT, dN, dI, fl, nIt)
    end subroutine gpu_update_cells

    ! The updateCell loop with grains (mc_dust_setup first): as
    ! gpu_update_cells for the gas, and Tdust(0:nSpeciesMax, 0:nSizes,
    ! 0:nCells), REAL as the host keeps it, rewritten in every reached cell
    ! (getDustT, update_mod.f90:1836-1945).  Without gas (lgGas false) the
    ! gas arrays are not read and lgConverged follows Tdust(0,0)
    ! (update_mod.f90:306-324).
    subroutine gpu_update_cells_dust(ctx, jscale, lgGas, Hden, Te, Ne, ionDen, Tdust, lgConverged, &
                                     lgBlack, noHit, noIonBal, noTeBal)
        type(c_ptr), intent(in) :: ctx
        real(c_double), intent(in) :: jscale
        logical, intent(in) :: lgGas
        real, intent(in) :: Hden(0:)
        real, intent(inout) :: Te(0:), Ne(0:), ionDen(0:, :, :), Tdust(0:, 0:, 0:)
        integer, intent(inout) :: lgConverged(0:)
        integer, intent(out) :: lgBlack(0:)
        integer, intent(out) :: noHit, noIonBal, noTeBal
        real(c_double), allocatable, target :: dH(:), dT(:), dN(:), dI(:, :, :), dD(:, :, :)
        integer(c_int32_t), allocatable, target :: fl(:), nIt(:)
        integer :: nc1
        nc1 = size(Tdust, 3)
        allocate (fl(0:nc1 - 1), nIt(0:nc1 - 1))
        allocate (dD(0:size(Tdust, 1) - 1, 0:size(Tdust, 2) - 1, 0:nc1 - 1))
        dD = real(Tdust, c_double)
        if (lgGas) then
            allocate (dH(0:nc1 - 1), dT(0:nc1 - 1), dN(0:nc1 - 1))
            allocate (dI(0:nc1 - 1, size(ionDen, 2), size(ionDen, 3)))
            dH = real(Hden, c_double)
            dT = real(Te, c_double)
            dN = real(Ne, c_double)
            dI = real(ionDen, c_double)
            call mc_check(mc_update_cells(ctx, jscale, c_null_ptr, c_loc(dH), c_loc(dT), c_loc(dN), &
                                          c_loc(dI), c_loc(dD), c_loc(fl), c_loc(nIt), c_null_ptr), &
                          ctx, 'updateCell')
            Te = real(dT)
            Ne = real(dN)
            ionDen = real(dI)
            deallocate (dH, dT, dN, dI)
        else
            call mc_check(mc_update_cells(ctx, jscale, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                          c_null_ptr, c_loc(dD), c_loc(fl), c_loc(nIt), c_null_ptr), &
                          ctx, 'updateCell')
        end if
        Tdust = real(dD)
        where (iand(fl, 16) == 0) lgConverged = iand(fl, 1)
        lgBlack = ishft(iand(fl, 2), -1)
        lgBlack(0) = 0
        noHit = count(iand(fl(1:), 16) /= 0)
        noIonBal = count(iand(fl(1:), 8) /= 0)
        noTeBal = count(iand(fl(1:), 4) /= 0)
        deallocate (fl, nIt, dD)
    end subroutine gpu_update_cells_dust

    ! outputGas's analytic forbidden-line and Hbeta luminosities of the grid's
    ! state (output_mod.f90:302-520), before the division by HbetaVol:
    ! forbVol(lower, upper, model ion in forLines order)
    subroutine gpu_line_luminosities(ctx, Hden, Te, Ne, ionDen, lgBlack, forbVol, HbetaVol)
        type(c_ptr), intent(in) :: ctx
        real, intent(in), target :: Hden(0:), Te(0:), Ne(0:), ionDen(0:, :, :)
        integer(c_int32_t), intent(in), target :: lgBlack(0:)
        real(c_double), intent(out), target :: forbVol(:, :, :)
        real(c_double), intent(out), target :: HbetaVol
        call mc_check(mc_line_luminosities(ctx, c_loc(Hden), c_loc(Te), c_loc(Ne), c_loc(ionDen), &
                                           c_loc(lgBlack), c_loc(forbVol), c_loc(HbetaVol), c_null_ptr), &
                      ctx, 'outputGas lines')
    end subroutine gpu_line_luminosities

end module mocassin_amd_mod
