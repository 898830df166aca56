! pt_gpu.f90 -- parallel-tempering driver with the reference's command line
!     pt_gpu Jmin Jmax NJ Dmin Dmax ND Bmin Bmax NB outputPath
! (src/PT.f90:93-138) and its output files (src/PT.f90:297-364), running
! every (J, D, B) slot x temperature replica in ONE device context: each
! parameter slot is a PT ladder of numTemps replicas, all swept by the same
! kernel launches, swaps done on device.
!
! Schedule as the reference (src/PT.f90:46-48, 219-287): numIterations x
! (numMCSSweepsPerSwap sweeps + one replica-exchange round), then 1000 relax
! sweeps without swaps, then statistics of each replica at each temperature.
! Environment variables ISING_PT_ITERATIONS, ISING_PT_SWEEPS_PER_SWAP and
! ISING_PT_RELAX override the three counts (for short test runs).
!
! Lattice as the reference's: 30 x 30 x 6 bcc cells (src/PT.f90:36-38,145)
! of edge a = 2 (bcc_unit_cell overwrites a_bravais, src/unit_cells.f90:21).
! Differences from the reference, all deliberate: seeded random-sphere start
! (the reference never sets the spins); swaps use the same beta = 1/(Kb T) as the
! sweeps and the textbook exponent (the reference uses 1/T and the opposite
! sign, src/PT.f90:256-259); one even/odd round of neighbour swaps per
! interval instead of numTemps random picks; no MPI ranks (one GPU holds all
! slots).
program pt_gpu
        use iso_c_binding
        use iso_fortran_env, only: error_unit, dp => real64
        use ising_gpu_c
        use PT_Utils, only: indicesFromSlot
        use reciprocal_space_processes, only: write_density_line
        implicit none

        real(dp), parameter :: Kb = 8.6173303e-5_dp
        real(dp), parameter :: TMax = 20000_dp, TMin = 400_dp
        integer, parameter :: numTemps = 10
        integer, parameter :: numCellsX = 30, numCellsY = 30, numCellsZ = 6
        real(dp), parameter :: a_bravais = 2.0_dp
        integer, parameter :: atomsPerCell = 2
        real(dp), parameter :: cone_r = 0.2_dp

        integer :: numIterations, numMCSSweepsPerSwap, numRelax
        real(dp) :: JMin, JMax, DMin, DMax, BMin, BMax
        integer :: NJ, ND, NB, NumParams, i, j, k, JIndex, DIndex, BIndex, slot, mid
        character(len=400) :: outputPath, outputPathSpins, outputPathDensity, outputPathDensityInterpolated
        character(len=500) :: string_buff
        character(len=:), allocatable :: output_string, filename_string
        real(dp), allocatable :: ParamArray(:,:)
        real(c_double), allocatable, target :: betas(:)
        real(dp), allocatable :: TemperatureArray(:)
        type(ising_ladder_desc), allocatable :: ladders(:)
        type(c_ptr) :: ctx
        integer(c_int) :: rc
        integer(c_int32_t), allocatable :: slot_of_temp(:)
        integer(c_int64_t), allocatable :: q(:)
        real(c_double), allocatable :: qr(:), mag(:), spins(:), density(:), density_fine(:)
        type(ising_skyrmion_params) :: sk_params
        integer(c_int32_t) :: sk_counts(1)
        integer :: scl, jx, ky
        type(ising_schedule) :: sched
        type(ising_series) :: series
        integer :: fileHandle, nsites, ix, iy, iz, site, atom
        real(dp) :: winding_middle, winding_spread

        if (command_argument_count() /= 10) then
                write(error_unit, '(A)') "Incorrect number of command line arguments, &
                        & input: Jmin Jmax NJ Dmin Dmax ND Bmin Bmax NB outputFilePath"
                error stop 1
        end if
        call arg_real(1, JMin); call arg_real(2, JMax); call arg_int(3, NJ)
        call arg_real(4, DMin); call arg_real(5, DMax); call arg_int(6, ND)
        call arg_real(7, BMin); call arg_real(8, BMax); call arg_int(9, NB)
        outputPath = " "
        call get_command_argument(10, outputPath)
        numIterations = env_int("ISING_PT_ITERATIONS", 40)
        numMCSSweepsPerSwap = env_int("ISING_PT_SWEEPS_PER_SWAP", 100)
        numRelax = env_int("ISING_PT_RELAX", 1000)
        ! upsampling of the interpolated density (makeChainMesh's sclx = scly = 10,
        ! src/chainMesh.f90:764-765)
        scl = env_int("ISING_PT_INTERP_SCALE", 10)

        NumParams = NJ * ND * NB
        if (NumParams < 1) error stop "ERROR: Number of Parameter Slots must be at least one"
        ! Parameter values of each slot (src/PT.f90:170-176).
        allocate(ParamArray(NumParams, 3))
        do i = 1, NumParams
                call indicesFromSlot(i, NJ, ND, NB, JIndex, DIndex, BIndex)
                ParamArray(i, 1) = dble(JIndex - 1) / dble(NJ) * (JMax - JMin) + JMin
                ParamArray(i, 2) = dble(DIndex - 1) / dble(ND) * (DMax - DMin) + DMin
                ParamArray(i, 3) = dble(BIndex - 1) / dble(NB) * (BMax - BMin) + BMin
        end do
        ! Geometric ladder (src/PT.f90:202).
        allocate(TemperatureArray(numTemps), betas(numTemps * NumParams))
        do j = 1, numTemps
                TemperatureArray(j) = TMin * (TMax / TMin) ** (dble(j - 1) / dble(numTemps - 1))
        end do
        allocate(ladders(NumParams))
        do i = 1, NumParams
                do j = 1, numTemps
                        betas((i - 1) * numTemps + j) = 1.0_dp / (Kb * TemperatureArray(j))
                end do
                ladders(i)%lattice = ising_lattice(numCellsX, numCellsY, numCellsZ, [1, 1, 1], ISING_CELL_BCC)
                ! Reference parameter convention: +J S.S exchange, interfacial DMI.
                ladders(i)%ham = ising_hamiltonian(-ParamArray(i, 1), ParamArray(i, 2), ISING_DMI_INTERFACIAL_REF, 0, &
                                                   ParamArray(i, 3), 0.0_dp, 2.0_dp)
                ladders(i)%n_replicas = numTemps
                ladders(i)%reserved = 0
                ladders(i)%cone_radius = cone_r
                ladders(i)%betas = c_loc(betas((i - 1) * numTemps + 1))
        end do
        rc = ising_create(ctx, ladders, int(NumParams, c_int), 20261203_c_int64_t, 0_c_int)
        call ising_check(rc, c_null_ptr, "pt_gpu: create")
        rc = ising_init_random(ctx, -1_c_int)
        call ising_check(rc, ctx, "pt_gpu: init")

        ! Sweeps + swaps, then relaxation without swaps (src/PT.f90:219-287).
        sched = ising_schedule(numIterations * numMCSSweepsPerSwap, numMCSSweepsPerSwap, 0, 0, 1, 0)
        rc = ising_pt_run(ctx, sched, series)
        call ising_check(rc, ctx, "pt_gpu: PT iterations")
        print *, "Completed", numIterations, "iterations of", numMCSSweepsPerSwap, "sweeps on", &
                 NumParams * numTemps, "replicas"
        if (numRelax > 0) then
                rc = ising_sweep(ctx, int(numRelax, c_int))
                call ising_check(rc, ctx, "pt_gpu: relax")
        end if

        ! Statistics and files (src/PT.f90:297-364).
        nsites = numCellsX * numCellsY * numCellsZ * atomsPerCell
        allocate(slot_of_temp(NumParams * numTemps), q(NumParams * numTemps * numCellsZ), &
                 qr(NumParams * numTemps * numCellsZ), mag(3 * NumParams * numTemps), spins(3 * nsites))
        rc = ising_temperature_map(ctx, slot_of_temp)
        call ising_check(rc, ctx, "pt_gpu: temperature map")
        rc = ising_winding_all(ctx, q, qr)
        call ising_check(rc, ctx, "pt_gpu: winding")
        rc = ising_magnetisation(ctx, mag)
        call ising_check(rc, ctx, "pt_gpu: magnetisation")
        outputPathSpins = trim(adjustl(outputPath)) // "_spins"
        outputPathDensity = trim(adjustl(outputPath)) // "_density"
        outputPathDensityInterpolated = trim(adjustl(outputPath)) // "_density_interpolated"
        call execute_command_line("mkdir -p " // trim(adjustl(outputPath)))
        call execute_command_line("mkdir -p " // trim(adjustl(outputPathSpins)))
        call execute_command_line("mkdir -p " // trim(adjustl(outputPathDensity)))
        call execute_command_line("mkdir -p " // trim(adjustl(outputPathDensityInterpolated)))
        allocate(density(numCellsX * numCellsY), density_fine(scl * numCellsX * scl * numCellsY))
        ! skyrmion count of the middle layer as chainMesh_statistics takes it
        ! (src/PT-Utils.f90:54-56): N = 1, thresholds 1e-5..0.95 in 60 steps
        sk_params%scale_x = scl
        sk_params%scale_y = scl
        sk_params%n_particles = 1
        sk_params%num_thresholds = 60
        sk_params%num_sigmas = 20
        sk_params%min_threshold = 1e-5_dp
        sk_params%max_threshold = 0.95_dp
        mid = max(1, numCellsZ / 2)
        do i = 1, NumParams
                do j = 1, numTemps
                        slot = slot_of_temp((i - 1) * numTemps + j)
                        k = slot * numCellsZ
                        winding_middle = qr(k + mid)
                        winding_spread = abs(maxval(qr(k + 1:k + numCellsZ)) - minval(qr(k + 1:k + numCellsZ)))
                        rc = ising_skyrmion_distribution(ctx, int(slot, c_int), int(mid - 1, c_int), sk_params, &
                                                         sk_counts, c_null_ptr)
                        call ising_check(rc, ctx, "pt_gpu: skyrmion count")
                        output_string = "J,D,B,T,winding_number_middle,&
                                &skyrmion_number_middle,winding_number_spread,sx,&
                                &sy,sz" // new_line('a')
                        string_buff = " "
                        write(string_buff, '((F0.4,",",F0.4,",",F0.4,",",F0.10,","))') ParamArray(i, 1), &
                                ParamArray(i, 2), ParamArray(i, 3), TemperatureArray(j)
                        output_string = output_string // trim(adjustl(string_buff))
                        string_buff = " "
                        write(string_buff, '((F0.4,",",I0.4,",",F0.2,",",F0.2,",",F0.2,",",F0.2))') winding_middle, &
                                int(sk_counts(1)), winding_spread, mag(3 * slot + 1), mag(3 * slot + 2), mag(3 * slot + 3)
                        output_string = output_string // string_buff // new_line('a')
                        string_buff = " "
                        write(string_buff, '((F0.4,"_",F0.4,"_",F0.4,"_",e0.4))') ParamArray(i, 1), ParamArray(i, 2), &
                                ParamArray(i, 3), TemperatureArray(j)
                        filename_string = trim(adjustl(outputPath)) // "/data_" // trim(adjustl(string_buff)) // ".csv"
                        open(newunit=fileHandle, status="replace", action="write", file=filename_string)
                        write(fileHandle, '(A)') output_string
                        close(fileHandle)
                        ! Spins as x,y,z,Sx,Sy,Sz (src/LLG.f90:143,154).
                        rc = ising_get_spins(ctx, int(slot, c_int), spins, int(nsites, c_int64_t))
                        call ising_check(rc, ctx, "pt_gpu: spins")
                        filename_string = trim(adjustl(outputPathSpins)) // "/spins_" // trim(adjustl(string_buff)) // ".csv"
                        open(newunit=fileHandle, status="replace", action="write", file=filename_string)
                        write(fileHandle, '(A)') "x,y,z,Sx,Sy,Sz"
                        do ix = 0, numCellsX - 1
                                do iy = 0, numCellsY - 1
                                        do iz = 0, numCellsZ - 1
                                                do atom = 0, atomsPerCell - 1
                                                        site = ((ix * numCellsY + iy) * numCellsZ + iz) * atomsPerCell &
                                                               + atom + 1
                                                        write(fileHandle, '(ES22.14, 5(",", ES22.14))') &
                                                                a_bravais * (ix + 0.5_dp * atom), &
                                                                a_bravais * (iy + 0.5_dp * atom), &
                                                                a_bravais * (iz + 0.5_dp * atom), spins(site), &
                                                                spins(nsites + site), spins(2 * nsites + site)
                                                end do
                                        end do
                                end do
                        end do
                        close(fileHandle)
                        ! Charge density per plaquette as x,y,z,Winding_Density
                        ! (write_winding_number_density, src/reciprocalSpaceProcesses.f90:544-577).
                        filename_string = trim(adjustl(outputPathDensity)) // "/density_" // &
                                trim(adjustl(string_buff)) // ".csv"
                        open(newunit=fileHandle, status="replace", action="write", file=filename_string)
                        write(fileHandle, '(A)') "x,y,z,Winding_Density"
                        do iz = 1, numCellsZ
                                rc = ising_winding_density(ctx, int(slot, c_int), int(iz - 1, c_int), density)
                                call ising_check(rc, ctx, "pt_gpu: density")
                                do ix = 1, numCellsX
                                        do iy = 1, numCellsY
                                                call write_density_line(fileHandle, a_bravais * ix, a_bravais * iy, &
                                                        a_bravais * iz, density(ix + numCellsX * (iy - 1)))
                                        end do
                                end do
                        end do
                        close(fileHandle)
                        ! Interpolated charge density, every layer (write_winding_number_density_interpolated,
                        ! src/reciprocalSpaceProcesses.f90:502-540; fine indices times the cell vectors).
                        filename_string = trim(adjustl(outputPathDensityInterpolated)) // "/density_interp_" // &
                                trim(adjustl(string_buff)) // ".csv"
                        open(newunit=fileHandle, status="replace", action="write", file=filename_string)
                        write(fileHandle, '(A)') "x,y,z,Winding_Density"
                        do iz = 1, numCellsZ
                                rc = ising_winding_density_interpolated(ctx, int(slot, c_int), int(iz - 1, c_int), &
                                                                        int(scl, c_int), int(scl, c_int), density_fine)
                                call ising_check(rc, ctx, "pt_gpu: interpolated density")
                                do jx = 1, scl * numCellsX
                                        do ky = 1, scl * numCellsY
                                                call write_density_line(fileHandle, a_bravais * jx, a_bravais * ky, &
                                                        a_bravais * iz, density_fine(jx + scl * numCellsX * (ky - 1)))
                                        end do
                                end do
                        end do
                        close(fileHandle)
                end do
        end do
        call ising_destroy(ctx)
        print *, "All okay"

contains
        subroutine arg_real(n, v)
                integer, intent(in) :: n
                real(dp), intent(out) :: v
                character(len=100) :: b
                b = " "
                call get_command_argument(n, b)
                read(b, *) v
        end subroutine arg_real

        subroutine arg_int(n, v)
                integer, intent(in) :: n
                integer, intent(out) :: v
                character(len=100) :: b
                b = " "
                call get_command_argument(n, b)
                read(b, *) v
        end subroutine arg_int

        integer function env_int(name, default)
                character(len=*), intent(in) :: name
                integer, intent(in) :: default
                character(len=32) :: b
                integer :: st, ln
                call get_environment_variable(name, b, ln, st)
                env_int = default
                if (st == 0 .and. ln > 0) read(b, *) env_int
        end function env_int
end program pt_gpu
