! pt_utils_gpu.f90 -- drop-in reciprocal_space_processes (winding number) and
! PT_Utils (slot maps, final statistics) modules on the device path.
module reciprocal_space_processes
        use iso_c_binding
        use ising_gpu_c
        use ChainMesh
        implicit none
contains
        ! Replaces calculate_winding_number2 (src/reciprocalSpaceProcesses.f90:320-364):
        ! Berg-Luescher charge of layer Z_index (1-based) of the current spins,
        ! in FP64 on the device (the integer value is exact, see ising_winding).
        function calculate_winding_number2(chainMesh, Z_index) result(winding_number)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(in) :: Z_index
                real(kind=8) :: winding_number
                integer(c_int64_t) :: q
                real(c_double) :: qr
                integer(c_int) :: rc
                if (Z_index < 1 .or. Z_index > chainMesh%numCellsZ) error stop "Z_index out of bounds"
                call chainmesh_push_spins(chainMesh)
                rc = ising_winding(chainMesh%gpu, 0_c_int, int(Z_index - 1, c_int), q, qr)
                call ising_check(rc, chainMesh%gpu, "calculate_winding_number2")
                winding_number = qr
        end function calculate_winding_number2

        ! Replaces calculate_winding_number_density (src/reciprocalSpaceProcesses.f90:438-500):
        ! density_matrix(i, j) = (sigma1 + sigma2) / (4 pi) of the plaquette with
        ! corner cell (i, j) of layer Z_index (1-based), from the device.
        subroutine calculate_winding_number_density(chainMesh, Z_index, density_matrix)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(in) :: Z_index
                real(kind=8), dimension(:,:), allocatable, intent(out) :: density_matrix
                integer(c_int) :: rc
                if (Z_index < 1 .or. Z_index > chainMesh%numCellsZ) error stop "Z_index out of bounds"
                allocate(density_matrix(chainMesh%numCellsX, chainMesh%numCellsY))
                call chainmesh_push_spins(chainMesh)
                rc = ising_winding_density(chainMesh%gpu, 0_c_int, int(Z_index - 1, c_int), density_matrix)
                call ising_check(rc, chainMesh%gpu, "calculate_winding_number_density")
        end subroutine calculate_winding_number_density

        ! Replaces write_winding_number_density (src/reciprocalSpaceProcesses.f90:544-577):
        ! header "x,y,z,Winding_Density", then for every layer i, cell j along x and
        ! k along y the position (j, k, i) * a (1-based cell indices, lattice
        ! constant a of the mesh) and the density, each F0.8.
        subroutine write_winding_number_density(chainMesh, filepath)
                type(ChainMesh_t), intent(inout) :: chainMesh
                character(len=*), intent(in) :: filepath
                real(kind=8), dimension(:,:), allocatable :: density_matrix
                integer :: i, j, k, unit, stat
                open(newunit=unit, file=filepath, status="replace", action="write", iostat=stat)
                if (stat /= 0) error stop "Error opening file to write density information"
                write(unit, '(A)') "x,y,z,Winding_Density"
                do i = 1, chainMesh%numCellsZ
                        call calculate_winding_number_density(chainMesh, i, density_matrix)
                        do j = 1, chainMesh%numCellsX
                                do k = 1, chainMesh%numCellsY
                                        call write_density_line(unit, chainMesh%a * dble(j), chainMesh%a * dble(k), &
                                                                chainMesh%a * dble(i), density_matrix(j, k))
                                end do
                        end do
                end do
                close(unit)
        end subroutine write_winding_number_density

        ! Replaces calculate_winding_density_interpolated (src/reciprocalSpaceProcesses.f90:365-431):
        ! the charge density of layer z_index (1-based) on the grid sclx times finer
        ! in x and scly in y, density(numX, numY), from band-limited interpolation of
        ! the corner spins (computed on the device).
        subroutine calculate_winding_density_interpolated(chainMesh, z_index, density)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(in) :: z_index
                real(kind=8), dimension(:,:), allocatable, intent(inout) :: density
                integer(c_int) :: rc
                integer :: numX, numY
                if (z_index < 1 .or. z_index > chainMesh%numCellsZ) error stop "Z_index out of bounds"
                numX = chainMesh%sclx * chainMesh%numCellsX
                numY = chainMesh%scly * chainMesh%numCellsY
                if (allocated(density)) then
                        if (any(shape(density) /= [numX, numY])) deallocate(density)
                end if
                if (.not. allocated(density)) allocate(density(numX, numY))
                call chainmesh_push_spins(chainMesh)
                rc = ising_winding_density_interpolated(chainMesh%gpu, 0_c_int, int(z_index - 1, c_int), &
                                                        int(chainMesh%sclx, c_int), int(chainMesh%scly, c_int), density)
                call ising_check(rc, chainMesh%gpu, "calculate_winding_density_interpolated")
        end subroutine calculate_winding_density_interpolated

        ! Replaces write_winding_number_density_interpolated
        ! (src/reciprocalSpaceProcesses.f90:502-540): header "x,y,z,Winding_Density",
        ! then for every layer i, fine point j along x and k along y the position
        ! (j, k, i) * a -- the reference writes fine indices times the cell
        ! vectors -- and the density, each F0.8.
        subroutine write_winding_number_density_interpolated(chainMesh, filepath)
                type(ChainMesh_t), intent(inout) :: chainMesh
                character(len=*), intent(in) :: filepath
                real(kind=8), dimension(:,:), allocatable :: density_matrix
                integer :: i, j, k, unit, stat
                open(newunit=unit, file=filepath, status="replace", action="write", iostat=stat)
                if (stat /= 0) error stop "Error opening file to write density information"
                write(unit, '(A)') "x,y,z,Winding_Density"
                do i = 1, chainMesh%numCellsZ
                        call calculate_winding_density_interpolated(chainMesh, i, density_matrix)
                        do j = 1, size(density_matrix, 1)
                                do k = 1, size(density_matrix, 2)
                                        call write_density_line(unit, chainMesh%a * dble(j), chainMesh%a * dble(k), &
                                                                chainMesh%a * dble(i), density_matrix(j, k))
                                end do
                        end do
                end do
                close(unit)
        end subroutine write_winding_number_density_interpolated

        ! Replaces compute_skyrmion_distribution (src/reciprocalSpaceProcesses.f90:930-983):
        ! winding_array(j) = number of regions of charge j (signed) of layer
        ! Z_index (1-based), searched over num_thresholds mask thresholds from
        ! max_threshold down to min_threshold and 20 filter widths, on the device.
        subroutine compute_skyrmion_distribution(chainMesh, N, winding_array, min_threshold, max_threshold, &
                                                 num_thresholds, Z_index)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(in) :: N
                integer, dimension(:), allocatable, intent(inout) :: winding_array
                real(kind=8), intent(in) :: min_threshold, max_threshold
                integer, intent(in) :: num_thresholds, Z_index
                type(ising_skyrmion_params) :: prm
                integer(c_int32_t), allocatable :: counts(:)
                integer(c_int) :: rc
                if (max_threshold < min_threshold) error stop "max_threshold must be greater than min_threshold"
                if (Z_index < 1 .or. Z_index > chainMesh%numCellsZ) error stop "Z_index out of bounds"
                if (allocated(winding_array)) then
                        if (size(winding_array) /= N) deallocate(winding_array)
                end if
                if (.not. allocated(winding_array)) allocate(winding_array(N))
                allocate(counts(N))
                prm%scale_x = chainMesh%sclx
                prm%scale_y = chainMesh%scly
                prm%n_particles = N
                prm%num_thresholds = num_thresholds
                prm%num_sigmas = 20
                prm%min_threshold = min_threshold
                prm%max_threshold = max_threshold
                call chainmesh_push_spins(chainMesh)
                rc = ising_skyrmion_distribution(chainMesh%gpu, 0_c_int, int(Z_index - 1, c_int), prm, counts, c_null_ptr)
                call ising_check(rc, chainMesh%gpu, "compute_skyrmion_distribution")
                winding_array = counts
        end subroutine compute_skyrmion_distribution

        ! Replaces calculate_demagnetisation_field (src/reciprocalSpaceProcesses.f90:202-281):
        ! the demagnetising field of the current spins, computed on the device
        ! (cloud-in-cell deposit, discrete Fourier transform, -(k.M) k / k^2,
        ! inverse, weights back to the atoms), outputArray(numAtoms, 3).
        subroutine calculate_demagnetisation_field(chainMesh, outputArray)
                type(ChainMesh_t), intent(inout) :: chainMesh
                real(kind=8), allocatable, dimension(:,:), intent(out) :: outputArray
                integer(c_int) :: rc
                allocate(outputArray(chainMesh%numAtoms, 3))
                call chainmesh_push_spins(chainMesh)
                rc = ising_demag_field(chainMesh%gpu, 0_c_int, outputArray, int(chainMesh%numAtoms, c_int64_t))
                call ising_check(rc, chainMesh%gpu, "calculate_demagnetisation_field")
        end subroutine calculate_demagnetisation_field

        subroutine write_density_line(unit, x, y, z, d)
                integer, intent(in) :: unit
                real(kind=8), intent(in) :: x, y, z, d
                character(len=160) :: buf
                write(buf, '(F0.8, A, F0.8, A, F0.8, A, F0.8)') x, ",", y, ",", z, ",", d
                write(unit, '(A)') trim(adjustl(buf))
        end subroutine write_density_line
end module reciprocal_space_processes

module PT_Utils
        use iso_c_binding
        use ising_gpu_c
        use ChainMesh
        use reciprocal_space_processes
        implicit none
        public :: indicesFromSlot, SlotFromIndices, chainMesh_statistics
contains
        ! Slot <-> (J, D, B) index maps, 1-based (src/PT-Utils.f90:11-30).
        subroutine indicesFromSlot(Slot, NJ, ND, NB, Jindex, Dindex, BIndex)
                integer, intent(in) :: Slot, NJ, ND, NB
                integer, intent(out) :: Jindex, Dindex, Bindex
                integer :: res
                Jindex = (Slot - 1) / (ND * NB) + 1
                res = mod(Slot - 1, ND * NB)
                Dindex = res / NB + 1
                Bindex = mod(res, NB) + 1
        end subroutine indicesFromSlot

        function SlotFromIndices(NJ, ND, NB, JIndex, DIndex, BIndex) result(res)
                integer, intent(in) :: NJ, ND, NB, JIndex, DIndex, BIndex
                integer :: res
                res = (JIndex - 1) * ND * NB + (DIndex - 1) * NB + BIndex
        end function SlotFromIndices

        ! Replaces chainMesh_statistics (src/PT-Utils.f90:33-66): winding number
        ! of every layer (one device pass), spread = max - min, the skyrmion
        ! count of the middle layer numCellsZ/2 (compute_skyrmion_distribution
        ! with N = 1, thresholds 1e-5..0.95 in 60 steps), and the mean spin vector.
        subroutine chainMesh_statistics(chainMesh, skyrmion_number_middle, winding_number_middle, &
                                        winding_number_spread, magnetisation)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(out) :: skyrmion_number_middle
                real(kind=8), intent(out) :: winding_number_middle, winding_number_spread
                real(kind=8), intent(out) :: magnetisation(3)
                integer(c_int64_t), allocatable :: q(:)
                real(c_double), allocatable :: qr(:)
                real(c_double) :: m(3)
                integer(c_int) :: rc
                integer :: mid
                integer, allocatable :: skyrmion_array(:)
                real(kind=8), parameter :: lower_bound = 1e-5_8, upper_bound = 0.95_8
                integer, parameter :: num_thresholds = 60
                call chainmesh_push_spins(chainMesh)
                allocate(q(chainMesh%numCellsZ), qr(chainMesh%numCellsZ))
                rc = ising_winding_all(chainMesh%gpu, q, qr)
                call ising_check(rc, chainMesh%gpu, "chainMesh_statistics: winding")
                winding_number_spread = abs(maxval(qr) - minval(qr))
                mid = max(1, chainMesh%numCellsZ / 2)
                winding_number_middle = qr(mid)
                call compute_skyrmion_distribution(chainMesh, 1, skyrmion_array, lower_bound, upper_bound, &
                                                   num_thresholds, mid)
                skyrmion_number_middle = skyrmion_array(1)
                rc = ising_magnetisation(chainMesh%gpu, m)
                call ising_check(rc, chainMesh%gpu, "chainMesh_statistics: magnetisation")
                magnetisation = m
        end subroutine chainMesh_statistics
end module PT_Utils
