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
        ! of every layer (one device pass), spread = max - min, the middle
        ! layer numCellsZ/2, and the mean spin vector. The flood-fill
        ! skyrmion count of the reference (compute_skyrmion_distribution) is
        ! not computed here: skyrmion_number_middle returns nint(|Q_middle|).
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
                call chainmesh_push_spins(chainMesh)
                allocate(q(chainMesh%numCellsZ), qr(chainMesh%numCellsZ))
                rc = ising_winding_all(chainMesh%gpu, q, qr)
                call ising_check(rc, chainMesh%gpu, "chainMesh_statistics: winding")
                winding_number_spread = abs(maxval(qr) - minval(qr))
                mid = max(1, chainMesh%numCellsZ / 2)
                winding_number_middle = qr(mid)
                skyrmion_number_middle = int(abs(q(mid)))
                rc = ising_magnetisation(chainMesh%gpu, m)
                call ising_check(rc, chainMesh%gpu, "chainMesh_statistics: magnetisation")
                magnetisation = m
        end subroutine chainMesh_statistics
end module PT_Utils
