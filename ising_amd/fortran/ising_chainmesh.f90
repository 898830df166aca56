! ising_chainmesh.f90 -- drop-in ChainMesh module for simple-cubic, bcc and
! hyper-kagome lattices.
!
! Keeps the fields of the reference's ChainMesh_t that the Monte Carlo path
! and the PT driver read (src/chainMesh.f90:17-47): numCellsX/Y/Z, numAtoms
! and atomSpins(numAtoms,3) real(8) in the reference's site order
! (IndexFromCoordinatesInteger, src/chainMesh.f90:673-685: z fastest; with n
! atoms per cell, atom a of cell c is site n*(c-1) + a, src/chainMesh.f90:878).
! Each mesh owns one device context (one replica) created on first use;
! the spins stay host-owned and are copied in and out per call, exactly as
! the reference's routines mutate chainMesh%atomSpins in place.
module ChainMesh
        use iso_c_binding
        use ising_gpu_c
        implicit none

        type ChainMesh_t
                integer :: numCellsX = 0, numCellsY = 0, numCellsZ = 0
                integer :: numAtoms = 0
                logical :: periodic(3) = [.true., .true., .true.]
                integer :: cell = 0           ! ISING_CELL_SC, ISING_CELL_BCC or ISING_CELL_HYPERKAGOME
                integer :: atomsPerCell = 1
                real(kind=8) :: a = 1.0_8     ! lattice constant (positions in output files)
                ! upsampling of the interpolated charge density (src/chainMesh.f90:44,764-773)
                integer :: sclx = 10, scly = 10
                real(kind=8), allocatable :: atomSpins(:,:)
                ! the last demagnetising field computed (src/chainMesh.f90:38),
                ! filled by the demag = .true. paths of EnergyMin
                real(kind=8), allocatable :: demagnetisation_array(:,:)
                type(c_ptr) :: gpu = c_null_ptr
                integer(c_int64_t) :: seed = 0
        end type ChainMesh_t

        ! Base seed of the per-mesh Philox streams; each new mesh context
        ! takes the next value so that meshes never share a stream.
        integer(c_int64_t), save :: chainmesh_seed_base = 20261203_c_int64_t
        integer(c_int64_t), save :: chainmesh_count = 0

        public :: ChainMesh_t, makeChainMesh, deallocateChainMesh, IndexFromCoordinatesInteger, &
                  chainmesh_context, chainmesh_push_spins, chainmesh_pull_spins, chainmesh_set_seed_base

contains

        ! Mesh of numCellsX x numCellsY x numCellsZ cubic cells, one atom per
        ! cell (SC), two (cell = ISING_CELL_BCC: atoms at (0,0,0) and
        ! (1/2,1/2,1/2), src/unit_cells.f90:7-30) or twelve (cell =
        ! ISING_CELL_HYPERKAGOME, src/unit_cells.f90:32-77). The reference
        ! builds general Bravais cells (src/chainMesh.f90:729-911); these three
        ! are provided. `a` is the cell edge used for positions in output files;
        ! by default the edge its unit-cell routine sets (bcc 2, hyper-kagome 6,
        ! src/unit_cells.f90:21,47), 1 for simple cubic.
        function makeChainMesh(numCellsX, numCellsY, numCellsZ, periodic, cell, a, sclx, scly) result(mesh)
                integer, intent(in) :: numCellsX, numCellsY, numCellsZ
                logical, intent(in), optional :: periodic(3)
                integer, intent(in), optional :: cell
                real(kind=8), intent(in), optional :: a
                integer, intent(in), optional :: sclx, scly
                type(ChainMesh_t) :: mesh
                integer :: stat
                mesh%numCellsX = numCellsX
                mesh%numCellsY = numCellsY
                mesh%numCellsZ = numCellsZ
                if (present(cell)) mesh%cell = cell
                select case (mesh%cell)
                case (ISING_CELL_BCC)
                        mesh%atomsPerCell = 2
                        mesh%a = 2.0_8
                case (ISING_CELL_HYPERKAGOME)
                        mesh%atomsPerCell = 12
                        mesh%a = 6.0_8
                case default
                        mesh%atomsPerCell = 1
                        mesh%a = 1.0_8
                end select
                if (present(a)) mesh%a = a
                if (present(sclx)) mesh%sclx = sclx
                if (present(scly)) mesh%scly = scly
                if (mesh%sclx < 2) error stop "x interpolation scale must be greater than or equal to 2"
                if (mesh%scly < 2) error stop "y interpolation scale must be greater than or equal to 2"
                mesh%numAtoms = numCellsX * numCellsY * numCellsZ * mesh%atomsPerCell
                if (present(periodic)) mesh%periodic = periodic
                allocate(mesh%atomSpins(mesh%numAtoms, 3), stat=stat)
                if (stat /= 0) error stop "Error: Failed to allocated atomSpins array"
                mesh%atomSpins = 0.0_8
                mesh%atomSpins(:, 3) = 1.0_8
                allocate(mesh%demagnetisation_array(mesh%numAtoms, 3), stat=stat)
                if (stat /= 0) error stop "Error: Failed to allocate demagnetisation_array"
                mesh%demagnetisation_array = 0.0_8
        end function makeChainMesh

        subroutine deallocateChainMesh(mesh)
                type(ChainMesh_t), intent(inout) :: mesh
                if (c_associated(mesh%gpu)) call ising_destroy(mesh%gpu)
                mesh%gpu = c_null_ptr
                if (allocated(mesh%atomSpins)) deallocate(mesh%atomSpins)
                if (allocated(mesh%demagnetisation_array)) deallocate(mesh%demagnetisation_array)
        end subroutine deallocateChainMesh

        ! 1-based cell index, z fastest (src/chainMesh.f90:673-685).
        pure function IndexFromCoordinatesInteger(mesh, i, j, k) result(res)
                type(ChainMesh_t), intent(in) :: mesh
                integer, intent(in) :: i, j, k
                integer :: res
                res = (i - 1) * mesh%numCellsY * mesh%numCellsZ + (j - 1) * mesh%numCellsZ + (k - 1) + 1
        end function IndexFromCoordinatesInteger

        subroutine chainmesh_set_seed_base(seed)
                integer(c_int64_t), intent(in) :: seed
                chainmesh_seed_base = seed
        end subroutine chainmesh_set_seed_base

        ! The mesh's device context (one ladder of one replica), created on
        ! first use with a Hamiltonian and beta that every caller overwrites.
        subroutine chainmesh_context(mesh)
                type(ChainMesh_t), intent(inout) :: mesh
                type(ising_ladder_desc) :: d(1)
                real(c_double), target, save :: beta_one(1) = [1.0_c_double]
                integer(c_int) :: rc
                integer :: a
                if (c_associated(mesh%gpu)) return
                d(1)%lattice%lx = mesh%numCellsX
                d(1)%lattice%ly = mesh%numCellsY
                d(1)%lattice%lz = mesh%numCellsZ
                d(1)%lattice%cell = int(mesh%cell, c_int32_t)
                do a = 1, 3
                        d(1)%lattice%periodic(a) = merge(1_c_int32_t, 0_c_int32_t, mesh%periodic(a))
                end do
                d(1)%ham = ising_hamiltonian(0.0_c_double, 0.0_c_double, ISING_DMI_NONE, 0_c_int32_t, &
                                             0.0_c_double, 0.0_c_double, 1.0_c_double)
                d(1)%n_replicas = 1
                d(1)%reserved = 0
                d(1)%cone_radius = 0.5_c_double
                d(1)%betas = c_loc(beta_one)
                mesh%seed = chainmesh_seed_base + chainmesh_count
                chainmesh_count = chainmesh_count + 1
                rc = ising_create(mesh%gpu, d, 1_c_int, mesh%seed, 0_c_int)
                call ising_check(rc, c_null_ptr, "makeChainMesh device context")
        end subroutine chainmesh_context

        subroutine chainmesh_push_spins(mesh)
                type(ChainMesh_t), intent(inout) :: mesh
                integer(c_int) :: rc
                call chainmesh_context(mesh)
                rc = ising_set_spins(mesh%gpu, 0_c_int, mesh%atomSpins, int(mesh%numAtoms, c_int64_t))
                call ising_check(rc, mesh%gpu, "copy-in of atomSpins")
        end subroutine chainmesh_push_spins

        subroutine chainmesh_pull_spins(mesh)
                type(ChainMesh_t), intent(inout) :: mesh
                integer(c_int) :: rc
                rc = ising_get_spins(mesh%gpu, 0_c_int, mesh%atomSpins, int(mesh%numAtoms, c_int64_t))
                call ising_check(rc, mesh%gpu, "copy-out of atomSpins")
        end subroutine chainmesh_pull_spins
end module ChainMesh
