! test_dropin.f90 -- exercises the drop-in Fortran entry points the way the
! reference's PT driver calls them (src/PT.f90:229-230, 251-254, 316) and
! prints one "name value" line per check; tests/test_fortran_dropin.py
! compares the values with the FP64 oracle and closed forms. With a directory
! argument it also writes the demagnetisation check's spins and fields there.
program test_dropin
        use iso_c_binding
        use omp_lib, only: OMP_LOCK_KIND
        use ChainMesh
        use EnergyMin
        use PT_Utils
        use reciprocal_space_processes
        implicit none
        type(ChainMesh_t) :: mesh, fm, dm, hk
        integer(kind=OMP_LOCK_KIND), allocatable :: lockArray(:)
        real(kind=8) :: e0, e1, eold, enew, winding_middle, winding_spread, m(3), sp(3), s0(3)
        real(kind=8) :: J, Dz, B, beta, pi
        integer :: skyr, i, ix, iy, iz, k
        real(kind=8) :: q, th
        real(kind=8), allocatable :: density(:,:), hfield(:,:)
        integer :: peak(2)
        character(len=512) :: outdir
        integer :: unit, nargs

        pi = acos(-1.0_8)
        J = -1.0_8; Dz = 0.6_8; B = 0.3_8; beta = 2.0_8
        mesh = makeChainMesh(8, 8, 4)
        allocate(lockArray(mesh%numAtoms))
        lockArray = 0
        ! Neel spiral along x in reference units: S = (sin qx, 0, cos qx).
        q = 2.0_8 * pi / 8.0_8
        do ix = 0, 7
                do iy = 0, 7
                        do iz = 0, 3
                                k = (ix * 8 + iy) * 4 + iz + 1
                                mesh%atomSpins(k, :) = [sin(q * ix), 0.0_8, cos(q * ix)]
                        end do
                end do
        end do
        call TotalHeisenbergEnergy(mesh, J, 0.0_8, Dz, 0.0_8, B, lockArray, e0, .false.)
        print '(A,ES24.16)', "spiral_energy ", e0
        print '(A,ES24.16)', "spiral_expect ", dble(mesh%numAtoms) * (J * (2.0_8 + cos(q)) - Dz * sin(q))

        ! Single-site energies against the total-energy difference.
        sp = [0.6_8, -0.64_8, 0.48_8]
        i = 37
        call calculateHeisenbergEnergy(mesh, i, J, 0.0_8, Dz, 0.0_8, B, lockArray, sp, eold, enew, &
                                       reshape([0.0_8], [1, 1]), .false.)
        s0 = mesh%atomSpins(i, :)
        mesh%atomSpins(i, :) = sp
        ! The single-site form uses Zeeman coefficient 1 and the total 2, as the
        ! reference: E_total(new) - E_total(old) = (enew - eold) - B (Sz' - Sz).
        call TotalHeisenbergEnergy(mesh, J, 0.0_8, Dz, 0.0_8, B, lockArray, e1, .false.)
        print '(A,ES24.16)', "local_delta ", enew - eold
        print '(A,ES24.16)', "total_delta ", (e1 - e0) + B * (sp(3) - s0(3))
        mesh%atomSpins(i, :) = s0

        ! Sweeps change the configuration and keep unit spins.
        call Metropolis_mcs(mesh, beta, 20, J, 0.0_8, Dz, 0.0_8, B, 0.5_8, lockArray, demag=.false.)
        print '(A,ES24.16)', "max_norm_dev ", maxval(abs(sqrt(sum(mesh%atomSpins**2, dim=2)) - 1.0_8))
        call TotalHeisenbergEnergy(mesh, J, 0.0_8, Dz, 0.0_8, B, lockArray, e1, .false.)
        print '(A,ES24.16)', "energy_after ", e1

        ! Statistics of a ferromagnet and of a skyrmion texture.
        fm = makeChainMesh(16, 16, 4)
        call chainMesh_statistics(fm, skyr, winding_middle, winding_spread, m)
        print '(A,3ES24.16)', "fm_mag ", m
        print '(A,ES24.16)', "fm_winding ", winding_middle
        do ix = 0, 15
                do iy = 0, 15
                        th = pi * max(0.0_8, 1.0_8 - sqrt((ix - 7.7_8)**2 + (iy - 7.8_8)**2) / 5.5_8)
                        do iz = 0, 3
                                k = (ix * 16 + iy) * 4 + iz + 1
                                fm%atomSpins(k, :) = [sin(th) * cos(atan2(iy - 7.8_8, ix - 7.7_8)), &
                                                      sin(th) * sin(atan2(iy - 7.8_8, ix - 7.7_8)), cos(th)]
                        end do
                end do
        end do
        call chainMesh_statistics(fm, skyr, winding_middle, winding_spread, m)
        print '(A,ES24.16)', "sk_winding_middle ", winding_middle
        print '(A,I0)', "sk_number ", skyr
        print '(A,ES24.16)', "sk_layer1 ", calculate_winding_number2(fm, 1)
        call calculate_winding_number_density(fm, 2, density)
        print '(A,ES24.16)', "sk_density_sum ", sum(density)
        print '(A,ES24.16)', "sk_density_max ", maxval(abs(density))
        peak = maxloc(abs(density))
        print '(A,I0,A,I0)', "sk_density_peak ", peak(1), " ", peak(2)
        ! Demagnetisation: bcc mesh 4 x 3 x 3 (periodic), a conical texture.
        nargs = command_argument_count()
        dm = makeChainMesh(4, 3, 3, cell=ISING_CELL_BCC)
        do i = 1, dm%numAtoms
                th = 0.37_8 * i
                dm%atomSpins(i, :) = [0.6_8 * cos(th), 0.6_8 * sin(th), 0.8_8]
        end do
        call calculate_demagnetisation_field(dm, hfield)
        call TotalHeisenbergEnergy(dm, J, 0.0_8, Dz, 0.0_8, B, lockArray, e0, .true.)
        call TotalHeisenbergEnergy(dm, J, 0.0_8, Dz, 0.0_8, B, lockArray, e1, .false.)
        print '(A,ES24.16)', "demag_energy_term ", e0 - e1
        print '(A,ES24.16)', "demag_array_vs_field ", maxval(abs(dm%demagnetisation_array - hfield))
        ! calculateHeisenbergEnergy takes the field as an argument; one large
        ! enough for mu_0 H to show at the FP32 resolution of the local field
        sp = [0.0_8, 0.6_8, -0.8_8]
        allocate(density(dm%numAtoms, 3))
        density(:, 1) = 2.0e5_8
        density(:, 2) = -1.0e5_8
        density(:, 3) = 3.0e5_8
        call calculateHeisenbergEnergy(dm, 5, J, 0.0_8, Dz, 0.0_8, B, lockArray, sp, eold, enew, density, .true.)
        print '(A,2ES24.16)', "demag_local ", eold, enew
        call calculateHeisenbergEnergy(dm, 5, J, 0.0_8, Dz, 0.0_8, B, lockArray, sp, e0, e1, density, .false.)
        print '(A,2ES24.16)', "nodemag_local ", e0, e1
        deallocate(density)
        if (nargs >= 1) then
                call get_command_argument(1, outdir)
                open(newunit=unit, file=trim(outdir)//"/demag_in.txt", status="replace", action="write")
                do i = 1, dm%numAtoms
                        write(unit, '(3ES24.16)') dm%atomSpins(i, :)
                end do
                close(unit)
                open(newunit=unit, file=trim(outdir)//"/demag_field.txt", status="replace", action="write")
                do i = 1, dm%numAtoms
                        write(unit, '(3ES24.16)') hfield(i, :)
                end do
                close(unit)
        end if
        ! r = 0.5: the field is recomputed after sweep int(pi sqrt 5) = 7, the last one
        call Metropolis_mcs(dm, beta, 7, J, 0.0_8, Dz, 0.0_8, B, 0.5_8, lockArray, demag=.true.)
        call calculate_demagnetisation_field(dm, hfield)
        print '(A,ES24.16)', "demag_after_sweeps ", maxval(abs(dm%demagnetisation_array - hfield))
        print '(A,ES24.16)', "demag_norm_dev ", maxval(abs(sqrt(sum(dm%atomSpins**2, dim=2)) - 1.0_8))
        ! Hyper-kagome mesh 3 x 3 x 2 (the lattice of src/ISING.f90:105): a
        ! ferromagnet along z has 2 bonds per atom, E = (2 J - 2 B) numAtoms in
        ! the reference's convention; its layer charge is zero; sweeps keep
        ! unit spins.
        hk = makeChainMesh(3, 3, 2, cell=ISING_CELL_HYPERKAGOME)
        hk%atomSpins(:, 1) = 0.0_8
        hk%atomSpins(:, 2) = 0.0_8
        hk%atomSpins(:, 3) = 1.0_8
        call TotalHeisenbergEnergy(hk, J, 0.0_8, Dz, 0.0_8, B, lockArray, e0, .false.)
        print '(A,2ES24.16)', "hk_fm_energy ", e0, (2.0_8 * J - 2.0_8 * B) * hk%numAtoms
        print '(A,ES24.16)', "hk_fm_winding ", calculate_winding_number2(hk, 1)
        call Metropolis_mcs(hk, beta, 5, J, 0.0_8, Dz, 0.0_8, B, 0.5_8, lockArray, demag=.false.)
        print '(A,ES24.16)', "hk_norm_dev ", maxval(abs(sqrt(sum(hk%atomSpins**2, dim=2)) - 1.0_8))
        print '(A,I0)', "hk_atoms ", hk%numAtoms
        call deallocateChainMesh(mesh)
        call deallocateChainMesh(fm)
        call deallocateChainMesh(dm)
        call deallocateChainMesh(hk)
        print '(A)', "done"
end program test_dropin
