! test_dropin.f90 -- exercises the drop-in Fortran entry points the way the
! reference's PT driver calls them (src/PT.f90:229-230, 251-254, 316) and
! prints one "name value" line per check; tests/test_fortran_dropin.py
! compares the values with the FP64 oracle and closed forms.
program test_dropin
        use iso_c_binding
        use omp_lib, only: OMP_LOCK_KIND
        use ChainMesh
        use EnergyMin
        use PT_Utils
        use reciprocal_space_processes
        implicit none
        type(ChainMesh_t) :: mesh, fm
        integer(kind=OMP_LOCK_KIND), allocatable :: lockArray(:)
        real(kind=8) :: e0, e1, eold, enew, winding_middle, winding_spread, m(3), sp(3), s0(3)
        real(kind=8) :: J, Dz, B, beta, pi
        integer :: skyr, i, ix, iy, iz, k
        real(kind=8) :: q, th
        real(kind=8), allocatable :: density(:,:)
        integer :: peak(2)

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
                        th = pi * exp(-((ix - 7.7_8)**2 + (iy - 7.8_8)**2) / 9.0_8)
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
        call deallocateChainMesh(mesh)
        call deallocateChainMesh(fm)
        print '(A)', "done"
end program test_dropin
