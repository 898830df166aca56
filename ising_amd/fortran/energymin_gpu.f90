! energymin_gpu.f90 -- drop-in EnergyMin module: the reference's Metropolis
! entry points with their argument lists, running on the MI355X through the
! C ABI (include/ising_gpu.h).
!
! Parameter convention (the reference's, src/energyMin.f90):
!   exchange +J S.S per bond (:139), so the library gets J_lib = -J;
!   interfacial DMI D_ij = Dz (c x r_{j->i}) (:114,137);
!   Zeeman -B S_z in single-site energies (:161-162) and -2 B S_z in
!   TotalHeisenbergEnergy (g * Bohr_magneton = 2, :222).
! J_prime (second shell) must be 0, as the reference's PT always calls it
! (src/PT.f90:229-230, 251-254); anything else stops with a message, as the
! reference stops on unsupported states. demag = .true. adds the
! demagnetisation term -mu_0 S.H (src/energyMin.f90:113-116,163-166,223-226)
! with the field of calculate_demagnetisation_field computed on the device;
! it needs a fully periodic mesh. With demag = .false. the field is not
! computed (the reference computes it on every call and discards it).
!
! Sampling: Metropolis_mcs runs exact colour-by-colour sweeps (two colours
! on simple-cubic and bcc meshes, three on hyper-kagome; every site one
! proposal per sweep, accept iff u < exp(-beta dE)) with the
! reference's ball proposal of radius r, instead of the reference's
! snapshot scheme with repeated picks (src/energyMin.f90:287-360).
module EnergyMin
        use iso_c_binding
        use omp_lib, only: OMP_LOCK_KIND
        use ising_gpu_c
        use ChainMesh
        implicit none

contains

        function reference_hamiltonian(J, Dz, B, zeeman_total_scale, demag) result(h)
                real(kind=8), intent(in) :: J, Dz, B, zeeman_total_scale
                logical, intent(in), optional :: demag
                type(ising_hamiltonian) :: h
                h%J = -J
                h%D = Dz
                h%dmi_type = ISING_DMI_INTERFACIAL_REF
                h%demag = 0
                if (present(demag)) h%demag = merge(1_c_int32_t, 0_c_int32_t, demag)
                h%Bz = B
                h%Kz = 0.0_8
                h%zeeman_total_scale = zeeman_total_scale
        end function reference_hamiltonian

        subroutine check_supported(J_prime)
                real(kind=8), intent(in) :: J_prime
                if (J_prime /= 0.0_8) error stop "EnergyMin (GPU): J_prime /= 0 (second-shell exchange) is not supported"
        end subroutine check_supported

        ! demag argument of the reference's routines: optional, default .true.
        ! (src/energyMin.f90:105-106, 276-277).
        logical function demag_on(demag)
                logical, intent(in), optional :: demag
                demag_on = .true.
                if (present(demag)) demag_on = demag
        end function demag_on

        ! chainMesh%demagnetisation_array <- the field the device holds now.
        subroutine pull_demag_field(chainMesh)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer(c_int) :: rc
                if (.not. allocated(chainMesh%demagnetisation_array)) &
                        allocate(chainMesh%demagnetisation_array(chainMesh%numAtoms, 3))
                rc = ising_get_demag_field(chainMesh%gpu, 0_c_int, chainMesh%demagnetisation_array, &
                                           int(chainMesh%numAtoms, c_int64_t))
                call ising_check(rc, chainMesh%gpu, "demagnetisation_array")
        end subroutine pull_demag_field

        ! Replaces Metropolis_mcs (src/energyMin.f90:234-374).
        subroutine Metropolis_mcs(chainMesh, beta, numMCSSweeps, J, J_prime, Dz, Dz_prime, B, r, lockArray, &
                                  demag)
                type(ChainMesh_t), intent(inout) :: chainMesh
                real(kind=8), intent(in) :: beta, J, J_prime, Dz, Dz_prime, B, r
                integer, intent(in) :: numMCSSweeps
                integer(kind=OMP_LOCK_KIND), intent(inout) :: lockArray(:)
                logical, optional, intent(in) :: demag
                integer(c_int) :: rc
                logical :: dm
                real(kind=8), parameter :: pi = 3.14159265358979323846_8
                call check_supported(J_prime)
                dm = demag_on(demag)
                call chainmesh_push_spins(chainMesh)
                rc = ising_set_hamiltonian(chainMesh%gpu, 0_c_int, reference_hamiltonian(J, Dz, B, 2.0_8, dm))
                call ising_check(rc, chainMesh%gpu, "Metropolis_mcs: Hamiltonian")
                if (dm) then
                        ! demag_update_interval = int(pi sqrt 5 / (2 r)) (src/energyMin.f90:258)
                        rc = ising_set_demag_interval(chainMesh%gpu, int((pi * sqrt(5.0_8)) / (2.0_8 * r), c_int))
                        call ising_check(rc, chainMesh%gpu, "Metropolis_mcs: demagnetisation interval")
                end if
                rc = ising_set_cone(chainMesh%gpu, 0_c_int, r)
                call ising_check(rc, chainMesh%gpu, "Metropolis_mcs: cone radius")
                rc = ising_set_beta(chainMesh%gpu, 0_c_int, beta)
                call ising_check(rc, chainMesh%gpu, "Metropolis_mcs: beta")
                rc = ising_sweep(chainMesh%gpu, int(numMCSSweeps, c_int))
                call ising_check(rc, chainMesh%gpu, "Metropolis_mcs: sweeps")
                call chainmesh_pull_spins(chainMesh)
                if (dm) call pull_demag_field(chainMesh)
        end subroutine Metropolis_mcs

        ! Replaces TotalHeisenbergEnergy (src/energyMin.f90:499-528).
        subroutine TotalHeisenbergEnergy(chainMesh, J, J_prime, Dz, Dz_prime, B, lockArray, totalEnergy, demag)
                type(ChainMesh_t), intent(inout) :: chainMesh
                real(kind=8), intent(in) :: J, J_prime, Dz, Dz_prime, B
                integer(kind=OMP_LOCK_KIND), intent(inout) :: lockArray(:)
                real(kind=8), intent(out) :: totalEnergy
                logical, intent(in) :: demag
                integer(c_int) :: rc
                real(c_double) :: e
                call check_supported(J_prime)
                call chainmesh_push_spins(chainMesh)
                rc = ising_energy_with(chainMesh%gpu, 0_c_int, reference_hamiltonian(J, Dz, B, 2.0_8, demag), e)
                call ising_check(rc, chainMesh%gpu, "TotalHeisenbergEnergy")
                totalEnergy = e
                if (demag) call pull_demag_field(chainMesh)
        end subroutine TotalHeisenbergEnergy

        ! Replaces calculateHeisenbergEnergy (src/energyMin.f90:86-171): local
        ! energies of atom atomIndex (1-based) before and after S_proposed,
        ! oldEnergy = -S.h and newEnergy = -S'.h from the device field h. The
        ! reference passes S_proposed as its own vector type; here it is a
        ! plain real(8) 3-vector. With demag (default .true., as the reference's)
        ! the energies include -mu_0 S.H with H = demagnetisation_array(atomIndex,:).
        subroutine calculateHeisenbergEnergy(chainMesh, atomIndex, J, J_prime, Dz, Dz_prime, B, &
                        lockArray, S_proposed, oldEnergy, newEnergy, demagnetisation_array, demag)
                type(ChainMesh_t), intent(inout) :: chainMesh
                integer, intent(in) :: atomIndex
                real(kind=8), intent(in) :: J, J_prime, Dz, Dz_prime, B
                integer(kind=OMP_LOCK_KIND), intent(inout) :: lockArray(:)
                real(kind=8), intent(in) :: S_proposed(3)
                real(kind=8), intent(inout) :: oldEnergy, newEnergy
                real(kind=8), dimension(:,:), intent(in) :: demagnetisation_array
                logical, optional, intent(in) :: demag
                integer(c_int32_t) :: sites(2)
                real(c_double) :: props(6), de(2)
                integer(c_int) :: rc
                logical :: dm
                call check_supported(J_prime)
                dm = demag_on(demag)
                if (atomIndex < 1 .or. atomIndex > chainMesh%numAtoms) error stop "calculateHeisenbergEnergy: atom index"
                call chainmesh_push_spins(chainMesh)
                rc = ising_set_hamiltonian(chainMesh%gpu, 0_c_int, reference_hamiltonian(J, Dz, B, 2.0_8, dm))
                call ising_check(rc, chainMesh%gpu, "calculateHeisenbergEnergy: Hamiltonian")
                if (dm) then
                        if (size(demagnetisation_array, 1) /= chainMesh%numAtoms .or. size(demagnetisation_array, 2) /= 3) &
                                error stop "calculateHeisenbergEnergy: demagnetisation_array must be (numAtoms, 3)"
                        rc = ising_set_demag_field(chainMesh%gpu, 0_c_int, demagnetisation_array, &
                                                   int(chainMesh%numAtoms, c_int64_t))
                        call ising_check(rc, chainMesh%gpu, "calculateHeisenbergEnergy: demagnetisation_array")
                end if
                sites = int(atomIndex - 1, c_int32_t)
                props(1:3) = 0.0_8
                props(4:6) = S_proposed
                rc = ising_delta_energy(chainMesh%gpu, 0_c_int, sites, props, 2_c_int, de)
                call ising_check(rc, chainMesh%gpu, "calculateHeisenbergEnergy")
                ! de(1) = E(0) - E(S) = S.h, so the local energy of S is -de(1).
                oldEnergy = -de(1)
                newEnergy = oldEnergy + de(2)
        end subroutine calculateHeisenbergEnergy
end module EnergyMin
