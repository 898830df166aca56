! ref_harness.f90 -- C-callable probes of the reference's leaf modules
! (constants, Rand, vecNd, Atom, unit_cells), compiled from the reference sources as they lie
! under /root/reference/src by oracle/ref/Makefile into oracle/_ref/.
! Test infrastructure only: tests/test_oracle_ref.py compares the oracle's
! restatements with these.
module ref_harness
        use iso_c_binding
        use constants, only: pi, mu_0
        use Rand, only: random, makeRandom, algor_uniform_random
        use vecNd
        use Atom, only: Atom_t
        use unit_cells, only: bcc_unit_cell, hyper_kagome_unit_cell
        implicit none
contains
        ! n successive draws of algor_uniform_random after makeRandom(seed)
        ! (src/rand.f90:11-95).
        subroutine ref_uniform_stream(seed, n, out) bind(C, name="ref_uniform_stream")
                integer(c_int32_t), value :: seed
                integer(c_int), value :: n
                real(c_double), intent(out) :: out(n)
                type(random) :: r
                integer :: i
                r = makeRandom(seed)
                do i = 1, n
                        out(i) = algor_uniform_random(r)
                end do
        end subroutine ref_uniform_stream

        ! The value of constants::pi (src/constants.f90:16).
        function ref_pi() bind(C, name="ref_pi") result(p)
                real(c_double) :: p
                p = pi
        end function ref_pi

        ! The value of constants::mu_0 (src/constants.f90:15), the
        ! permeability the demagnetisation energy is scaled by.
        function ref_mu0() bind(C, name="ref_mu0") result(p)
                real(c_double) :: p
                p = mu_0
        end function ref_mu0

        ! a .x. b and a * b of vecNd (src/vecNd.f90:213-237, 413-422).
        subroutine ref_cross_dot(a, b, c, d) bind(C, name="ref_cross_dot")
                real(c_double), intent(in) :: a(3), b(3)
                real(c_double), intent(out) :: c(3), d
                type(vecNd_t) :: va, vb, vc
                va = a
                vb = b
                vc = va .x. vb
                c = vc%coords
                d = va * vb
        end subroutine ref_cross_dot

        ! Imaginary part of cmplx(0.0_8, x) with no kind argument, the form
        ! src/reciprocalSpaceProcesses.f90:293 uses: Fortran gives it the
        ! default (single precision) complex kind.
        function ref_cmplx_default_kind_imag(x) bind(C, name="ref_cmplx_default_kind_imag") result(y)
                real(c_double), value :: x
                real(c_double) :: y
                complex(kind=8) :: z
                z = 1.0_8 + cmplx(0.0_8, x)
                y = aimag(z)
        end function ref_cmplx_default_kind_imag

        ! The unit cells of src/unit_cells.f90 (kind 0: bcc_unit_cell, 1:
        ! hyper_kagome_unit_cell): edge lengths, angles in degrees, and the
        ! atoms' fractional positions as stored (default real).
        subroutine ref_unit_cell(kind, abc, angles, n, xyz) bind(C, name="ref_unit_cell")
                integer(c_int), value :: kind
                real(c_double), intent(out) :: abc(3), angles(3)
                integer(c_int), intent(out) :: n
                real(c_double), intent(out) :: xyz(3, 16)
                type(Atom_t), allocatable :: atoms(:)
                real(kind=8) :: a, b, c, al, be, ga
                integer :: i
                if (kind == 0) then
                        call bcc_unit_cell(a, b, c, al, be, ga, atoms)
                else
                        call hyper_kagome_unit_cell(a, b, c, al, be, ga, atoms)
                end if
                abc = [a, b, c]
                angles = [al, be, ga]
                n = size(atoms)
                xyz = 0.0_c_double
                do i = 1, n
                        xyz(:, i) = [real(atoms(i)%x, c_double), real(atoms(i)%y, c_double), real(atoms(i)%z, c_double)]
                end do
        end subroutine ref_unit_cell
end module ref_harness
