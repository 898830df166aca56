/*
 * demag_oracle.c -- FP64 restatement of the reference's demagnetising field.
 * TEST INFRASTRUCTURE ONLY (see ising_oracle.h).
 *
 * orc_ref_demag_field follows calculate_demagnetisation_field
 * (src/reciprocalSpaceProcesses.f90:202-281) step by step:
 *   1. cloud-in-cell deposit of the spins onto one mesh point per unit cell
 *      (interpolate_to_fft_array, :76-129): the point of cell c gathers, from
 *      the atoms of the 27 cells around it (getNeighboringCells,
 *      src/chainMesh.f90:85-131, which lists a cell twice or three times
 *      when an axis has fewer than three cells), weight
 *      prod_a max(1 - |d_a|, 0), d = minimum-image Bravais coefficients of
 *      atom - centre(c) (distance_points_vec_bravais,
 *      src/chainMesh.f90:244-276; nint rounds half away from zero);
 *   2. forward real-to-complex 3D transform, FFTW's sign e^{-2 pi i k n/N},
 *      x halved to kx = 0..nx/2 (fftw_plan_dft_r2c_3d(M,L,N,...) on the
 *      Fortran array (padX, L, M), src/chainMesh.f90:717);
 *   3. projection H(k) = -(k.M(k)) k / k^2 with k = 2 pi (wx/N ar + wy/L br
 *      + wz/M cr), wx = kx, wy, wz folded to (-n/2, n/2], the k = 0 mode
 *      and modes beyond the wave-guide cutoff |k_a| > pi/a zeroed (:237-270);
 *      lattice units, a = 1, so ar = e_x etc.;
 *   4. complex-to-real inverse (unnormalised) divided by N L M (:271-275):
 *      inverse complex transforms along z and y, then the real inverse along
 *      x, which reads only the real parts of its DC and Nyquist terms (the
 *      c2r convention of FFTW's backward real codelets);
 *   5. the same cloud-in-cell weights carry the mesh field back to the atoms
 *      (interpolate_fft_to_atoms, :131-178).
 *
 * Pinning (tests/test_oracle_demag.py): closed forms of a single
 * magnetisation mode m0 cos(q.r + phi) on simple-cubic and bcc cells
 * (longitudinal, transverse, uniform); a brute-force enumeration of the
 * deposit on hyper-kagome cells (atoms outside their own cell); an
 * independent numpy.fft
 * (rfftn/irfftn) route for the transform and projection, including the
 * Nyquist planes; a brute-force enumeration of the 27-cell deposit with its
 * repeated cells on axes of one and two cells; the reference's mu_0 from its
 * own constants.f90 (tests/golden/ref_leaf.json).
 */
#include "ising_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* constants::pi of the reference: a default-real literal, i.e. the
 * single-precision value (src/constants.f90:16). */
#define REF_PI_F 3.1415927410125732421875

/* getNeighboringCells (src/chainMesh.f90:98-117) wraps one step: n -> 0,
 * -1 -> n-1. */
static inline int wrap_cell(int i, int n)
{
    if (i == n) return 0;
    if (i == -1) return n - 1;
    return i;
}

/* distance_points_vec_bravais (src/chainMesh.f90:268-270): d - n nint(d/n). */
static inline double min_image(double d, int n)
{
    return d - (double)n * round(d / (double)n);
}

/* Cloud-in-cell weight of atom `a` of cell (ni, nj, nk) at the mesh point
 * (centre) of cell (ci, cj, ck); atom positions in lattice units, the cell
 * corner plus the atom's fraction of the cell (orc_atom_frac). An atom that
 * lies outside its own cell (hyper-kagome fractions down to -3/4) still
 * deposits only on the 27 cells around the cell it belongs to, as the
 * reference's loop over getNeighboringCells does. */
static double cic_weight(const orc_lattice* L, int ci, int cj, int ck, int ni, int nj, int nk, int a)
{
    double f[3];
    orc_atom_frac(L, a, f);
    const double dx = min_image(((double)ni + f[0]) - ((double)ci + 0.5), L->lx);
    const double dy = min_image(((double)nj + f[1]) - ((double)cj + 0.5), L->ly);
    const double dz = min_image(((double)nk + f[2]) - ((double)ck + 0.5), L->lz);
    const double wx = fmax(1.0 - fabs(dx), 0.0);
    const double wy = fmax(1.0 - fabs(dy), 0.0);
    const double wz = fmax(1.0 - fabs(dz), 0.0);
    return wx * wy * wz;
}

/* e^{sign 2 pi i m / n} */
static inline double complex twiddle(int64_t m, int n, int sign)
{
    const double t = 2.0 * M_PI * (double)(m % n) / (double)n;
    return cos(t) + (double)sign * sin(t) * I;
}

void orc_demag_deposit(const orc_lattice* L, const double* s, double* mesh)
{
    const int nx = L->lx, ny = L->ly, nz = L->lz, na = orc_atoms_per_cell(L);
    const int64_t Nc = (int64_t)nx * ny * nz, N = Nc * na;
    memset(mesh, 0, 3 * (size_t)Nc * sizeof(double));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t cell = ((int64_t)i * ny + j) * nz + k;
                for (int di = -1; di <= 1; ++di)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int dk = -1; dk <= 1; ++dk) {
                            const int ni = wrap_cell(i + di, nx), nj = wrap_cell(j + dj, ny), nk = wrap_cell(k + dk, nz);
                            for (int a = 0; a < na; ++a) {
                                const double w = cic_weight(L, i, j, k, ni, nj, nk, a);
                                const int64_t site = (((int64_t)ni * ny + nj) * nz + nk) * na + a;
                                for (int c = 0; c < 3; ++c) mesh[c * Nc + cell] += w * s[c * N + site];
                            }
                        }
            }
}

void orc_demag_mesh_field(const orc_lattice* L, const double* mesh, double* hmesh)
{
    const int nx = L->lx, ny = L->ly, nz = L->lz, hx = nx / 2 + 1;
    const int64_t Nc = (int64_t)nx * ny * nz, Nk = (int64_t)hx * ny * nz;
    double complex* F = malloc(3 * (size_t)Nk * sizeof(double complex));
    double complex* T = malloc(3 * (size_t)Nk * sizeof(double complex));
#define KI(kx, ky, kz) ((((int64_t)(kx)) * ny + (ky)) * nz + (kz))
    /* forward: x (real to half complex), then y, then z */
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < hx; ++kx)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double complex acc = 0.0;
                    for (int n = 0; n < nx; ++n)
                        acc += mesh[c * Nc + ((int64_t)n * ny + j) * nz + k] * twiddle((int64_t)kx * n, nx, -1);
                    T[c * Nk + KI(kx, j, k)] = acc;
                }
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < hx; ++kx)
            for (int ky = 0; ky < ny; ++ky)
                for (int k = 0; k < nz; ++k) {
                    double complex acc = 0.0;
                    for (int n = 0; n < ny; ++n) acc += T[c * Nk + KI(kx, n, k)] * twiddle((int64_t)ky * n, ny, -1);
                    F[c * Nk + KI(kx, ky, k)] = acc;
                }
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < hx; ++kx)
            for (int ky = 0; ky < ny; ++ky)
                for (int kz = 0; kz < nz; ++kz) {
                    double complex acc = 0.0;
                    for (int n = 0; n < nz; ++n) acc += F[c * Nk + KI(kx, ky, n)] * twiddle((int64_t)kz * n, nz, -1);
                    T[c * Nk + KI(kx, ky, kz)] = acc;
                }
    /* projection */
    for (int kx = 0; kx < hx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            for (int kz = 0; kz < nz; ++kz) {
                const int64_t e = KI(kx, ky, kz);
                const int wx = kx;
                const int wy = ky > ny / 2 ? ky - ny : ky;
                const int wz = kz > nz / 2 ? kz - nz : kz;
                const double kv[3] = { 2.0 * REF_PI_F * ((double)wx / (double)nx),
                                       2.0 * REF_PI_F * ((double)wy / (double)ny),
                                       2.0 * REF_PI_F * ((double)wz / (double)nz) };
                if (kx == 0 && ky == 0 && kz == 0) {
                    for (int c = 0; c < 3; ++c) T[c * Nk + e] = 0.0;
                    continue;
                }
                const double complex kdm = kv[0] * T[e] + kv[1] * T[Nk + e] + kv[2] * T[2 * Nk + e];
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                const int cut = fabs(kv[0]) > REF_PI_F || fabs(kv[1]) > REF_PI_F || fabs(kv[2]) > REF_PI_F;
                for (int c = 0; c < 3; ++c) T[c * Nk + e] = cut ? 0.0 : -(kdm / k2) * kv[c];
            }
    /* inverse: z, then y (complex), then x (half complex to real) */
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < hx; ++kx)
            for (int ky = 0; ky < ny; ++ky)
                for (int k = 0; k < nz; ++k) {
                    double complex acc = 0.0;
                    for (int n = 0; n < nz; ++n) acc += T[c * Nk + KI(kx, ky, n)] * twiddle((int64_t)k * n, nz, +1);
                    F[c * Nk + KI(kx, ky, k)] = acc;
                }
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < hx; ++kx)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double complex acc = 0.0;
                    for (int n = 0; n < ny; ++n) acc += F[c * Nk + KI(kx, n, k)] * twiddle((int64_t)j * n, ny, +1);
                    T[c * Nk + KI(kx, j, k)] = acc;
                }
    const int nyq = (nx % 2 == 0) ? nx / 2 : -1;
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double acc = creal(T[c * Nk + KI(0, j, k)]);
                    for (int kx = 1; kx < hx; ++kx) {
                        const double complex v = T[c * Nk + KI(kx, j, k)];
                        if (kx == nyq) acc += creal(v) * ((i & 1) ? -1.0 : 1.0);
                        else acc += 2.0 * creal(v * twiddle((int64_t)kx * i, nx, +1));
                    }
                    hmesh[c * Nc + ((int64_t)i * ny + j) * nz + k] = acc / (double)Nc;
                }
#undef KI
    free(F);
    free(T);
}

void orc_demag_gather(const orc_lattice* L, const double* hmesh, double* H)
{
    const int nx = L->lx, ny = L->ly, nz = L->lz, na = orc_atoms_per_cell(L);
    const int64_t Nc = (int64_t)nx * ny * nz, N = Nc * na;
    memset(H, 0, 3 * (size_t)N * sizeof(double));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t cell = ((int64_t)i * ny + j) * nz + k;
                for (int di = -1; di <= 1; ++di)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int dk = -1; dk <= 1; ++dk) {
                            const int ni = wrap_cell(i + di, nx), nj = wrap_cell(j + dj, ny), nk = wrap_cell(k + dk, nz);
                            for (int a = 0; a < na; ++a) {
                                const double w = cic_weight(L, i, j, k, ni, nj, nk, a);
                                const int64_t site = (((int64_t)ni * ny + nj) * nz + nk) * na + a;
                                for (int c = 0; c < 3; ++c) H[c * N + site] += w * hmesh[c * Nc + cell];
                            }
                        }
            }
}

int orc_ref_demag_field(const orc_lattice* L, const double* s, double* H)
{
    if (!L->periodic[0] || !L->periodic[1] || !L->periodic[2]) return -1;
    const int64_t Nc = (int64_t)L->lx * L->ly * L->lz;
    double* m = malloc(3 * (size_t)Nc * sizeof(double));
    double* hm = malloc(3 * (size_t)Nc * sizeof(double));
    orc_demag_deposit(L, s, m);
    orc_demag_mesh_field(L, m, hm);
    orc_demag_gather(L, hm, H);
    free(m);
    free(hm);
    return 0;
}

/* Demagnetisation part of TotalHeisenbergEnergy: -0.5 mu_0 S.H per atom
 * (calculateTotalHeisenbergEnergyHelper, src/energyMin.f90:223-226). */
double orc_demag_energy(const orc_lattice* L, const double* s, const double* H, double mu0)
{
    const int64_t N = orc_nsites(L);
    double e = 0.0;
    for (int64_t i = 0; i < N; ++i)
        e += -(0.5 * mu0 * (s[i] * H[i] + s[N + i] * H[N + i] + s[2 * N + i] * H[2 * N + i]));
    return e;
}
