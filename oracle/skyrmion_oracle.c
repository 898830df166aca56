/*
 * skyrmion_oracle.c -- FP64 CPU oracle of the skyrmion count of one layer
 * (the statistic chainMesh_statistics reports as skyrmion_number_middle).
 * TEST INFRASTRUCTURE ONLY: tests/ compare the device against it; the product's
 * kernels live in ising_amd/csrc/ising_hip.hip.
 *
 * Reference: compute_skyrmion_distribution (src/reciprocalSpaceProcesses.f90:
 * 930-983) -> calculate_skyrmion_number (:627-812) ->
 * calculate_winding_density_interpolated (:365-431) ->
 * spectral_interpolate_to_fine (:12-70), Gaussian_filter_2d (:814-857),
 * called from chainMesh_statistics (src/PT-Utils.f90:54-56). The reference's
 * pipeline cannot be restated as written (FFTW's overlapping output rows,
 * negative y frequencies copied to positive ones, neighbour indices wrapped
 * with the coarse cell counts on the fine grid, a count incremented once per
 * pixel of a qualifying region); this file states the intended semantics,
 * step by step:
 *
 *   1. coarse grid: the spin of each cell's corner atom of layer z (the atom
 *      of map_spins_to_std_grid, chainMesh.f90:54-83), Nx x Ny, periodic;
 *   2. fine grid (sx Nx) x (sy Ny): band-limited (trigonometric)
 *      interpolation of each component, the Nyquist term of an even axis
 *      split evenly between +-N/2 so the interpolant is real and passes
 *      through the samples; separable, x first then y, sums in ascending
 *      order with kernel tables D_N(r/s + m);
 *   3. density per fine plaquette (i, j): [sigma(s1,s2,s3) + sigma(s1,s3,s4)]
 *      / (4 pi), s1 = (i,j), s2 = (i+1,j), s3 = (i+1,j+1), s4 = (i,j+1),
 *      renormalised spins, periodic (:410-426); a zero interpolated spin
 *      gives the plaquette density 0;
 *   4. 5x5 Gaussian filter, centred (the reference's kernel is offset by 3
 *      cells in x and y, a translation of the filtered field), weights
 *      exp(-(a^2+b^2)/(2 sigma^2)) normalised by their sum, periodic;
 *   5. mask |f| > q max|f|;
 *   6. regions: every masked pixel joins the local maximum its steepest
 *      ascent over the 8 neighbours reaches (pixels ordered by |f|, ties by
 *      the reference's scan order x-major, earlier first); the reference's
 *      scan-order flood fill from each local maximum gives the same regions
 *      where peaks are separated by unmasked pixels;
 *   7. region charge: sum of the filtered density over the region, in fixed
 *      point (round(f 2^40), int64) so the sum is exact and order-free;
 *   8. count: each region with | |charge| - particle_number | < 0.05 adds
 *      sign(charge) once (:748-756 add it once per pixel of the region);
 *   9. distribution: thresholds q_i = max - (i/T)(max - min), i = 1..T,
 *      widths sigma_k = 1 - (k/S)(1 - 1e-5), k = 1..S, visited i-major; the
 *      first (i, k) whose sum_j j count_j lies within 0.1 of the layer's
 *      winding number Q is returned, else the counts whose nint(sum)
 *      came closest to nint(Q) first (:955-981).
 *
 * Pinning (tests/test_oracle_skyrmion.py): the interpolation against an
 * independent numpy FFT zero-padding route and exactly at the coarse sites;
 * the density's total against the integer charge of smooth textures
 * (Belavin-Polyakov skyrmions, ferromagnet); the filter against
 * scipy.ndimage with wrap-around; counts of textures with a known number of
 * separated skyrmions, their sign under S -> -S and translation invariance;
 * the distribution search against hand-built count tables.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ising_oracle.h"

#define SK_FIXED_SCALE 1099511627776.0   /* 2^40 */
#define SK_FIXED_INV (1.0 / SK_FIXED_SCALE)

static int64_t wrapi(int64_t v, int64_t n) { v %= n; return v < 0 ? v + n : v; }

/* D_N(u): the band-limited interpolation kernel of N periodic samples. */
static double interp_kernel(int n, long double u)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    long double acc = 1.0L;
    const int kmax = (n % 2 == 0) ? n / 2 - 1 : (n - 1) / 2;
    for (int k = 1; k <= kmax; ++k) acc += 2.0L * cosl(2.0L * pi * (long double)k * u / (long double)n);
    if (n % 2 == 0) acc += cosl(pi * u);
    return (double)(acc / (long double)n);
}

/* Table K[r][m] = D_N(r/s + m), r < s, m < N. */
void orc_sk_kernel_table(int n, int s, double* K)
{
    for (int r = 0; r < s; ++r)
        for (int m = 0; m < n; ++m) K[(size_t)r * n + m] = interp_kernel(n, (long double)r / (long double)s + m);
}

/* Coarse grid of layer z: corner spin of cell (i, j) at S[c][j][i]. */
void orc_sk_coarse(const orc_lattice* L, const double* s, int z, double* S)
{
    const int64_t N = orc_nsites(L);
    const int nx = L->lx, ny = L->ly;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int64_t ids[4];
            orc_plaquette(L, i, j, z, ids);    /* ids[0]: the corner atom of cell (i, j, z) */
            /* the device stores FP32 spins */
            for (int c = 0; c < 3; ++c) S[((size_t)c * ny + j) * nx + i] = (double)(float)s[c * N + ids[0]];
        }
}

/* Fine grid F[c][J][I] (NX = sx nx, NY = sy ny) from the coarse grid. */
void orc_sk_interpolate(int nx, int ny, int sx, int sy, const double* S, double* F)
{
    const int NX = sx * nx, NY = sy * ny;
    double* Kx = malloc(sizeof(double) * (size_t)sx * nx);
    double* Ky = malloc(sizeof(double) * (size_t)sy * ny);
    double* T = malloc(sizeof(double) * 3 * (size_t)ny * NX);
    orc_sk_kernel_table(nx, sx, Kx);
    orc_sk_kernel_table(ny, sy, Ky);
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < ny; ++j)
            for (int I = 0; I < NX; ++I) {
                const int n0 = I / sx, r = I % sx;
                double acc = 0.0;
                for (int m = 0; m < nx; ++m)
                    acc += Kx[(size_t)r * nx + m] * S[((size_t)c * ny + j) * nx + wrapi(n0 - m, nx)];
                T[((size_t)c * ny + j) * NX + I] = acc;
            }
    for (int c = 0; c < 3; ++c)
        for (int J = 0; J < NY; ++J)
            for (int I = 0; I < NX; ++I) {
                const int m0 = J / sy, r = J % sy;
                double acc = 0.0;
                for (int m = 0; m < ny; ++m)
                    acc += Ky[(size_t)r * ny + m] * T[((size_t)c * ny + wrapi(m0 - m, ny)) * NX + I];
                F[((size_t)c * NY + J) * NX + I] = acc;
            }
    free(Kx); free(Ky); free(T);
}

static int unit(const double* F, size_t plane, size_t idx, double o[3])
{
    const double x = F[idx], y = F[plane + idx], z = F[2 * plane + idx];
    const double nn = sqrt(x * x + y * y + z * z);
    if (!(nn > 0.0)) return 0;
    o[0] = x / nn; o[1] = y / nn; o[2] = z / nn;
    return 1;
}

static double sigma3(const double a[3], const double b[3], const double c[3])
{
    const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    const double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    const double x0 = b[1] * c[2] - b[2] * c[1];
    const double x1 = b[2] * c[0] - b[0] * c[2];
    const double x2 = b[0] * c[1] - b[1] * c[0];
    const double re = ((1.0 + ab) + bc) + ca;
    const double im = a[0] * x0 + a[1] * x1 + a[2] * x2;
    return 2.0 * atan2(im, re);
}

/* Density d[J][I] of the fine grid F[c][J][I]. */
void orc_sk_density(int NX, int NY, const double* F, double* d)
{
    const size_t plane = (size_t)NX * NY;
    for (int J = 0; J < NY; ++J)
        for (int I = 0; I < NX; ++I) {
            const int I2 = (I + 1) % NX, J2 = (J + 1) % NY;
            double s1[3], s2[3], s3[3], s4[3];
            double v = 0.0;
            if (unit(F, plane, (size_t)J * NX + I, s1) && unit(F, plane, (size_t)J * NX + I2, s2) &&
                unit(F, plane, (size_t)J2 * NX + I2, s3) && unit(F, plane, (size_t)J2 * NX + I, s4))
                v = (sigma3(s1, s2, s3) + sigma3(s1, s3, s4)) / (4.0 * M_PI);
            d[(size_t)J * NX + I] = v;
        }
}

/* Normalised 5x5 weights G[a+2][b+2] (a over x, b over y). */
void orc_sk_gauss_weights(double sigma, double G[25])
{
    double acc = 0.0;
    for (int a = -2; a <= 2; ++a)
        for (int b = -2; b <= 2; ++b) {
            const double w = exp(-(double)(a * a + b * b) / (2.0 * sigma * sigma));
            G[(a + 2) * 5 + (b + 2)] = w;
            acc += w;
        }
    for (int k = 0; k < 25; ++k) G[k] /= acc;
}

/* f = G (*) d, centred, periodic. */
void orc_sk_filter(int NX, int NY, const double* d, double sigma, double* f)
{
    double G[25];
    orc_sk_gauss_weights(sigma, G);
    for (int J = 0; J < NY; ++J)
        for (int I = 0; I < NX; ++I) {
            double acc = 0.0;
            for (int a = -2; a <= 2; ++a)
                for (int b = -2; b <= 2; ++b)
                    acc += d[(size_t)wrapi(J - b, NY) * NX + wrapi(I - a, NX)] * G[(a + 2) * 5 + (b + 2)];
            f[(size_t)J * NX + I] = acc;
        }
}

/* Pixel (I, J) before (I', J') in the reference's scan order (I outer). */
static int64_t scan_rank(int NY, int I, int J) { return (int64_t)I * NY + J; }

/* Steepest-ascent root over the 8 neighbours of every pixel with |f| > thr
 * (J*NX + I; -1 for the others, which no threshold masks). */
static void roots_of(int NX, int NY, const double* f, double thr, int64_t* root)
{
    const size_t n = (size_t)NX * NY;
    int64_t* up = malloc(sizeof(int64_t) * n);
    for (int J = 0; J < NY; ++J)
        for (int I = 0; I < NX; ++I) {
            const size_t p = (size_t)J * NX + I;
            double bv = fabs(f[p]);
            int64_t br = scan_rank(NY, I, J), best = (int64_t)p;
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b) {
                    if (a == 0 && b == 0) continue;
                    const int I2 = (int)wrapi(I + a, NX), J2 = (int)wrapi(J + b, NY);
                    const size_t q = (size_t)J2 * NX + I2;
                    const double v = fabs(f[q]);
                    const int64_t r = scan_rank(NY, I2, J2);
                    if (v > bv || (v == bv && r < br)) { bv = v; br = r; best = (int64_t)q; }
                }
            up[p] = best;
        }
    for (size_t p = 0; p < n; ++p) {
        if (!(fabs(f[p]) > thr)) { root[p] = -1; continue; }
        int64_t q = (int64_t)p;
        while (up[q] != q) q = up[q];
        root[p] = q;
    }
    free(up);
}

/* counts[t][k][j] (t < T thresholds, k < S widths, j < n_part) of the
 * density d[J][I]. */
void orc_sk_count_table(int NX, int NY, const double* d, int n_part, int T, int S, double min_t, double max_t,
                        int32_t* counts)
{
    const size_t n = (size_t)NX * NY;
    double* f = malloc(sizeof(double) * n);
    int64_t* root = malloc(sizeof(int64_t) * n);
    int64_t* acc = malloc(sizeof(int64_t) * n);
    memset(counts, 0, sizeof(int32_t) * (size_t)T * S * n_part);
    for (int k = 1; k <= S; ++k) {
        const double sigma = 1.0 - ((double)k / (double)S) * (1.0 - 0.00001);
        orc_sk_filter(NX, NY, d, sigma, f);
        double mx = 0.0;
        for (size_t p = 0; p < n; ++p) mx = fmax(mx, fabs(f[p]));
        /* the lowest threshold, by the same expression as in the loop below */
        roots_of(NX, NY, f, (max_t - ((double)T / (double)T) * (max_t - min_t)) * mx, root);
        for (int t = 1; t <= T; ++t) {
            const double q = max_t - ((double)t / (double)T) * (max_t - min_t);
            const double thr = q * mx;
            memset(acc, 0, sizeof(int64_t) * n);
            for (size_t p = 0; p < n; ++p)
                if (fabs(f[p]) > thr) acc[root[p]] += llrint(f[p] * SK_FIXED_SCALE);
            for (size_t p = 0; p < n; ++p) {
                if ((size_t)root[p] != p || !(fabs(f[p]) > thr)) continue;
                const double ch = (double)acc[p] * SK_FIXED_INV;
                for (int j = 1; j <= n_part; ++j)
                    if (fabs(fabs(ch) - (double)j) < 0.05)
                        counts[((size_t)(t - 1) * S + (k - 1)) * n_part + (j - 1)] += ch >= 0.0 ? 1 : -1;
            }
        }
    }
    free(f); free(root); free(acc);
}

/* The search of compute_skyrmion_distribution over a count table. */
void orc_sk_select(const int32_t* counts, int n_part, int T, int S, double q, int32_t* out)
{
    int32_t* closest = calloc((size_t)n_part, sizeof(int32_t));
    const long long nq = llround(q);
    for (int t = 0; t < T; ++t)
        for (int k = 0; k < S; ++k) {
            const int32_t* arr = counts + ((size_t)t * S + k) * n_part;
            double winding = 0.0;
            long long current = 0;
            for (int j = 1; j <= n_part; ++j) {
                winding += (double)j * arr[j - 1];
                current += (long long)j * closest[j - 1];
            }
            if (llabs(llround(winding) - nq) < llabs(current - nq)) memcpy(closest, arr, sizeof(int32_t) * n_part);
            if (fabs(winding - q) < 0.1) {
                memcpy(out, arr, sizeof(int32_t) * n_part);
                free(closest);
                return;
            }
        }
    memcpy(out, closest, sizeof(int32_t) * n_part);
    free(closest);
}

/* The whole statistic of layer z of spins s: counts[n_part]; the fine
 * density (NX*NY doubles) is returned in d when d is not NULL. Returns -1
 * unless x and y are periodic. */
int orc_sk_distribution(const orc_lattice* L, const double* s, int z, int sx, int sy, int n_part, double min_t,
                        double max_t, int T, int S, int32_t* out, int32_t* table, double* d_out)
{
    if (!L->periodic[0] || !L->periodic[1]) return -1;
    const int nx = L->lx, ny = L->ly, NX = sx * nx, NY = sy * ny;
    double* Sg = malloc(sizeof(double) * 3 * (size_t)nx * ny);
    double* F = malloc(sizeof(double) * 3 * (size_t)NX * NY);
    double* d = malloc(sizeof(double) * (size_t)NX * NY);
    int32_t* tab = table ? table : malloc(sizeof(int32_t) * (size_t)T * S * n_part);
    orc_sk_coarse(L, s, z, Sg);
    orc_sk_interpolate(nx, ny, sx, sy, Sg, F);
    orc_sk_density(NX, NY, F, d);
    orc_sk_count_table(NX, NY, d, n_part, T, S, min_t, max_t, tab);
    int64_t qi;
    const double q = orc_winding_layer(L, s, z, &qi);
    orc_sk_select(tab, n_part, T, S, q, out);
    if (d_out) memcpy(d_out, d, sizeof(double) * (size_t)NX * NY);
    if (!table) free(tab);
    free(Sg); free(F); free(d);
    return 0;
}
