/*
 * ising_oracle.c -- FP64 CPU oracle. TEST INFRASTRUCTURE ONLY (see header).
 *
 * Compiled with -O2 -ffp-contract=off so every expression is evaluated as
 * written (no fused multiply-add), which the topological-charge comparison
 * relies on.
 *
 * Pinning of each function (where the check lives):
 *   orc_philox4x32_10       Random123 known-answer vectors (tests/test_oracle_rng.py)
 *   orc_ref_uniform         reference rand.f90 built into oracle/_ref (tests/test_oracle_ref.py)
 *   orc_neighbour           brute-force enumeration on tiny lattices (tests/test_oracle_geometry.py)
 *   orc_bcc_bonds           brute-force distance enumeration of bcc positions (tests/test_oracle_bcc.py)
 *   orc_hk_bonds            brute-force minimum-image distance enumeration over the whole
 *                           periodic lattice; four bonds per site, each bond seen from both
 *                           ends (tests/test_oracle_hyperkagome.py)
 *   orc_site_colour         no bond joins two sites of one colour (same tests)
 *   hyper-kagome energies   closed forms (ferromagnet, uniform rotation of the bond sum),
 *                           dE = E(after) - E(before), dE linear in the proposal
 *   orc_plaquette           the reference's firstAtomInMeshCell list order and
 *                           computeAdjacentAtoms restated on atom positions (bcc,
 *                           hyper-kagome)
 *   bcc energies            closed forms (ferromagnet, Neel, bulk and interfacial helices),
 *                           ref restatement in reference mode (tests/test_oracle_bcc.py)
 *   orc_total_energy        closed forms: ferromagnet, bulk helix (tests/test_oracle_energy.py)
 *   orc_delta_energy        E(after) - E(before) of orc_total_energy, plus closed forms
 *   orc_local_field         via orc_delta_energy linearity in the proposal
 *   orc_ref_local_energy    Neel-spiral closed form in reference units; ref dE = orc dE in
 *                           reference mode (independent geometric construction)
 *   orc_ref_total_energy    Neel-spiral closed form; equals orc_total_energy in reference mode
 *   orc_ref_winding_number  ferromagnet -> 0, skyrmion -> +-1, Q(-S) = -Q(S), mirror -> -Q
 *   orc_winding_layer       same invariants; equals orc_ref_winding_number to 1e-6
 *   orc_ref_winding_density sums to orc_ref_winding_number per layer; FM -> 0
 *   orc_winding_density     sums to orc_winding_layer; odd under S -> -S; equals
 *                           orc_ref_winding_density to 1e-6 (tests/test_oracle_winding.py)
 *   orc_propose             analytic distribution of S.S' (KS test), symmetry
 *   orc_sweeps              single-spin Langevin <Sz>, two-spin dimer quadrature,
 *                           high-temperature series (tests/test_oracle_metropolis.py)
 *   orc_pt_swap_round       two-replica detailed balance (tests/test_oracle_pt.py)
 *   orc_delta_energy_field  orc_delta_energy plus the field term; the field itself is
 *                           pinned in demag_oracle.c (tests/test_oracle_demag.py)
 *   orc_sweeps_demag        field schedule of Metropolis_mcs; single spin in a fixed
 *                           external field follows the Langevin law (tests/test_oracle_demag.py)
 */
#include "ising_oracle.h"
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of
 * two 32x32->64 multiplies; key bumped by the Weyl constants between rounds. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Counter layout of one site update: key = (seed_lo, slot),
 * ctr = (site in reference order, sweep, stream, seed_hi). */
void orc_metropolis_random(uint64_t seed, uint32_t slot, uint32_t site, uint32_t sweep, uint32_t out[4])
{
    uint32_t ctr[4] = { site, sweep, ORC_STREAM_METROPOLIS, (uint32_t)(seed >> 32) };
    uint32_t key[2] = { (uint32_t)seed, slot };
    orc_philox4x32_10(ctr, key, out);
}

static inline double u24(uint32_t x) { return (double)(x >> 8) * 0x1p-24; }       /* [0,1)  */
static inline double s24(uint32_t x) { return (double)(x >> 8) * 0x1p-23 - 1.0; } /* [-1,1) */

/* Reference generator, src/rand.f90:11-18 (seeding) and :75-92 (one draw):
 * xorshift (13,-17,5) on ix, Park-Miller 16807 by Schrage on iy,
 * combined as ((ix xor iy) and (2^31-1)) or 1, scaled by 1/(2^31-1). */
orc_ref_rng orc_ref_make_random(int32_t seed)
{
    orc_ref_rng r;
    r.seed = seed;
    r.ix = (int32_t)(777755555 ^ seed);
    r.iy = (int32_t)((888889999 ^ seed) | 1);
    return r;
}

double orc_ref_uniform(orc_ref_rng* r)
{
    uint32_t ix = (uint32_t)r->ix;
    ix ^= ix << 13;
    ix ^= ix >> 17;           /* ishft with negative count is a logical shift */
    ix ^= ix << 5;
    int32_t iy = r->iy;
    int32_t q = iy / 127773;
    iy = 16807 * (iy - q * 127773) - q * 2836;
    if (iy < 0) iy += 2147483647;
    r->ix = (int32_t)ix;
    r->iy = iy;
    int32_t comb = (int32_t)((((uint32_t)ix ^ (uint32_t)iy) & 2147483647u) | 1u);
    return (1.0 / 2147483647.0) * (double)comb;
}

/* ---- geometry ------------------------------------------------------- */

int64_t orc_site(const orc_lattice* L, int ix, int iy, int iz)
{
    return ((int64_t)ix * L->ly + iy) * L->lz + iz;
}

static inline void decode(const orc_lattice* L, int64_t s, int c[3])
{
    c[2] = (int)(s % L->lz); s /= L->lz;
    c[1] = (int)(s % L->ly); s /= L->ly;
    c[0] = (int)s;
}

/* Neighbour of `site` along axis (0,1,2) in direction dir (+1/-1); -1 when the
 * bond crosses an open boundary. Periodic wrap as src/chainMesh.f90:625-627. */
int64_t orc_neighbour(const orc_lattice* L, int64_t site, int axis, int dir)
{
    int c[3];
    const int n[3] = { L->lx, L->ly, L->lz };
    decode(L, site, c);
    c[axis] += dir;
    if (c[axis] < 0 || c[axis] >= n[axis]) {
        if (!L->periodic[axis]) return -1;
        c[axis] = (c[axis] + n[axis]) % n[axis];
    }
    return orc_site(L, c[0], c[1], c[2]);
}

/* bcc: site ((ix*ly + iy)*lz + iz)*2 + atom; atom 1 sits at (1/2,1/2,1/2) of
 * its cell. */
int64_t orc_bcc_site(const orc_lattice* L, int ix, int iy, int iz, int atom)
{
    return (((int64_t)ix * L->ly + iy) * L->lz + iz) * 2 + atom;
}

static inline void bcc_decode(const orc_lattice* L, int64_t s, int c[3], int* atom)
{
    *atom = (int)(s & 1);
    s >>= 1;
    c[2] = (int)(s % L->lz); s /= L->lz;
    c[1] = (int)(s % L->ly); s /= L->ly;
    c[0] = (int)s;
}

/* The eight nearest neighbours of a bcc site, at distance sqrt(3)/2 cells:
 * atom 0 of cell c bonds to atom 1 of the cells c - (dx,dy,dz), dx,dy,dz in
 * {0,1}, along (1/2 - dx, 1/2 - dy, 1/2 - dz); atom 1 bonds to atom 0 of the
 * cells c + (dx,dy,dz) along (dx - 1/2, dy - 1/2, dz - 1/2). sgn[k] holds the
 * signs of bond k's vector (site -> neighbour). Bonds across open faces are
 * left out; returns the number of bonds. */
int orc_bcc_bonds(const orc_lattice* L, int64_t site, int64_t nb[8], int sgn[8][3])
{
    int c[3], a;
    bcc_decode(L, site, c, &a);
    const int n[3] = { L->lx, L->ly, L->lz };
    int k = 0;
    for (int b = 0; b < 8; ++b) {
        const int d[3] = { (b >> 2) & 1, (b >> 1) & 1, b & 1 };
        int cc[3], ok = 1;
        for (int ax = 0; ax < 3 && ok; ++ax) {
            cc[ax] = a == 0 ? c[ax] - d[ax] : c[ax] + d[ax];
            if (cc[ax] < 0 || cc[ax] >= n[ax]) {
                if (!L->periodic[ax]) ok = 0;
                else cc[ax] = (cc[ax] + n[ax]) % n[ax];
            }
        }
        if (!ok) continue;
        nb[k] = orc_bcc_site(L, cc[0], cc[1], cc[2], 1 - a);
        for (int ax = 0; ax < 3; ++ax) sgn[k][ax] = a == 0 ? 1 - 2 * d[ax] : 2 * d[ax] - 1;
        ++k;
    }
    return k;
}

/* ---- hyper-kagome cell ---------------------------------------------- */

/* hyper_kagome_unit_cell (src/unit_cells.f90:32-77): a = b = c = 6, right
 * angles, twelve atoms at these fractions of the cell edges (every value is
 * exact in binary). makeChainMesh places atom i of cell (ix, iy, iz) at
 * (ix, iy, iz) + frac_i in units of the edge (src/chainMesh.f90:873-880). */
#define HK_ATOMS 12
#define HK_MAXB 12
static const double hk_frac[HK_ATOMS][3] = {
    { -0.5, 0.0, 0.5 },   { -0.25, 0.75, 0.5 }, { -0.5, 0.75, 0.25 }, { -0.25, 0.25, 0.0 },
    { -0.5, 0.25, 0.75 }, { -0.25, 0.5, 0.75 }, { -0.75, 0.5, 0.25 }, { 0.0, 0.5, 0.5 },
    { 0.0, 0.25, 0.25 },  { -0.75, 0.75, 0.0 }, { 0.0, 0.0, 0.0 },    { -0.75, 0.0, 0.75 } };

typedef struct { int u; int shift[3]; double r[3]; } hk_bond;
static hk_bond hk_bonds[HK_ATOMS][HK_MAXB];
static int hk_nb[HK_ATOMS];
static int hk_col[HK_ATOMS], hk_ncol;

/* Proper colouring of the atom types' bond graph with the fewest colours:
 * types in order, each the smallest colour that no earlier bonded type has,
 * backtracking when none is left (the lexicographically first colouring). */
static int hk_colour_from(int t, int k, const int adj[HK_ATOMS][HK_ATOMS])
{
    if (t == HK_ATOMS) return 1;
    for (int c = 0; c < k; ++c) {
        int ok = 1;
        for (int u = 0; u < t && ok; ++u)
            if (adj[t][u] && hk_col[u] == c) ok = 0;
        if (!ok) continue;
        hk_col[t] = c;
        if (hk_colour_from(t + 1, k, adj)) return 1;
    }
    return 0;
}

/* First neighbour shell of every atom, as initAtomShells
 * (src/chainMesh.f90:495-591) forms it with makeChainMesh's threshold 0.1 a
 * (:783): every atom of the cells within two cells, its distance, sorted;
 * the shell is the nearest atom and each next one whose distance exceeds
 * the previous by at most the threshold. Bond vectors are the displacements
 * (site -> neighbour) divided by their length. */
__attribute__((constructor)) static void hk_build(void)
{
    const double a = 6.0, thr = 0.1 * a;
    enum { NC = HK_ATOMS * 125 };
    static double dist[NC];
    static int cand[NC][4];
    int adj[HK_ATOMS][HK_ATOMS];
    memset(adj, 0, sizeof adj);
    for (int t = 0; t < HK_ATOMS; ++t) {
        int n = 0;
        for (int u = 0; u < HK_ATOMS; ++u)
            for (int sx = -2; sx <= 2; ++sx)
                for (int sy = -2; sy <= 2; ++sy)
                    for (int sz = -2; sz <= 2; ++sz) {
                        if (u == t && sx == 0 && sy == 0 && sz == 0) continue;
                        const double d[3] = { (hk_frac[u][0] + sx - hk_frac[t][0]) * a,
                                              (hk_frac[u][1] + sy - hk_frac[t][1]) * a,
                                              (hk_frac[u][2] + sz - hk_frac[t][2]) * a };
                        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        /* insertion into the sorted list */
                        int j = n++;
                        while (j > 0 && dist[j - 1] > r) {
                            dist[j] = dist[j - 1];
                            memcpy(cand[j], cand[j - 1], sizeof cand[j]);
                            --j;
                        }
                        dist[j] = r;
                        cand[j][0] = u; cand[j][1] = sx; cand[j][2] = sy; cand[j][3] = sz;
                    }
        int m = 0;
        for (int j = 0; j < n && m < HK_MAXB; ++j) {
            if (j > 0 && fabs(dist[j] - dist[j - 1]) > thr) break;
            hk_bond* b = &hk_bonds[t][m++];
            b->u = cand[j][0];
            for (int k = 0; k < 3; ++k) {
                b->shift[k] = cand[j][1 + k];
                b->r[k] = (hk_frac[b->u][k] + b->shift[k] - hk_frac[t][k]) * a / dist[j];
            }
            adj[t][b->u] = adj[b->u][t] = 1;
        }
        hk_nb[t] = m;
    }
    for (hk_ncol = 1; hk_ncol <= HK_ATOMS; ++hk_ncol)
        if (hk_colour_from(0, hk_ncol, adj)) break;
}

int orc_atoms_per_cell(const orc_lattice* L)
{
    return L->cell == ORC_CELL_BCC ? 2 : L->cell == ORC_CELL_HYPERKAGOME ? HK_ATOMS : 1;
}

/* Position of atom `atom` in its cell, in units of the cell edge. */
void orc_atom_frac(const orc_lattice* L, int atom, double f[3])
{
    for (int k = 0; k < 3; ++k)
        f[k] = L->cell == ORC_CELL_HYPERKAGOME ? hk_frac[atom][k] : (L->cell == ORC_CELL_BCC && atom == 1 ? 0.5 : 0.0);
}

/* Site (cell*natoms + atom) of atom `atom` of cell (ix, iy, iz), the
 * reference order (src/chainMesh.f90:684,878). */
int64_t orc_cell_site(const orc_lattice* L, int ix, int iy, int iz, int atom)
{
    return (((int64_t)ix * L->ly + iy) * L->lz + iz) * orc_atoms_per_cell(L) + atom;
}

/* Nearest neighbours of a hyper-kagome site and the unit bond vectors
 * (site -> neighbour); bonds across open faces are left out. */
int orc_hk_bonds(const orc_lattice* L, int64_t site, int64_t* nb, double (*r)[3])
{
    const int t = (int)(site % HK_ATOMS);
    int64_t cs = site / HK_ATOMS;
    const int c[3] = { (int)(cs / ((int64_t)L->ly * L->lz)), (int)((cs / L->lz) % L->ly), (int)(cs % L->lz) };
    const int n[3] = { L->lx, L->ly, L->lz };
    int k = 0;
    for (int b = 0; b < hk_nb[t]; ++b) {
        const hk_bond* B = &hk_bonds[t][b];
        int cc[3], ok = 1;
        for (int ax = 0; ax < 3 && ok; ++ax) {
            cc[ax] = c[ax] + B->shift[ax];
            if (cc[ax] < 0 || cc[ax] >= n[ax]) {
                if (!L->periodic[ax]) ok = 0;
                else cc[ax] = ((cc[ax] % n[ax]) + n[ax]) % n[ax];
            }
        }
        if (!ok) continue;
        nb[k] = orc_cell_site(L, cc[0], cc[1], cc[2], B->u);
        for (int ax = 0; ax < 3; ++ax) r[k][ax] = B->r[ax];
        ++k;
    }
    return k;
}

int orc_num_colours(const orc_lattice* L) { return L->cell == ORC_CELL_HYPERKAGOME ? hk_ncol : 2; }

/* Colour of a site in the sweep: (ix+iy+iz) mod 2 on SC, the atom of the cell
 * on bcc (both nearest-neighbour graphs are bipartite), the colour of the
 * atom type on hyper-kagome (its triangles need three). */
int orc_site_colour(const orc_lattice* L, int64_t i)
{
    if (L->cell == ORC_CELL_BCC) return (int)(i & 1);
    if (L->cell == ORC_CELL_HYPERKAGOME) return hk_col[i % HK_ATOMS];
    int64_t iz = i % L->lz, r = i / L->lz;
    int64_t iy = r % L->ly, ix = r / L->ly;
    return (int)((ix + iy + iz) & 1);
}

/* DMI vector d_a of the bond i -> i + r_a.  Bulk: D r_a.  Reference
 * interfacial form D_ij = Dz (c x r_{j->i}) with r_{j->i} = -r_a
 * (src/energyMin.f90:114,135-137). */
void orc_dmi_vectors(const orc_ham* h, double d[3][3])
{
    memset(d, 0, sizeof(double) * 9);
    if (h->dmi_type == ORC_DMI_BULK) {
        d[0][0] = h->D; d[1][1] = h->D; d[2][2] = h->D;
    } else if (h->dmi_type == ORC_DMI_INTERFACIAL_REF) {
        d[0][1] = -h->D;   /* z x (-x) = -y */
        d[1][0] =  h->D;   /* z x (-y) = +x */
    }
}

static inline void cross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void load(const double* s, int64_t N, int64_t i, double v[3])
{
    v[0] = s[i]; v[1] = s[N + i]; v[2] = s[2 * N + i];
}
static inline int64_t nsites(const orc_lattice* L)
{
    return (int64_t)L->lx * L->ly * L->lz * orc_atoms_per_cell(L);
}
int64_t orc_nsites(const orc_lattice* L) { return nsites(L); }

/* ---- energies (build form) ----------------------------------------- */

/* h_i = J sum_j S_j - sum_a (S_{i+a} - S_{i-a}) x d_a + Bz z.  The
 * anisotropy term is not part of h (it depends on the site's own spin). */
/* bcc: h_i = J sum_b S_b - sum_b S_b x d_b + Bz z over the eight bonds, with
 * d_b = (1/sqrt 3) sum_a sgn_a(b) d_a for the unit bond vector sgn(b)/sqrt 3
 * (the per-axis vectors d_a of orc_dmi_vectors: bulk d_b = D r_b, interfacial
 * d_b = D (z x r_{j->i})). */
static void bcc_bond_dmi(const double d[3][3], const int sg[3], double db[3])
{
    const double inv = 1.0 / sqrt(3.0);
    for (int m = 0; m < 3; ++m) db[m] = inv * (sg[0] * d[0][m] + sg[1] * d[1][m] + sg[2] * d[2][m]);
}

static void bcc_local_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, double out[3])
{
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    int64_t nb[8];
    int sg[8][3];
    const int nbonds = orc_bcc_bonds(L, site, nb, sg);
    double sum[3] = { 0, 0, 0 }, dm[3] = { 0, 0, 0 };
    for (int k = 0; k < nbonds; ++k) {
        double sj[3], db[3], c[3];
        load(s, N, nb[k], sj);
        bcc_bond_dmi(d, sg[k], db);
        cross(sj, db, c);
        for (int m = 0; m < 3; ++m) { sum[m] += sj[m]; dm[m] += c[m]; }
    }
    for (int m = 0; m < 3; ++m) out[m] = h->J * sum[m] - dm[m];
    out[2] += h->Bz;
}

/* DMI vector of a bond with unit vector r: d_b = sum_a r_a d_a (bulk D r,
 * reference interfacial D (z x r_{j->i})), the form bcc_bond_dmi takes for
 * r = sgn/sqrt 3. */
static void bond_dmi(const double d[3][3], const double r[3], double db[3])
{
    for (int m = 0; m < 3; ++m) db[m] = r[0] * d[0][m] + r[1] * d[1][m] + r[2] * d[2][m];
}

/* hyper-kagome: h_i = J sum_b S_b - sum_b S_b x d_b + Bz z over the first
 * shell (four bonds). */
static void hk_local_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, double out[3])
{
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    int64_t nb[HK_MAXB];
    double r[HK_MAXB][3];
    const int nbonds = orc_hk_bonds(L, site, nb, r);
    double sum[3] = { 0, 0, 0 }, dm[3] = { 0, 0, 0 };
    for (int k = 0; k < nbonds; ++k) {
        double sj[3], db[3], c[3];
        load(s, N, nb[k], sj);
        bond_dmi(d, r[k], db);
        cross(sj, db, c);
        for (int m = 0; m < 3; ++m) { sum[m] += sj[m]; dm[m] += c[m]; }
    }
    for (int m = 0; m < 3; ++m) out[m] = h->J * sum[m] - dm[m];
    out[2] += h->Bz;
}

void orc_local_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, double out[3])
{
    if (L->cell == ORC_CELL_BCC) { bcc_local_field(L, h, s, site, out); return; }
    if (L->cell == ORC_CELL_HYPERKAGOME) { hk_local_field(L, h, s, site, out); return; }
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    double sum[3] = { 0, 0, 0 }, dm[3] = { 0, 0, 0 };
    for (int a = 0; a < 3; ++a) {
        double sp[3] = { 0, 0, 0 }, sm[3] = { 0, 0, 0 };
        int64_t jp = orc_neighbour(L, site, a, +1), jm = orc_neighbour(L, site, a, -1);
        if (jp >= 0) load(s, N, jp, sp);
        if (jm >= 0) load(s, N, jm, sm);
        double diff[3] = { sp[0] - sm[0], sp[1] - sm[1], sp[2] - sm[2] }, c[3];
        cross(diff, d[a], c);
        for (int k = 0; k < 3; ++k) { sum[k] += sp[k] + sm[k]; dm[k] += c[k]; }
    }
    for (int k = 0; k < 3; ++k) out[k] = h->J * sum[k] - dm[k];
    out[2] += h->Bz;
}

/* dE of replacing S_i by sn: -(sn - S).h - Kz (sn_z - S_z)(sn_z + S_z). */
double orc_delta_energy(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, const double sn[3])
{
    const int64_t N = nsites(L);
    double hf[3], so[3];
    orc_local_field(L, h, s, site, hf);
    load(s, N, site, so);
    double del[3] = { sn[0] - so[0], sn[1] - so[1], sn[2] - so[2] };
    return -dot(del, hf) - h->Kz * del[2] * (sn[2] + so[2]);
}

/* bcc: every bond once, from its atom-0 end: -J S_i.S_j + d_b.(S_i x S_j),
 * plus the on-site terms of every site. */
static double bcc_total_energy(const orc_lattice* L, const orc_ham* h, const double* s)
{
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    double e = 0.0;
    #pragma omp parallel for reduction(+:e) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double si[3], sj[3], c[3], db[3];
        load(s, N, i, si);
        double ei = 0.0;
        if ((i & 1) == 0) {
            int64_t nb[8];
            int sg[8][3];
            const int nbonds = orc_bcc_bonds(L, i, nb, sg);
            for (int k = 0; k < nbonds; ++k) {
                load(s, N, nb[k], sj);
                cross(si, sj, c);
                bcc_bond_dmi(d, sg[k], db);
                ei += -h->J * dot(si, sj) + dot(db, c);
            }
        }
        ei -= h->zeeman_total_scale * h->Bz * si[2] + h->Kz * si[2] * si[2];
        e += ei;
    }
    return e;
}

/* hyper-kagome: every bond once, from the end whose atom comes first in the
 * cell (bonds never join two atoms of one type), plus the on-site terms. */
static double hk_total_energy(const orc_lattice* L, const orc_ham* h, const double* s)
{
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    double e = 0.0;
    #pragma omp parallel for reduction(+:e) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double si[3], sj[3], c[3], db[3];
        load(s, N, i, si);
        double ei = 0.0;
        int64_t nb[HK_MAXB];
        double r[HK_MAXB][3];
        const int nbonds = orc_hk_bonds(L, i, nb, r);
        for (int k = 0; k < nbonds; ++k) {
            if (nb[k] % HK_ATOMS < i % HK_ATOMS) continue;
            load(s, N, nb[k], sj);
            cross(si, sj, c);
            bond_dmi(d, r[k], db);
            ei += -h->J * dot(si, sj) + dot(db, c);
        }
        ei -= h->zeeman_total_scale * h->Bz * si[2] + h->Kz * si[2] * si[2];
        e += ei;
    }
    return e;
}

double orc_total_energy(const orc_lattice* L, const orc_ham* h, const double* s)
{
    if (L->cell == ORC_CELL_BCC) return bcc_total_energy(L, h, s);
    if (L->cell == ORC_CELL_HYPERKAGOME) return hk_total_energy(L, h, s);
    const int64_t N = nsites(L);
    double d[3][3];
    orc_dmi_vectors(h, d);
    double e = 0.0;
    #pragma omp parallel for reduction(+:e) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double si[3], sj[3], c[3];
        load(s, N, i, si);
        double ei = 0.0;
        for (int a = 0; a < 3; ++a) {
            int64_t j = orc_neighbour(L, i, a, +1);
            if (j < 0) continue;
            load(s, N, j, sj);
            cross(si, sj, c);
            ei += -h->J * dot(si, sj) + dot(d[a], c);
        }
        ei -= h->zeeman_total_scale * h->Bz * si[2] + h->Kz * si[2] * si[2];
        e += ei;
    }
    return e;
}

void orc_magnetisation(const orc_lattice* L, const double* s, double m[3])
{
    const int64_t N = nsites(L);
    double mx = 0, my = 0, mz = 0;
    for (int64_t i = 0; i < N; ++i) { mx += s[i]; my += s[N + i]; mz += s[2 * N + i]; }
    m[0] = mx / (double)N; m[1] = my / (double)N; m[2] = mz / (double)N;
}

/* Corners (i,j), (i+1,j), (i,j+1), (i+1,j+1) of the plaquette of layer z:
 * one atom of each of those cells, the cell's firstAtomInMeshCell. makeChainMesh
 * adds the atoms of a cell in order and addAtomToChainCell puts each new one
 * at the head of the cell's list (src/chainMesh.f90:663-671,873-881), so that
 * atom is the LAST of the unit cell: the body-centre atom on bcc, atom 12 on
 * hyper-kagome. derivativeList then steps to the same atom of the next cell
 * (src/chainMesh.f90:931-1031; no other atom lies on its x or y line).
 * Returns 0 when a corner lies across an open face. ids = {s1, s2, s3, s4}
 * with s2 = +x, s3 = +y, s4 = +x+y. */
static int plaquette(const orc_lattice* L, int i, int j, int z, int64_t ids[4])
{
    int i2 = i + 1, j2 = j + 1;
    if (i2 == L->lx) { if (!L->periodic[0]) return 0; i2 = 0; }
    if (j2 == L->ly) { if (!L->periodic[1]) return 0; j2 = 0; }
    const int a = orc_atoms_per_cell(L) - 1;
    ids[0] = orc_cell_site(L, i, j, z, a); ids[1] = orc_cell_site(L, i2, j, z, a);
    ids[2] = orc_cell_site(L, i, j2, z, a); ids[3] = orc_cell_site(L, i2, j2, z, a);
    return 1;
}

int orc_plaquette(const orc_lattice* L, int i, int j, int z, int64_t ids[4]) { return plaquette(L, i, j, z, ids); }

/* ---- reference restatements ----------------------------------------- */

/* Minimum-image displacement p2 - p1 on a periodic SC lattice of unit cell
 * length 1, src/chainMesh.f90:206-241 (coefficients in the Bravais basis,
 * then c - width*nint(c/width); nint rounds half away from zero). */
static void ref_min_image(const orc_lattice* L, const double p1[3], const double p2[3], double d[3])
{
    const double w[3] = { (double)L->lx, (double)L->ly, (double)L->lz };
    for (int k = 0; k < 3; ++k) {
        double c = p2[k] - p1[k];
        d[k] = c - w[k] * round(c / w[k]);
    }
}

/* Atom position in cell units: the cell corner plus the atom's fraction
 * (src/unit_cells.f90; makeChainMesh, src/chainMesh.f90:873-880). */
static void ref_position(const orc_lattice* L, int64_t site, double p[3])
{
    const int na = orc_atoms_per_cell(L);
    int c[3];
    double f[3];
    decode(L, site / na, c);
    orc_atom_frac(L, (int)(site % na), f);
    for (int k = 0; k < 3; ++k) p[k] = c[k] + f[k];
}

/* Nearest-neighbour list of a site of a periodic lattice, as the reference's
 * atomShells(i,1) holds it (src/chainMesh.f90:495-591): 6 on SC (-x, +x,
 * -y, +y, -z, +z), 8 on bcc, 4 on hyper-kagome. */
static int ref_nn(const orc_lattice* L, int64_t site, int64_t nb[HK_MAXB])
{
    if (L->cell == ORC_CELL_BCC) {
        int sg[8][3];
        return orc_bcc_bonds(L, site, nb, sg);
    }
    if (L->cell == ORC_CELL_HYPERKAGOME) {
        double r[HK_MAXB][3];
        return orc_hk_bonds(L, site, nb, r);
    }
    int k = 0;
    for (int a = 0; a < 3; ++a)
        for (int dir = -1; dir <= 1; dir += 2) nb[k++] = orc_neighbour(L, site, a, dir);
    return k;
}

/* calculateHeisenbergEnergy (src/energyMin.f90:86-171) for a periodic SC,
 * bcc or hyper-kagome lattice with J' = 0 and demag off: old/new local energy of site i for
 * proposal sp. For each nearest neighbour j: r = -(x_j - x_i)_min/|.|,
 * D = Dz (c x r), energy += J S.S_j + D.(S x S_j); then -B S_z. */
void orc_ref_local_energy(const orc_lattice* L, double J_r, double Dz, double B, const double* s,
                          int64_t site, const double sp[3], double* old_e, double* new_e)
{
    const int64_t N = nsites(L);
    const double chat[3] = { 0.0, 0.0, 1.0 };
    double S[3], p1[3];
    load(s, N, site, S);
    ref_position(L, site, p1);
    double tv[3] = { Dz * chat[0], Dz * chat[1], Dz * chat[2] };
    double oe = 0.0, ne = 0.0;
    int64_t nb[HK_MAXB];
    const int nn = ref_nn(L, site, nb);
    for (int k = 0; k < nn; ++k) {
        const int64_t j = nb[k];
        double Sj[3], p2[3], r[3], D[3], x1[3], x2[3];
        load(s, N, j, Sj);
        ref_position(L, j, p2);
        ref_min_image(L, p1, p2, r);
        double nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int m = 0; m < 3; ++m) r[m] = (-1.0) * r[m] / nr;
        cross(tv, r, D);
        cross(S, Sj, x1);
        cross(sp, Sj, x2);
        oe = oe + (J_r * dot(S, Sj) + dot(D, x1));
        ne = ne + (J_r * dot(sp, Sj) + dot(D, x2));
    }
    *old_e = oe - B * S[2];
    *new_e = ne - B * sp[2];
}

/* TotalHeisenbergEnergy + calculateTotalHeisenbergEnergyHelper
 * (src/energyMin.f90:499-528, 174-229), demag off: sum over atoms of
 * 0.5 * sum_j (J S.S_j + D.(S x S_j)) - g*muB*B*S_z with g*muB = 2
 * (src/constants.f90:11,14) and D = [0,0,Dz] x r. */
double orc_ref_total_energy(const orc_lattice* L, double J_r, double Dz, double B, const double* s)
{
    const int64_t N = nsites(L);
    double total = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double S[3], p1[3], e = 0.0;
        const double tv[3] = { 0.0, 0.0, Dz };
        load(s, N, i, S);
        ref_position(L, i, p1);
        int64_t nb[HK_MAXB];
        const int nn = ref_nn(L, i, nb);
        for (int k = 0; k < nn; ++k) {
            const int64_t j = nb[k];
            double Sj[3], p2[3], r[3], D[3], x1[3];
            load(s, N, j, Sj);
            ref_position(L, j, p2);
            ref_min_image(L, p1, p2, r);
            double nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            for (int m = 0; m < 3; ++m) r[m] = (-1.0) * r[m] / nr;
            cross(tv, r, D);
            cross(S, Sj, x1);
            e = e + 0.5 * (J_r * dot(S, Sj) + dot(D, x1));
        }
        e = e - (2.0 * 1.0 * B * S[2]);
        total = total + e;
    }
    return total;
}

/* Scalar product as src/vecNd.f90:413-422: accumulate from 0 in index order. */
static inline double vdot(const double a[3], const double b[3])
{
    double r = 0.0;
    r = r + a[0] * b[0];
    r = r + a[1] * b[1];
    r = r + a[2] * b[2];
    return r;
}

/* arc_winding (src/reciprocalSpaceProcesses.f90:284-318). The imaginary part
 * is built with cmplx() without a kind argument, which yields a default
 * (single precision) complex: the triple product is rounded to float. */
double orc_ref_arc_winding(const double s1[3], const double s2[3], const double s3[3])
{
    double c[3];
    cross(s2, s3, c);
    double re = 1.0 + vdot(s1, s2) + vdot(s2, s3) + vdot(s3, s1);
    double im = (double)(float)vdot(s1, c);
    re = re + 0.0;
    double norm = sqrt(2.0 * (1.0 + vdot(s1, s2)) * (1.0 + vdot(s2, s3)) * (1.0 + vdot(s3, s1)));
    re = re / norm;
    im = im / norm;
    return 2.0 * atan2(im, re);
}

/* calculate_winding_number2 (src/reciprocalSpaceProcesses.f90:320-364):
 * triangles (s1,s2,s4) and (s1,s4,s3) per plaquette with s2 = +x, s3 = +y,
 * s4 = +x+y, spins renormalised in FP64, sum divided by 4*pi where pi is the
 * single-precision literal of src/constants.f90:16. */
double orc_ref_winding_number(const orc_lattice* L, const double* s, int zi)
{
    const int64_t N = nsites(L);
    const double pi_ref = (double)3.14159265358979323846f;
    double wn = 0.0;
    for (int i = 0; i < L->lx; ++i) {
        for (int j = 0; j < L->ly; ++j) {
            double v[4][3];
            int64_t ids[4];
            if (!plaquette(L, i, j, zi - 1, ids)) continue;
            for (int k = 0; k < 4; ++k) {
                load(s, N, ids[k], v[k]);
                double n = 0.0;
                n = n + v[k][0] * v[k][0];
                n = n + v[k][1] * v[k][1];
                n = n + v[k][2] * v[k][2];
                n = sqrt(n);
                v[k][0] /= n; v[k][1] /= n; v[k][2] /= n;
            }
            double s1 = orc_ref_arc_winding(v[0], v[1], v[3]);
            double s2 = orc_ref_arc_winding(v[0], v[3], v[2]);
            wn = wn + s1 + s2;
        }
    }
    return wn / (4.0 * pi_ref);
}

/* calculate_winding_number_density (src/reciprocalSpaceProcesses.f90:438-500):
 * per plaquette (i, j) of layer zi (1-based), the same two triangles as
 * calculate_winding_number2, density(i, j) = (sigma1 + sigma2) / (4 pi) with
 * the single-precision pi; out[j*lx + i] (x fastest, the column-major order of
 * density_matrix(numCellsX, numCellsY)). */
void orc_ref_winding_density(const orc_lattice* L, const double* s, int zi, double* out)
{
    const int64_t N = nsites(L);
    const double pi_ref = (double)3.14159265358979323846f;
    for (int i = 0; i < L->lx; ++i) {
        for (int j = 0; j < L->ly; ++j) {
            double v[4][3];
            int64_t ids[4];
            if (!plaquette(L, i, j, zi - 1, ids)) continue;
            for (int k = 0; k < 4; ++k) {
                load(s, N, ids[k], v[k]);
                double n = 0.0;
                n = n + v[k][0] * v[k][0];
                n = n + v[k][1] * v[k][1];
                n = n + v[k][2] * v[k][2];
                n = sqrt(n);
                v[k][0] /= n; v[k][1] /= n; v[k][2] /= n;
            }
            out[(int64_t)j * L->lx + i] = (orc_ref_arc_winding(v[0], v[1], v[3]) + orc_ref_arc_winding(v[0], v[3], v[2]))
                                          / (4.0 * pi_ref);
        }
    }
}

/* ---- topological charge, build form --------------------------------- */

/* Solid angle of one triangle from FP32 spins, every operation in FP64 in a
 * fixed order (the GPU kernel evaluates the identical sequence):
 * renormalise a, b, c; Re = ((1 + a.b) + b.c) + c.a; Im = a.(b x c);
 * sigma = 2 atan2(Im, Re). The reference's positive normalisation factor
 * does not change atan2 and is omitted (it is NaN for antiparallel bonds). */
static inline void renorm(const float in[3], double o[3])
{
    double x = (double)in[0], y = (double)in[1], z = (double)in[2];
    double n = sqrt(x * x + y * y + z * z);
    o[0] = x / n; o[1] = y / n; o[2] = z / n;
}

double orc_triangle_sigma(const float af[3], const float bf[3], const float cf[3])
{
    double a[3], b[3], c[3];
    renorm(af, a); renorm(bf, b); renorm(cf, c);
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    double x0 = b[1] * c[2] - b[2] * c[1];
    double x1 = b[2] * c[0] - b[0] * c[2];
    double x2 = b[0] * c[1] - b[1] * c[0];
    double re = ((1.0 + ab) + bc) + ca;
    double im = a[0] * x0 + a[1] * x1 + a[2] * x2;
    return 2.0 * atan2(im, re);
}

/* Per-layer Q (z 0-based). Plaquettes whose +x or +y corner lies across an
 * open boundary are skipped. q_int = nearest integer of the FP64 sum. */
double orc_winding_layer(const orc_lattice* L, const double* s, int z, int64_t* q_int)
{
    const int64_t N = nsites(L);
    double acc = 0.0;
    for (int i = 0; i < L->lx; ++i) {
        for (int j = 0; j < L->ly; ++j) {
            float v[4][3];
            int64_t ids[4];
            if (!plaquette(L, i, j, z, ids)) continue;
            for (int k = 0; k < 4; ++k)
                for (int m = 0; m < 3; ++m) v[k][m] = (float)s[m * N + ids[k]];
            acc += orc_triangle_sigma(v[0], v[1], v[3]) + orc_triangle_sigma(v[0], v[3], v[2]);
        }
    }
    double q = acc / (4.0 * M_PI);
    if (q_int) *q_int = (int64_t)llround(q);
    return q;
}

/* Per-plaquette topological charge density of layer z (0-based), build form:
 * out[j*lx + i] = (sigma(s1,s2,s4) + sigma(s1,s4,s3)) / (4 pi), the terms of
 * orc_winding_layer; 0 for plaquettes skipped at open boundaries. */
void orc_winding_density(const orc_lattice* L, const double* s, int z, double* out)
{
    const int64_t N = nsites(L);
    for (int i = 0; i < L->lx; ++i) {
        for (int j = 0; j < L->ly; ++j) {
            double d = 0.0;
            int64_t ids[4];
            if (plaquette(L, i, j, z, ids)) {
                float v[4][3];
                for (int k = 0; k < 4; ++k)
                    for (int m = 0; m < 3; ++m) v[k][m] = (float)s[m * N + ids[k]];
                d = (orc_triangle_sigma(v[0], v[1], v[3]) + orc_triangle_sigma(v[0], v[3], v[2])) / (4.0 * M_PI);
            }
            out[(int64_t)j * L->lx + i] = d;
        }
    }
}

/* ---- Metropolis ------------------------------------------------------ */

/* Cone (ball) proposal, src/energyMin.f90:49-68 and :326-329: u uniform in a
 * ball of radius `radius` (direction cos(theta) = 2U-1, phi = 2 pi U',
 * radius * U''^(1/3)), S' = (S + u)/|S + u|. rnd[3] becomes the acceptance
 * uniform. Returns 0 when |S + u| vanishes (proposal discarded). */
int orc_propose(const double so[3], double radius, const uint32_t rnd[4], double sn[3], double* u_acc)
{
    double ct = s24(rnd[0]);
    double uphi = u24(rnd[1]);
    double urad = u24(rnd[2]);
    *u_acc = u24(rnd[3]);
    double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    double phi = 2.0 * M_PI * uphi;
    double rr = radius * cbrt(urad);
    double t[3] = { so[0] + rr * st * cos(phi), so[1] + rr * st * sin(phi), so[2] + rr * ct };
    double n2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    if (!(n2 > 1e-30)) { sn[0] = so[0]; sn[1] = so[1]; sn[2] = so[2]; return 0; }
    double inv = 1.0 / sqrt(n2);
    sn[0] = t[0] * inv; sn[1] = t[1] * inv; sn[2] = t[2] * inv;
    return 1;
}

/* One colour of the exact sweep: every site of that colour (orc_site_colour)
 * gets one proposal, accepted iff u < exp(-beta dE). Same-colour sites share
 * no bond, so the order inside a colour is free. H (may be NULL) is an
 * external per-site field entering dE as -mu0 (S' - S).H_i, the reference's
 * demagnetisation term (src/energyMin.f90:113-116,163-166). */
static int64_t half_sweep_field(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                                uint64_t seed, uint32_t slot, uint32_t sweep, int colour, const double* H, double mu0)
{
    const int64_t N = nsites(L);
    int64_t acc = 0;
    #pragma omp parallel for reduction(+:acc) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        if (orc_site_colour(L, i) != colour) continue;
        uint32_t rnd[4];
        orc_metropolis_random(seed, slot, (uint32_t)i, sweep, rnd);
        double so[3], sn[3], u;
        load(s, N, i, so);
        if (!orc_propose(so, radius, rnd, sn, &u)) continue;
        double de = H ? orc_delta_energy_field(L, h, s, i, sn, H, mu0) : orc_delta_energy(L, h, s, i, sn);
        if (u < exp(-beta * de)) {
            s[i] = sn[0]; s[N + i] = sn[1]; s[2 * N + i] = sn[2];
            acc++;
        }
    }
    return acc;
}

int64_t orc_half_sweep(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                       uint64_t seed, uint32_t slot, uint32_t sweep, int colour)
{
    return half_sweep_field(L, h, s, beta, radius, seed, slot, sweep, colour, NULL, 0.0);
}

int64_t orc_sweeps(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                   uint64_t seed, uint32_t slot, uint32_t sweep0, int n_sweeps)
{
    int64_t acc = 0;
    for (int k = 0; k < n_sweeps; ++k)
        for (int c = 0; c < orc_num_colours(L); ++c)
            acc += orc_half_sweep(L, h, s, beta, radius, seed, slot, sweep0 + (uint32_t)k, c);
    return acc;
}

double orc_delta_energy_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site,
                              const double sn[3], const double* H, double mu0)
{
    const int64_t N = nsites(L);
    double so[3], hd[3];
    load(s, N, site, so);
    load(H, N, site, hd);
    const double del[3] = { sn[0] - so[0], sn[1] - so[1], sn[2] - so[2] };
    return orc_delta_energy(L, h, s, site, sn) - mu0 * dot(del, hd);
}

/* Metropolis_mcs with demagnetisation (src/energyMin.f90:258-259,362-369):
 * the field is computed from the spins before the first sweep and again
 * after every sweep whose 1-based number is a multiple of `interval`
 * (interval <= 0: only before the first); H receives the last field. */
int64_t orc_sweeps_demag(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                         uint64_t seed, uint32_t slot, uint32_t sweep0, int n_sweeps, double mu0, int interval,
                         double* H)
{
    if (orc_ref_demag_field(L, s, H)) return -1;
    int64_t acc = 0;
    for (int k = 0; k < n_sweeps; ++k) {
        for (int c = 0; c < orc_num_colours(L); ++c)
            acc += half_sweep_field(L, h, s, beta, radius, seed, slot, sweep0 + (uint32_t)k, c, H, mu0);
        if (interval > 0 && (k + 1) % interval == 0) orc_ref_demag_field(L, s, H);
    }
    return acc;
}

void orc_init_random(const orc_lattice* L, double* s, uint64_t seed, uint32_t slot)
{
    const int64_t N = nsites(L);
    for (int64_t i = 0; i < N; ++i) {
        uint32_t ctr[4] = { (uint32_t)i, 0xFFFFFFFFu, ORC_STREAM_INIT, (uint32_t)(seed >> 32) };
        uint32_t key[2] = { (uint32_t)seed, slot }, r[4];
        orc_philox4x32_10(ctr, key, r);
        double z = s24(r[0]);
        double phi = 2.0 * M_PI * u24(r[1]);
        double st = sqrt(fmax(0.0, (1.0 - z) * (1.0 + z)));
        s[i] = st * cos(phi); s[N + i] = st * sin(phi); s[2 * N + i] = z;
    }
}

/* Ferromagnet along +z with site i reversed iff U < frac (init stream, word 2). */
void orc_init_ferro(const orc_lattice* L, double* s, uint64_t seed, uint32_t slot, double frac)
{
    const int64_t N = nsites(L);
    for (int64_t i = 0; i < N; ++i) {
        uint32_t ctr[4] = { (uint32_t)i, 0xFFFFFFFFu, ORC_STREAM_INIT, (uint32_t)(seed >> 32) };
        uint32_t key[2] = { (uint32_t)seed, slot }, r[4];
        orc_philox4x32_10(ctr, key, r);
        s[i] = 0.0; s[N + i] = 0.0;
        s[2 * N + i] = (u24(r[2]) < frac) ? -1.0 : 1.0;
    }
}

/* ---- parallel tempering -------------------------------------------- */

/* 53-bit uniform in (0,1] for swap decisions: key = (seed_lo, ladder),
 * ctr = (pair, round, swap stream, seed_hi). */
double orc_swap_uniform(uint64_t seed, uint32_t ladder, uint32_t pair, uint32_t round)
{
    uint32_t ctr[4] = { pair, round, ORC_STREAM_SWAP, (uint32_t)(seed >> 32) };
    uint32_t key[2] = { (uint32_t)seed, ladder }, r[4];
    orc_philox4x32_10(ctr, key, r);
    uint64_t m = ((uint64_t)(r[0] >> 5) << 26) | (uint64_t)(r[1] >> 6);
    return ((double)m + 1.0) * 0x1p-53;
}

/* One replica-exchange round (the step of src/PT.f90:234-272 with the
 * textbook exponent): pairs (t, t+1) with t = round mod 2, 2 apart; accept iff
 * log u < (beta_t - beta_{t+1}) (E_t - E_{t+1}); swap which slot sits at each
 * temperature (labels move, spins stay, as src/PT.f90:264-268). */
int orc_pt_swap_round(int n, const double* beta_of_temp, const double* e_of_slot, int32_t* slot_of_temp,
                      uint64_t seed, uint32_t ladder, uint32_t round, int64_t* attempts, int64_t* accepts)
{
    int nacc = 0;
    for (int t = (int)(round & 1u); t + 1 < n; t += 2) {
        int32_t a = slot_of_temp[t], b = slot_of_temp[t + 1];
        double delta = (beta_of_temp[t] - beta_of_temp[t + 1]) * (e_of_slot[a] - e_of_slot[b]);
        double u = orc_swap_uniform(seed, ladder, (uint32_t)t, round);
        if (attempts) attempts[t]++;
        if (log(u) < delta) {
            slot_of_temp[t] = b; slot_of_temp[t + 1] = a;
            if (accepts) accepts[t]++;
            nacc++;
        }
    }
    return nacc;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
