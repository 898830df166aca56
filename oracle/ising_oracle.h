/*
 * ising_oracle.h -- FP64 CPU restatement of the reference's Metropolis /
 * parallel-tempering hot path (SamuelJosephs/ISING), used ONLY as test
 * infrastructure: by tests/, by __graft_entry__.smoke() to check its result,
 * and by bench.py's cpu_baseline leg. The product (libising_hip.so and the
 * ising_amd package) never links or calls it.
 *
 * Two families of functions:
 *   orc_ref_*  follow the reference's own code path line by line (positions,
 *              minimum-image bond vectors, cross products, its RNG);
 *   orc_*      state the same physics in the form the GPU path uses
 *              (checkerboard sweep, Philox counters, per-axis DMI vectors),
 *              in FP64, so GPU results can be compared to them.
 *
 * Spin arrays: double s[3*N], column-major (all x, then y, then z), sites in
 * the reference order site = (ix*Ly + iy)*Lz + iz (src/chainMesh.f90:673-685).
 *
 * Pinning (which check fixes each function) is listed per function in
 * ising_oracle.c and in the test files under tests/.
 */
#ifndef ISING_ORACLE_H
#define ISING_ORACLE_H
#include <stdint.h>

/* cell: ORC_CELL_SC, one atom per cell; ORC_CELL_BCC, atoms at (0,0,0) and
 * (1/2,1/2,1/2) of each cell (src/unit_cells.f90:7-30); ORC_CELL_HYPERKAGOME,
 * the twelve atoms of src/unit_cells.f90:32-77. Site index
 * ((ix*Ly + iy)*Lz + iz)*natoms + atom, the reference order for multi-atom
 * cells (src/chainMesh.f90:684,878). */
typedef struct { int32_t lx, ly, lz; int32_t periodic[3]; int32_t cell; } orc_lattice;
#define ORC_CELL_SC 0
#define ORC_CELL_BCC 1
#define ORC_CELL_HYPERKAGOME 2
typedef struct {
    double J, D; int32_t dmi_type; int32_t reserved;
    double Bz, Kz, zeeman_total_scale;
} orc_ham;

#define ORC_DMI_NONE 0
#define ORC_DMI_BULK 1
#define ORC_DMI_INTERFACIAL_REF 2

/* Philox counter streams (word 2 of the counter). */
#define ORC_STREAM_METROPOLIS 0u
#define ORC_STREAM_INIT       1u
#define ORC_STREAM_SWAP       2u

/* ---- RNG ---- */
void    orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void    orc_metropolis_random(uint64_t seed, uint32_t slot, uint32_t site, uint32_t sweep, uint32_t out[4]);
typedef struct { int32_t seed, ix, iy; } orc_ref_rng;
orc_ref_rng orc_ref_make_random(int32_t seed);
double  orc_ref_uniform(orc_ref_rng* r);

/* ---- geometry ---- */
int64_t orc_site(const orc_lattice* L, int ix, int iy, int iz);
int64_t orc_neighbour(const orc_lattice* L, int64_t site, int axis, int dir);
int64_t orc_bcc_site(const orc_lattice* L, int ix, int iy, int iz, int atom);
int     orc_bcc_bonds(const orc_lattice* L, int64_t site, int64_t nb[8], int sgn[8][3]);
int     orc_atoms_per_cell(const orc_lattice* L);
void    orc_atom_frac(const orc_lattice* L, int atom, double f[3]);
int64_t orc_cell_site(const orc_lattice* L, int ix, int iy, int iz, int atom);
int     orc_hk_bonds(const orc_lattice* L, int64_t site, int64_t* nb, double (*r)[3]);
int     orc_num_colours(const orc_lattice* L);
int     orc_site_colour(const orc_lattice* L, int64_t site);
int     orc_plaquette(const orc_lattice* L, int i, int j, int z, int64_t ids[4]);
int64_t orc_nsites(const orc_lattice* L);
void    orc_dmi_vectors(const orc_ham* h, double d[3][3]);

/* ---- energies (build form) ---- */
void    orc_local_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, double out[3]);
double  orc_delta_energy(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site, const double sn[3]);
double  orc_total_energy(const orc_lattice* L, const orc_ham* h, const double* s);
void    orc_magnetisation(const orc_lattice* L, const double* s, double m[3]);

/* ---- reference restatements ---- */
void    orc_ref_local_energy(const orc_lattice* L, double J_r, double Dz, double B, const double* s,
                             int64_t site, const double sp[3], double* old_e, double* new_e);
double  orc_ref_total_energy(const orc_lattice* L, double J_r, double Dz, double B, const double* s);
double  orc_ref_arc_winding(const double s1[3], const double s2[3], const double s3[3]);
double  orc_ref_winding_number(const orc_lattice* L, const double* s, int z_index_1based);

/* ---- topological charge (build form, same op order as the GPU) ---- */
double  orc_triangle_sigma(const float a[3], const float b[3], const float c[3]);
double  orc_winding_layer(const orc_lattice* L, const double* s, int z, int64_t* q_int);
void    orc_winding_density(const orc_lattice* L, const double* s, int z, double* out);
void    orc_ref_winding_density(const orc_lattice* L, const double* s, int z_index_1based, double* out);

/* ---- Metropolis ---- */
int     orc_propose(const double sold[3], double radius, const uint32_t rnd[4], double snew[3], double* u_acc);
int64_t orc_half_sweep(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                       uint64_t seed, uint32_t slot, uint32_t sweep, int colour);
int64_t orc_sweeps(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                   uint64_t seed, uint32_t slot, uint32_t sweep0, int n_sweeps);
int64_t orc_sweeps_learned(const orc_lattice* L, const orc_ham* h, double* s, double beta, const double* params,
                           int quantize, double log_rmin, double log_rmax, uint64_t seed, uint32_t slot,
                           uint32_t sweep0, int n_sweeps);       /* mlp_oracle.c */
void    orc_init_random(const orc_lattice* L, double* s, uint64_t seed, uint32_t slot);
void    orc_init_ferro(const orc_lattice* L, double* s, uint64_t seed, uint32_t slot, double frac);

/* ---- demagnetisation (demag_oracle.c) ----
 * H: double[3*N] in the spin layout (the reference's demagnetisation_array);
 * mesh arrays: double[3*Nc], cell (i*Ly + j)*Lz + k. mu0 is the reference's
 * constants::mu_0 (a single-precision literal, 1.256637006008532e-6). */
#define ORC_MU0_REF 1.256637006008532e-06
int     orc_ref_demag_field(const orc_lattice* L, const double* s, double* H);   /* -1: not fully periodic */
void    orc_demag_deposit(const orc_lattice* L, const double* s, double* mesh);
void    orc_demag_mesh_field(const orc_lattice* L, const double* mesh, double* hmesh);
void    orc_demag_gather(const orc_lattice* L, const double* hmesh, double* H);
double  orc_demag_energy(const orc_lattice* L, const double* s, const double* H, double mu0);
double  orc_delta_energy_field(const orc_lattice* L, const orc_ham* h, const double* s, int64_t site,
                               const double sn[3], const double* H, double mu0);
int64_t orc_sweeps_demag(const orc_lattice* L, const orc_ham* h, double* s, double beta, double radius,
                         uint64_t seed, uint32_t slot, uint32_t sweep0, int n_sweeps, double mu0, int interval,
                         double* H);

/* ---- skyrmion count of a layer (skyrmion_oracle.c) ----
 * Coarse grid S[3][ny][nx]; fine grids and densities [NY][NX] with
 * NX = sx nx, NY = sy ny; count tables [T][S][n_part]. */
void    orc_sk_kernel_table(int n, int s, double* K);
void    orc_sk_coarse(const orc_lattice* L, const double* s, int z, double* S);
void    orc_sk_interpolate(int nx, int ny, int sx, int sy, const double* S, double* F);
void    orc_sk_density(int NX, int NY, const double* F, double* d);
void    orc_sk_gauss_weights(double sigma, double G[25]);
void    orc_sk_filter(int NX, int NY, const double* d, double sigma, double* f);
void    orc_sk_count_table(int NX, int NY, const double* d, int n_part, int T, int S, double min_t, double max_t,
                           int32_t* counts);
void    orc_sk_select(const int32_t* counts, int n_part, int T, int S, double q, int32_t* out);
int     orc_sk_distribution(const orc_lattice* L, const double* s, int z, int sx, int sy, int n_part, double min_t,
                            double max_t, int T, int S, int32_t* out, int32_t* table, double* d_out);

/* ---- parallel tempering ---- */
double  orc_swap_uniform(uint64_t seed, uint32_t ladder, uint32_t pair, uint32_t round);
int     orc_pt_swap_round(int n_temps, const double* beta_of_temp, const double* e_of_slot,
                          int32_t* slot_of_temp, uint64_t seed, uint32_t ladder, uint32_t round,
                          int64_t* attempts, int64_t* accepts);

/* ---- threads ---- */
int     orc_num_threads(void);
void    orc_set_threads(int n);

#endif
