/*
 * ising_gpu.h -- C ABI of libising_hip.so, the MI355X (gfx950) Metropolis +
 * parallel-tempering engine for classical Heisenberg spins on simple-cubic,
 * body-centred-cubic and hyper-kagome lattices (exchange + DMI + Zeeman +
 * uniaxial anisotropy + the reference's demagnetisation term).
 *
 * Plain C types only; no exceptions cross this boundary. Every function
 * returns ISING_OK (0) on success or a negative ISING_ERR_* code; the text of
 * the last error of a context is available from ising_last_error().
 *
 * Ownership: the library owns all device memory of a context. Host buffers
 * passed in are caller-owned and are read/written synchronously before the
 * call returns. One host thread per context; contexts are independent.
 *
 * Layout of host spin buffers: the reference's `atomSpins(numAtoms,3)`
 * real(8) array in column-major order (all x components, then all y, then
 * all z), atoms in the reference's site order
 *     site = (ix*Ly + iy)*Lz + iz        (0-based, z fastest)
 * (src/chainMesh.f90:673-685 for the cell index, :878 for one atom per cell).
 *
 * Hamiltonian (one bond per nearest-neighbour pair, r_a = unit vector +a):
 *     E = -J sum_<ij> S_i.S_j + sum_i sum_a d_a.(S_i x S_{i+a})
 *         - zeeman_total_scale*Bz sum_i S_i^z - Kz sum_i (S_i^z)^2
 * with d_a = D r_a for bulk DMI and d_a = D (z x (-r_a)) for the reference's
 * interfacial form (src/energyMin.f90:114,135-137). Single-site energy
 * changes use Bz with coefficient 1 (src/energyMin.f90:161-162). On bcc and
 * hyper-kagome cells a bond i -> j with unit vector r carries
 * d_ij = sum_a r_a d_a (D r for bulk DMI, D (z x (-r)) for the interfacial
 * form), the nearest-neighbour shell being the reference's first atom shell
 * (src/chainMesh.f90:495-591).
 *
 * With ising_hamiltonian.demag = 1 the reference's demagnetisation term is
 * added: single-site changes -mu0 (S'-S).H_i and totals -mu0/2 sum S_i.H_i
 * (src/energyMin.f90:113-116,163-166,223-226), mu0 = ISING_MU0_REF, H the
 * field of calculate_demagnetisation_field (src/reciprocalSpaceProcesses.f90:
 * 202-281) in lattice units: cloud-in-cell deposit onto one mesh point per
 * cell centre, H(k) = -(k.M(k)) k / k^2 with the k = 0 mode removed, the
 * same weights back to the atoms. It needs a fully periodic lattice of at
 * most 1024 cells per axis (ISING_ERR_UNSUPPORTED otherwise).
 */
#ifndef ISING_GPU_H
#define ISING_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define ISING_OK                 0
#define ISING_ERR_INVALID_ARG   -1  /* null pointer, bad index, bad size      */
#define ISING_ERR_BAD_DIMS      -2  /* lattice shape the checkerboard rejects */
#define ISING_ERR_BAD_SPIN      -3  /* NaN / inf / non-unit spin uploaded     */
#define ISING_ERR_NO_DEVICE     -4  /* no HIP device or device index invalid  */
#define ISING_ERR_HIP           -5  /* a HIP runtime call failed              */
#define ISING_ERR_UNSUPPORTED   -6  /* e.g. demagnetisation on an open lattice */
#define ISING_ERR_OUT_OF_MEMORY -7
#define ISING_ERR_STATE         -8  /* call not valid in the current state    */

/* ---- DMI forms -------------------------------------------------------- */
#define ISING_DMI_NONE            0
#define ISING_DMI_BULK            1 /* d_a = D r_a                           */
#define ISING_DMI_INTERFACIAL_REF 2 /* d_a = D z x (-r_a), reference form    */

typedef struct ising_ctx ising_ctx;

/* The reference's constants::mu_0 (src/constants.f90:15): the literal
 * 1.25663706127e-6 is default (single-precision) real, so its value is the
 * float nearest to it. */
#define ISING_MU0_REF 1.256637006008532e-06

/* ---- unit cells ------------------------------------------------------- */
#define ISING_CELL_SC  0  /* one atom per cell, 6 nearest neighbours          */
#define ISING_CELL_BCC 1  /* atoms at (0,0,0) and (1/2,1/2,1/2), 8 nearest
                             neighbours (src/unit_cells.f90:7-30)             */
#define ISING_CELL_HYPERKAGOME 2 /* the 12 atoms of hyper_kagome_unit_cell
                             (src/unit_cells.f90:32-77), 4 nearest neighbours
                             at sqrt(2)/4 of the cell edge; swept in three
                             colours; periodic axes need >= 2 cells; no fused
                             sweep or learned proposal                        */

typedef struct {
    int32_t lx, ly, lz;      /* cells per axis                              */
    int32_t periodic[3];     /* 1 = periodic, 0 = open, per axis x,y,z      */
    int32_t cell;            /* ISING_CELL_*; sites = lx*ly*lz*(atoms per
                                cell), in the reference order
                                ((ix*ly + iy)*lz + iz)*atoms + atom
                                (src/chainMesh.f90:684,878)                 */
} ising_lattice;

typedef struct {
    double  J;                   /* E = -J S_i.S_j per bond                  */
    double  D;                   /* DMI magnitude                            */
    int32_t dmi_type;            /* ISING_DMI_*                              */
    int32_t demag;               /* 1: demagnetisation term, see below; 0: off */
    double  Bz;                  /* Zeeman field along z                     */
    double  Kz;                  /* uniaxial anisotropy, E -= Kz Sz^2        */
    double  zeeman_total_scale;  /* Zeeman coefficient in total energies:
                                    1 (default, consistent) or 2 (reference
                                    TotalHeisenbergEnergy, energyMin.f90:222) */
} ising_hamiltonian;

/* One parallel-tempering ladder: n_replicas copies of one lattice at
 * inverse temperatures betas[0..n_replicas-1]; swaps happen only between
 * neighbouring entries of betas. */
typedef struct {
    ising_lattice     lattice;
    ising_hamiltonian ham;
    int32_t           n_replicas;
    int32_t           reserved;
    double            cone_radius; /* proposal ball radius r (energyMin.f90:49-68) */
    const double*     betas;       /* n_replicas values, read at creation          */
} ising_ladder_desc;

typedef struct {
    int32_t n_sweeps;        /* sweeps to run                                 */
    int32_t swap_every;      /* replica-exchange round every k sweeps (0=off)  */
    int32_t measure_every;   /* E and M recorded every k sweeps (0=off)        */
    int32_t winding_every;   /* per-layer Q recorded every k sweeps (0=off)    */
    int32_t use_graph;       /* 1: replay one captured hipGraph per period     */
    int32_t winding_layer;   /* 0: Q of every layer; z+1: Q of layer z only ([n_q][R] series) */
} ising_schedule;

/* Measurement series of the last ising_pt_run, indexed by temperature
 * position t in the ladder concatenation (not by storage slot). Any pointer
 * may be NULL. Sizes: n_meas = n_sweeps/measure_every,
 * n_q = n_sweeps/winding_every, R = total replicas, Zmax = max lz. With a
 * schedule that names one layer (winding_layer = z+1) the charge series hold
 * that layer alone: [n_q][R]. */
typedef struct {
    double*  energy;         /* [n_meas][R]                                   */
    double*  magnetisation;  /* [n_meas][R][3] mean spin vector               */
    int64_t* winding;        /* [n_q][R][Zmax] nint(Q) per layer (0 padded)    */
    double*  winding_raw;    /* [n_q][R][Zmax] FP64 Q per layer                */
} ising_series;

/* ---- lifecycle -------------------------------------------------------- */
const char* ising_version(void);
int  ising_device_count(int* n);
/* Replaces: the makeChainMesh + lock-array set-up the reference PT driver does
 * per (param, T) mesh, src/PT.f90:180-211, batched into one context. */
int  ising_create(ising_ctx** out, const ising_ladder_desc* ladders, int n_ladders,
                  uint64_t seed, int device);
void ising_destroy(ising_ctx* ctx);                /* idempotent on NULL */
/* ctx == NULL: message of the last failed ising_create on this thread. */
const char* ising_last_error(const ising_ctx* ctx);
int  ising_num_replicas(const ising_ctx* ctx, int* n);
int  ising_replica_sites(const ising_ctx* ctx, int replica, int64_t* n);

/* ---- parameters (per ladder / per replica storage slot) --------------- */
int  ising_set_hamiltonian(ising_ctx* ctx, int ladder, const ising_hamiltonian* h);
int  ising_set_cone(ising_ctx* ctx, int ladder, double radius);
int  ising_set_beta(ising_ctx* ctx, int replica, double beta);
int  ising_get_betas(ising_ctx* ctx, double* beta_of_replica);

/* ---- spins ------------------------------------------------------------ */
/* Copy-in / copy-out of chainMesh%atomSpins (src/chainMesh.f90:889). */
int  ising_set_spins(ising_ctx* ctx, int replica, const double* soa, int64_t n);
int  ising_get_spins(ising_ctx* ctx, int replica, double* soa, int64_t n);
/* Seeded uniform-on-sphere start (replica < 0: all replicas). */
int  ising_init_random(ising_ctx* ctx, int replica);
/* Ferromagnet along +z with a seeded fraction of spins reversed. */
int  ising_init_ferro(ising_ctx* ctx, int replica, double flip_fraction);

/* ---- hot path --------------------------------------------------------- */
/* Replaces Metropolis_mcs(chainMesh, beta, numMCSSweeps, ...)
 * (src/energyMin.f90:234-374) for every replica of the context at once:
 * n_sweeps exact checkerboard Metropolis sweeps at each replica's beta. */
int  ising_sweep(ising_ctx* ctx, int n_sweeps);
/* Replaces TotalHeisenbergEnergy (src/energyMin.f90:499-528): FP64 total
 * energy of every replica slot, e[R]; ladders with demag = 1 first
 * recompute the field from the current spins, as the reference does. */
int  ising_energy(ising_ctx* ctx, double* e);
/* Total energy of one replica under an arbitrary Hamiltonian (drop-in for
 * TotalHeisenbergEnergy's per-call J, Dz, B arguments). */
int  ising_energy_with(ising_ctx* ctx, int replica, const ising_hamiltonian* h, double* e);
/* Mean spin vector per replica slot, m[R][3] (src/PT-Utils.f90:59-65). */
int  ising_magnetisation(ising_ctx* ctx, double* m);
/* Replaces calculate_winding_number2(chainMesh, Z_index)
 * (src/reciprocalSpaceProcesses.f90:320-364): z is 0-based here. */
int  ising_winding(ising_ctx* ctx, int replica, int z, int64_t* q, double* q_raw);
/* All layers of all replicas: q[R][Zmax], q_raw[R][Zmax]. */
int  ising_winding_all(ising_ctx* ctx, int64_t* q, double* q_raw);
/* Replaces calculate_winding_number_density(chainMesh, Z_index, density_matrix)
 * (src/reciprocalSpaceProcesses.f90:438-500): per-plaquette charge density of
 * layer z (0-based), density[j*lx + i] = (sigma1 + sigma2) / (4 pi) for the
 * plaquette with corner (i, j) -- the column-major order of
 * density_matrix(numCellsX, numCellsY); it sums to the layer's q_raw. */
int  ising_winding_density(ising_ctx* ctx, int replica, int z, double* density);
/* ---- skyrmion count -------------------------------------------------- */
/* Replaces calculate_winding_density_interpolated(chainMesh, z, density)
 * (src/reciprocalSpaceProcesses.f90:365-431): the charge density of layer z
 * (0-based) on a grid scale_x times finer in x and scale_y in y (the
 * reference's sclx, scly, 10 by default, src/chainMesh.f90:764-765):
 * band-limited interpolation of the corner spins (each component; an even
 * axis splits its Nyquist term between +-N/2), then per fine plaquette
 * (sigma(s1,s2,s3) + sigma(s1,s3,s4)) / (4 pi), s2 = +x, s3 = +x+y, s4 = +y,
 * periodic; density[J*NX + I], NX = scale_x lx. Needs periodic x and y
 * (ISING_ERR_UNSUPPORTED) and at most 2^26 fine points. */
int  ising_winding_density_interpolated(ising_ctx* ctx, int replica, int z, int scale_x, int scale_y,
                                        double* density);

typedef struct {
    int32_t scale_x, scale_y;  /* fine grid per cell, 10 x 10 in the reference           */
    int32_t n_particles;       /* N: counts of regions of charge 1..N                     */
    int32_t num_thresholds;    /* mask thresholds, 60 in chainMesh_statistics (<= 254)    */
    int32_t num_sigmas;        /* Gaussian widths, 20 in the reference                    */
    int32_t reserved;
    double  min_threshold;     /* 1e-5 in chainMesh_statistics                            */
    double  max_threshold;     /* 0.95                                                    */
} ising_skyrmion_params;

/* Replaces compute_skyrmion_distribution(chainMesh, N, winding_array,
 * min_threshold, max_threshold, num_thresholds, Z_index)
 * (src/reciprocalSpaceProcesses.f90:930-983) and the calculate_skyrmion_number
 * calls it makes (:627-812), with the intended semantics: for each
 * threshold q_t = max - (t/T)(max - min) (t = 1..T) and width
 * sigma_k = 1 - (k/S)(1 - 1e-5) (k = 1..S) the interpolated density is
 * filtered by a centred 5x5 Gaussian, pixels with |f| > q_t max|f| are
 * masked, every masked pixel joins the local maximum its steepest ascent over
 * the 8 neighbours reaches, and each region whose summed density lies within
 * 0.05 of a particle number j adds its sign once to count j. counts[N]
 * receives the first (t, k) entry (t outer) whose sum_j j count_j lies within
 * 0.1 of the layer's winding number, else the first whose nearest integer came
 * closest; table (may be NULL) receives every entry, [T][S][N]. Deviations
 * from the reference's arithmetic are listed with the kernels. */
int  ising_skyrmion_distribution(ising_ctx* ctx, int replica, int z, const ising_skyrmion_params* params,
                                 int32_t* counts, int32_t* table);

/* Test hook standing in for calculateHeisenbergEnergy
 * (src/energyMin.f90:86-171): FP32 device dE = E(S') - E(S) of single-site
 * proposals props[n][3] at sites[n] (reference site order) of one replica. */
int  ising_delta_energy(ising_ctx* ctx, int replica, const int32_t* sites,
                        const double* props, int n, double* de);

/* Device-resident checkpoint of the chain state: spins, which slot sits at
 * which temperature, the betas and the sweep / record / swap counters (so a
 * restored run repeats the same random numbers). One checkpoint per context;
 * the first save allocates it. */
int  ising_save_state(ising_ctx* ctx);
int  ising_restore_state(ising_ctx* ctx);

/* ---- demagnetisation -------------------------------------------------- */
/* Replaces calculate_demagnetisation_field(chainMesh, outputArray)
 * (src/reciprocalSpaceProcesses.f90:202-281): the field H of the current
 * spins of `replica`, h[3*n] in the atomSpins layout (n = sites), FP64. */
int  ising_demag_field(ising_ctx* ctx, int replica, double* h, int64_t n);
/* The field a replica's sweeps and ising_delta_energy use now (the last one
 * computed or set; zero before any), without recomputing it: the state of
 * chainMesh%demagnetisation_array after Metropolis_mcs. */
int  ising_get_demag_field(ising_ctx* ctx, int replica, double* h, int64_t n);
/* Sets the field a replica's sweeps and ising_delta_energy use until the
 * next recomputation (calculateHeisenbergEnergy takes the field as its
 * demagnetisation_array argument, src/energyMin.f90:87,113-115). */
int  ising_set_demag_field(ising_ctx* ctx, int replica, const double* h, int64_t n);
/* ising_sweep on ladders with demag = 1 recomputes the field from the spins
 * before its first sweep and after every sweep whose 1-based number in the
 * call is a multiple of `sweeps` (0: only before the first), as
 * Metropolis_mcs does with int(pi sqrt 5 / (2 r)) (src/energyMin.f90:258-259,
 * 362-369). Default 1. */
int  ising_set_demag_interval(ising_ctx* ctx, int sweeps);

/* ---- parallel tempering ---------------------------------------------- */
/* Replaces the PT iteration loop src/PT.f90:219-273 (sweeps + swaps) plus
 * the per-temperature statistics the reference takes at the end. Ladders
 * with demag = 1 are refused (the reference's PT runs without the term,
 * src/PT.f90:229-230,251-254). */
int  ising_pt_run(ising_ctx* ctx, const ising_schedule* sched, const ising_series* out);
/* Allocation-free variant of ising_pt_run. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc). The same chains,
 * records and results as ising_pt_run, but the series are staged in a
 * statically allocated device array of ISING_STATIC_SCRATCH_BYTES per GPU
 * (held until the process ends) and every period is launched eagerly, so the
 * call makes no hipMalloc, hipFree or graph instantiation. The staging needs
 * 256 + (n_meas*R*4 + n_q*R*Zq*2)*8 bytes plus alignment (n_meas, n_q, R as
 * in ising_series; Zq = Zmax, or 1 when the schedule names one layer);
 * larger schedules return ISING_ERR_INVALID_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several host threads, or several contexts on one device used at
 * once) race. */
#define ISING_STATIC_SCRATCH_BYTES (16u << 20)
int ising_pt_run_static(ising_ctx* ctx, const ising_schedule* sched, const ising_series* out);
/* slot_of_temp[R]: storage slot currently at temperature position t
 * (the reference's TemperatureMeshArray, src/PT.f90:58,201,264-268). */
int  ising_temperature_map(ising_ctx* ctx, int32_t* slot_of_temp);
/* Per neighbouring pair (t, t+1) of each ladder, at index t of the ladder
 * concatenation: cumulative attempts and accepts; entry of a ladder's last
 * temperature is 0. */
int  ising_swap_stats(ising_ctx* ctx, int64_t* attempts, int64_t* accepts);
/* Accepted single-site moves per replica slot since creation. */
int  ising_accept_counts(ising_ctx* ctx, int64_t* accepted);

/* ---- measurement (no reference counterpart) --------------------------- */
/* Runs n_sweeps sweeps like ising_sweep with one hipEvent pair around their
 * back-to-back launches on the context's stream (two half-sweep launches per
 * sweep, or one fused launch); *ms_per_launch = elapsed / (2*n_sweeps), the
 * time per half sweep (one update of every site of one colour). */
int  ising_time_sweeps(ising_ctx* ctx, int n_sweeps, double* ms_per_launch);
/* Runs n_sweeps sweeps of a PT schedule like ising_pt_run (swaps, E/M and
 * Q measurements; records are computed, not stored), launching every period
 * eagerly; each plain (non-measuring) sweep launch carries start/stop
 * events in its dispatch packet (hipExtLaunchKernel), and *ms_per_launch =
 * their mean device time per half sweep: the time of the dominant kernel
 * inside the workload it belongs to (back-to-back launches alone run at a
 * lower clock under the board's power limit). */
int  ising_time_pt_sweeps(ising_ctx* ctx, const ising_schedule* sched, int n_sweeps, double* ms_per_launch);
/* Recomputes the demagnetising field of every replica whose ladder has
 * demag = 1 from its current spins, n_fields times back to back, with one
 * hipEvent pair around them on the context's stream; *ms_per_field = elapsed
 * / n_fields. The stored fields are left as one recomputation leaves them.
 * ISING_ERR_INVALID_ARG when no ladder has the term. */
int  ising_time_demag(ising_ctx* ctx, int n_fields, double* ms_per_field);
/* hipDeviceSynchronize on the context's device. */
int  ising_synchronize(ising_ctx* ctx);
/* Plain device-to-device copy kernel over `bytes` bytes (two scratch buffers
 * allocated and freed by the call); *ms = device time of one copy. Gives the
 * achievable-bandwidth reference and the known-size kernel that scales HBM
 * traffic counters. */
int  ising_bandwidth_probe(ising_ctx* ctx, int64_t bytes, double* ms);

/* ---- launch tuning (no reference counterpart) ------------------------- */
/* Every flag set gives bit-identical chains; only the launch shape changes. */
#define ISING_TUNE_XCD_REMAP 1u  /* contiguous block ranges per accelerator die */
#define ISING_TUNE_NT_OWN    2u  /* non-temporal access to the updated colour (half sweeps) */
/* One fused launch per sweep instead of two half-sweep launches: each
 * workgroup stages z planes in LDS and updates both colours of lc
 * consecutive planes (1..255); simple-cubic contexts only. A plane of up to
 * 2048 sites of one colour is one workgroup's; larger planes, or any plane
 * with ISING_TUNE_SLAB(hy) (2..255 rows, with a fused lc), are cut into y
 * slabs of hy owned rows per workgroup (hy = 0: the tallest that fits), each
 * recomputing two halo rows on either side. ISING_ERR_UNSUPPORTED when a
 * replica is not simple cubic or a row has more than 341 sites of one colour
 * (no window of six rows fits). While a ladder takes its cone radius from
 * the network (ising_mlp_enable) the sweeps run as half sweeps under the
 * same flags. */
#define ISING_TUNE_FUSED_SHIFT 8
#define ISING_TUNE_FUSED(lc) ((uint32_t)(lc) << ISING_TUNE_FUSED_SHIFT)
#define ISING_TUNE_SLAB_SHIFT 16
#define ISING_TUNE_SLAB(hy) ((uint32_t)(hy) << ISING_TUNE_SLAB_SHIFT)
#define ISING_TUNE_MAX_CANDIDATES 16
int  ising_set_tuning(ising_ctx* ctx, uint32_t flags);
int  ising_get_tuning(const ising_ctx* ctx, uint32_t* flags);
/* The flag sets ising_autotune times, in order (at most
 * ISING_TUNE_MAX_CANDIDATES); flags may be NULL to query the count: the half
 * sweeps with and without ISING_TUNE_XCD_REMAP and ISING_TUNE_NT_OWN, then
 * (simple-cubic contexts without the demagnetising term) fused chunks of 8,
 * 16, 32 or all z planes with and without ISING_TUNE_XCD_REMAP; when every
 * replica's planes exceed one workgroup, also 24 planes, each chunk size with
 * the slab height of the fewest estimated rounds of workgroups. */
int  ising_tuning_candidates(const ising_ctx* ctx, uint32_t* flags, int* count);
/* Times every candidate over n_sweeps real sweeps each (they advance the
 * chains like ising_sweep) and keeps the fastest; ms (may be NULL, room for
 * ISING_TUNE_MAX_CANDIDATES) receives the time per half sweep of each
 * candidate in ising_tuning_candidates order. */
int  ising_autotune(ising_ctx* ctx, int n_sweeps, uint32_t* best, double* ms);

/* Sweeps-per-swap schedule tuning (no reference counterpart; the reference
 * fixes 100 sweeps per swap round, src/PT.f90:46-48). Times the schedule
 * *sched with swap_every = each of cand[n_cand] (at most
 * ISING_TUNE_MAX_CANDIDATES, every value >= 1) over n_sweeps real sweeps
 * each (they advance the chains; records are computed, not stored), and sets
 * sched->swap_every to the smallest candidate whose time per sweep is within
 * (1 + max_overhead) of the fastest candidate's: deterministic even/odd
 * exchange mixes temperatures fastest when it runs most often, so the tuner
 * trades only device time. ms_per_sweep (may be NULL, n_cand entries)
 * receives each candidate's time per sweep. */
int  ising_autotune_schedule(ising_ctx* ctx, ising_schedule* sched, const int32_t* cand, int n_cand,
                             int n_sweeps, double max_overhead, double* ms_per_sweep);

/* ---- learned proposal (no reference counterpart) ---------------------- */
/* A per-site network sets the proposal radius from the six neighbour spins
 * and beta*h (never the site's own spin, so the proposal stays symmetric):
 * x[21] -> 32 -> 32 -> 1 with SiLU in base 2 (z / (1 + 2^-z)), bfloat16
 * weights, FP32 accumulation,
 * output log r clamped to [ln 0.01, ln 8]. Parameters (FP32 master copy):
 * W1[32][21], b1[32], W2[32][32], b2[32], w3[32], b3. */
#define ISING_MLP_NPARAM 1793
#define ISING_MLP_RECORD 32  /* floats per training record: x[21], S[3], h[3], beta, K, 3 pad */
int  ising_mlp_set_params(ising_ctx* ctx, const float* params, int n);
int  ising_mlp_get_params(ising_ctx* ctx, float* params, int n);
/* Sweeps of `ladder` (< 0: all) take their radius from the network. */
int  ising_mlp_enable(ising_ctx* ctx, int ladder, int on);
/* Appends per_replica records per replica from the current configurations. */
int  ising_mlp_collect(ising_ctx* ctx, int per_replica);
int  ising_mlp_num_records(const ising_ctx* ctx, int64_t* n);
int  ising_mlp_clear_records(ising_ctx* ctx);
int  ising_mlp_get_records(ising_ctx* ctx, float* out, int64_t n);
/* Adam steps on the collected records maximising the expected squared jump
 * distance of the proposal; loss[steps] (may be NULL) = mean batch loss. */
int  ising_mlp_train(ising_ctx* ctx, int steps, int batch, double lr, double* loss);
/* Test hooks: the device forward pass on rows x[n][21]; the gradient of the
 * mean loss over records[n] at optimiser step `step` (no update); one
 * 32x32x16 bfloat16 MFMA D = A B with the operand layout the kernels use. */
int  ising_mlp_forward(ising_ctx* ctx, const float* x, int n, float* f);
int  ising_mlp_grad(ising_ctx* ctx, const float* records, int n, uint32_t step, double* loss, float* grad);
int  ising_mfma_probe(ising_ctx* ctx, const float* A, const float* B, float* D);

#ifdef __cplusplus
}
#endif
#endif /* ISING_GPU_H */
