// ising_hip.hip -- CDNA4 (gfx950) kernels and C ABI of libising_hip.so.
//
// Hot path: exact colour-by-colour Metropolis (two-colour checkerboard on
// simple-cubic and bcc lattices, three colours on hyper-kagome) for
// classical unit spins (exchange + DMI + Zeeman + uniaxial anisotropy, and the
// reference's demagnetising field), many replicas per launch, FP32 spins,
// Philox4x32-10 keyed by (seed, replica slot) with counter (site, sweep).
// Energies, magnetisation and topological charge are reduced in FP64; replica
// exchange runs on device.
//
// HBM layout of one simple-cubic replica with N = lx*ly*lz sites, hx = lx/2:
//   spins[off + c*N + k*N/2 + (iz*ly + iy)*hx + (ix >> 1)]
// component c in {x,y,z}, colour k = (ix+iy+iz) & 1. One half-sweep of colour
// k streams the k arrays once (coalesced: consecutive lanes take consecutive
// ix) and reads the other colour's rows iy+-1, iz+-1 and its own row as
// neighbours. bcc and hyper-kagome replicas store each colour as one block
// of the same arrays (see bcc_steps and HkTable). Replicas are packed back to back; a block table maps each
// workgroup to (replica, chunk), so replicas of different sizes share one
// launch.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>
#include <stdarg.h>
#include <math.h>
#include <cmath>
#include <string>
#include <vector>
#include <mutex>
#include <algorithm>
#include <numeric>
#include <type_traits>

#include "ising_gpu.h"
#include "philox.hpp"
#include "mlp.hpp"

namespace ising {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kXcds = 8;          // MI355X: 8 accelerator dies, workgroups dealt round robin

// Sweep tuning flags (bit mask, chosen by ising_autotune).
constexpr uint32_t kTuneXcdRemap = 1u;     // give each die a contiguous range of blocks
constexpr uint32_t kTuneNtOwn = 2u;        // non-temporal loads/stores of the updated colour
constexpr uint32_t kTuneFusedShift = 8;    // bits 8..15: z planes per fused-sweep workgroup (0: half sweeps)
constexpr uint32_t kTuneFusedMask = 0xFFu << kTuneFusedShift;
constexpr uint32_t kTuneSlabShift = 16;    // bits 16..23: rows per y slab of the fused sweep (0: whole planes where they fit)
constexpr uint32_t kTuneSlabMask = 0xFFu << kTuneSlabShift;

// Logical workgroup index: with kTuneXcdRemap the blocks a die receives
// (b mod 8 == die) map to one contiguous range, so neighbouring rows of a
// lattice stay in that die's L2. A bijection on [0, n).
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t n, uint32_t tune)
{
    if (!(tune & kTuneXcdRemap)) return b;
    const uint32_t die = b % kXcds, q = n / kXcds, r = n % kXcds;
    return die * q + min(die, r) + b / kXcds;
}

// ---------------------------------------------------------------------------
// Division by an invariant 32-bit divisor (Granlund & Montgomery, PLDI'94,
// fig. 4.1): n / d = (t + ((n - t) >> s1)) >> s2 with t = mulhi(n, m).
// A power-of-two divisor is stored with m = 0 and s2 = log2(d): the quotient
// is a shift and the remainder a mask (a wavefront-uniform branch).
struct FastDiv { uint32_t m, s1, s2, d; };

static FastDiv make_fastdiv(uint32_t d)
{
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    if ((1ull << l) == d) return FastDiv{ 0u, 0u, l, d };
    uint64_t m = (((1ull << 32) * ((1ull << l) - d)) / d) + 1;
    return FastDiv{ (uint32_t)m, l ? 1u : 0u, l ? l - 1u : 0u, d };
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f)
{
    if (f.m == 0u) return n >> f.s2;
    uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// n - quot * d for quot = fdiv(n, f).
__device__ __forceinline__ uint32_t fmodq(uint32_t n, uint32_t quot, const FastDiv& f)
{
    return f.m == 0u ? (n & (f.d - 1u)) : n - quot * f.d;
}

// Per replica storage slot.
struct Rep {
    int32_t lx, ly, lz;
    int32_t hx;             // row length in one colour's arrays: lx/2 (SC), lx (bcc, hyper-kagome)
    int32_t cell;           // ISING_CELL_*
    int32_t per;            // bit 0/1/2: axis x/y/z periodic
    int32_t ladder;
    int32_t first_block;    // first workgroup in the sweep/measure grid
    int32_t nblocks;
    int32_t q_first_block;  // first workgroup in the winding grid
    int32_t q_blocks_layer; // winding workgroups per z layer
    int32_t l_first_block;  // first workgroup in the one-layer winding grid
    int32_t zbase;          // first entry of this slot in [R][Zmax] arrays
    int32_t f_first_block;  // first workgroup in the fused-sweep grid
    int32_t f_nblocks;      // fused-sweep workgroups (z chunks)
    int32_t plane;          // ly*hx: sites of one colour on one z plane
    int32_t ncol;           // colours of the sweep: 2, or 3 on hyper-kagome
    int32_t f_rows;         // fused sweep: rows of a workgroup's window (ly: the whole plane)
    int32_t f_hy;           // fused sweep: rows a y slab owns (ly: the whole plane)
    int32_t f_nslab;        // fused sweep: y slabs per z chunk (1: whole planes)
    uint32_t half;          // sites of one colour: N/2, or N/3 on hyper-kagome
    int64_t off;            // float offset of the x component block
    int64_t n;              // N
    FastDiv div_hx, div_ly, div_lx, div_lz;
};

// Per ladder Hamiltonian, FP32 copy for sweeps and FP64 copy for energies.
struct Lad {
    float J, Bz, Kz, r;
    float lr8;              // log2(r) - 8 (-inf for r = 0)
    float zsf;              // Zeeman coefficient of total energies
    float d[9];
    int32_t dmi;
    int32_t learned;        // cone radius from the network (mlp.hpp)
    double Jd, Bzd, Kzd, zs;
    double dd[9];
    int32_t demag;          // demagnetisation term on (ising_hamiltonian.demag)
    int32_t pad_;
    double mu0;             // its coefficient (0 when off)
};

struct f3 { float x, y, z; };

// Element i of a component plane: a 32-bit byte offset from a wavefront-
// uniform base, so loads use the scalar-base + 32-bit vector-offset form.
__device__ __forceinline__ float ldf(const float* __restrict__ base, uint32_t i)
{
    return *(const float*)((const char*)base + (i << 2));
}
__device__ __forceinline__ void stf(float* __restrict__ base, uint32_t i, float v)
{
    *(float*)((char*)base + (i << 2)) = v;
}

__device__ __forceinline__ float ldf_nt(const float* __restrict__ base, uint32_t i)
{
    return __builtin_nontemporal_load((const float*)((const char*)base + (i << 2)));
}
__device__ __forceinline__ void stf_nt(float* __restrict__ base, uint32_t i, float v)
{
    __builtin_nontemporal_store(v, (float*)((char*)base + (i << 2)));
}

// The three component planes (x, y, z) of one colour of one replica.
struct Planes { const float* x; const float* y; const float* z; };

__device__ __forceinline__ Planes planes(const float* spins, const struct Rep& R, int colour);

__device__ __forceinline__ f3 ld3(const Planes& P, uint32_t i)
{
    return f3{ ldf(P.x, i), ldf(P.y, i), ldf(P.z, i) };
}

__device__ __forceinline__ Planes planes(const float* spins, const Rep& R, int colour)
{
    const float* b = spins + R.off + (int64_t)colour * R.half;
    return Planes{ b, b + R.n, b + 2 * R.n };
}

// Neighbour indices (inside the other colour's arrays, which share the
// (row, half index) layout) of the site at q = row*hx + h on row (iy, iz),
// parity p = ix & 1: q plus a fixed step per direction, corrected by one
// period at a periodic face. ok[] false across open faces.
// kPer: every axis periodic (every neighbour exists).
struct Nbr { uint32_t i[6]; bool ok[6]; };

template <bool kPer>
__device__ __forceinline__ Nbr neighbours(const Rep& R, uint32_t q, uint32_t h, uint32_t iy, uint32_t iz, uint32_t p)
{
    Nbr nb;
    const uint32_t hx = (uint32_t)R.hx, ly = (uint32_t)R.ly, lz = (uint32_t)R.lz;
    const uint32_t plane = ly * hx, vol = R.half;
    const bool px = kPer || (R.per & 1), py = kPer || ((R.per >> 1) & 1), pz = kPer || ((R.per >> 2) & 1);
    // +x: half index h+p; ix+1 wraps to 0 when p = 1 and h = hx-1.
    const bool xp_w = (p != 0u) && (h == hx - 1u);
    nb.i[0] = q + p - (xp_w ? hx : 0u);
    nb.ok[0] = !xp_w || px;
    // -x: half index h+p-1; ix-1 wraps to lx-1 when p = 0 and h = 0.
    const bool xm_w = (p == 0u) && (h == 0u);
    nb.i[1] = q + p - 1u + (xm_w ? hx : 0u);
    nb.ok[1] = !xm_w || px;
    const bool yp_w = (iy + 1u == ly), ym_w = (iy == 0u);
    nb.i[2] = q + hx - (yp_w ? plane : 0u);
    nb.ok[2] = !yp_w || py;
    nb.i[3] = q - hx + (ym_w ? plane : 0u);
    nb.ok[3] = !ym_w || py;
    const bool zp_w = (iz + 1u == lz), zm_w = (iz == 0u);
    nb.i[4] = q + plane - (zp_w ? vol : 0u);
    nb.ok[4] = !zp_w || pz;
    nb.i[5] = q - plane + (zm_w ? vol : 0u);
    nb.ok[5] = !zm_w || pz;
    return nb;
}

// The six neighbour spins (+x, -x, +y, -y, +z, -z); zero across open faces.
// Every index of neighbours() is in range (an open face yields the periodic
// image), so all lanes load unconditionally in one straight-line block and
// open lattices clear the absent neighbours afterwards (a uniform branch).
__device__ __forceinline__ void load_nbrs(const Rep& R, const Planes& oth, const Nbr& nb, f3 (&v)[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = ld3(oth, nb.i[k]);
    if (R.per != 7) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (!nb.ok[k]) v[k] = f3{ 0.f, 0.f, 0.f };
    }
}

// Local field h = J sum - sum_a A_a x d_a + Bz z from the sum of the
// neighbour spins and the signed sums A_a = sum_b sgn_a(b) S_b over the bonds
// (SC: A_a = S_{+a} - S_{-a}; bcc: eight bonds, d_a pre-scaled by 1/sqrt 3).
__device__ __forceinline__ f3 field_from_sums(const Lad& L, const f3& sum, const f3& dx, const f3& dy, const f3& dz)
{
    f3 dm{ 0.f, 0.f, 0.f };
    if (L.dmi == ISING_DMI_BULK) {
        const float D = L.d[0];
        dm = f3{ D * (dz.y - dy.z), D * (dx.z - dz.x), D * (dy.x - dx.y) };
    } else if (L.dmi != ISING_DMI_NONE) {
        const float* d = L.d;
        dm.x = (dx.y * d[2] - dx.z * d[1]) + (dy.y * d[5] - dy.z * d[4]) + (dz.y * d[8] - dz.z * d[7]);
        dm.y = (dx.z * d[0] - dx.x * d[2]) + (dy.z * d[3] - dy.x * d[5]) + (dz.z * d[6] - dz.x * d[8]);
        dm.z = (dx.x * d[1] - dx.y * d[0]) + (dy.x * d[4] - dy.y * d[3]) + (dz.x * d[7] - dz.y * d[6]);
    }
    return f3{ L.J * sum.x - dm.x, L.J * sum.y - dm.y, L.J * sum.z - dm.z + L.Bz };
}

// SC: h = J sum_j S_j - sum_a (S_{+a} - S_{-a}) x d_a + Bz z.
__device__ __forceinline__ f3 field_from(const Lad& L, const f3 (&v)[6])
{
    const f3 &xp = v[0], &xm = v[1], &yp = v[2], &ym = v[3], &zp = v[4], &zm = v[5];
    const f3 sum{ (xp.x + xm.x) + (yp.x + ym.x) + (zp.x + zm.x),
                  (xp.y + xm.y) + (yp.y + ym.y) + (zp.y + zm.y),
                  (xp.z + xm.z) + (yp.z + ym.z) + (zp.z + zm.z) };
    const f3 dx{ xp.x - xm.x, xp.y - xm.y, xp.z - xm.z };
    const f3 dy{ yp.x - ym.x, yp.y - ym.y, yp.z - ym.z };
    const f3 dz{ zp.x - zm.x, zp.y - zm.y, zp.z - zm.z };
    return field_from_sums(L, sum, dx, dy, dz);
}

// bcc: the eight neighbours (the other atom) of the site at q = row*lx + ix
// of cell (ix, iy, iz) with colour (= atom) k are the cells (ix - dx, iy - dy,
// iz - dz) for k = 0 and (ix + dx, iy + dy, iz + dz) for k = 1, dx, dy, dz in
// {0, 1}; the bond vector's sign on an axis is +1 (k = 0) or -1 (k = 1) for
// d = 0 and the opposite for d = 1. Neighbours across open faces count as
// zero (the index is then the periodic image, always in range). Returns the
// sum of the spins and the signed sums A_a.
struct BccSums { f3 sum, ax, ay, az; };

// Steps (modular uint32) from q to the shifted cell along x, y, z and whether
// that cell exists; bond b = (dx, dy, dz) reaches q + dx*ox + dy*oy + dz*oz.
struct BccSteps { uint32_t ox, oy, oz; bool okx, oky, okz; };

__device__ __forceinline__ BccSteps bcc_steps(const Rep& R, uint32_t ix, uint32_t iy, uint32_t iz, int colour)
{
    const uint32_t lx = (uint32_t)R.lx, ly = (uint32_t)R.ly, lz = (uint32_t)R.lz;
    const uint32_t plane = ly * lx;
    BccSteps t;
    if (colour == 0) {
        t.ox = ix == 0u ? lx - 1u : 0u - 1u;                     t.okx = ix > 0u || (R.per & 1);
        t.oy = iy == 0u ? (ly - 1u) * lx : 0u - lx;              t.oky = iy > 0u || ((R.per >> 1) & 1);
        t.oz = iz == 0u ? (lz - 1u) * plane : 0u - plane;        t.okz = iz > 0u || ((R.per >> 2) & 1);
    } else {
        t.ox = ix + 1u == lx ? 0u - (lx - 1u) : 1u;              t.okx = ix + 1u < lx || (R.per & 1);
        t.oy = iy + 1u == ly ? 0u - (ly - 1u) * lx : lx;         t.oky = iy + 1u < ly || ((R.per >> 1) & 1);
        t.oz = iz + 1u == lz ? 0u - (lz - 1u) * plane : plane;   t.okz = iz + 1u < lz || ((R.per >> 2) & 1);
    }
    return t;
}

__device__ __forceinline__ uint32_t bcc_index(const BccSteps& t, uint32_t q, int b)
{
    return q + (((b >> 2) & 1) ? t.ox : 0u) + (((b >> 1) & 1) ? t.oy : 0u) + ((b & 1) ? t.oz : 0u);
}

__device__ __forceinline__ bool bcc_ok(const BccSteps& t, int b)
{
    return (!((b >> 2) & 1) || t.okx) && (!((b >> 1) & 1) || t.oky) && (!(b & 1) || t.okz);
}

__device__ __forceinline__ BccSums bcc_sums(const Rep& R, const Planes& oth, uint32_t q, uint32_t ix, uint32_t iy,
                                            uint32_t iz, int colour)
{
    const BccSteps st = bcc_steps(R, ix, iy, iz, colour);
    f3 v[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = ld3(oth, bcc_index(st, q, b));
    if (R.per != 7) {
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!bcc_ok(st, b)) v[b] = f3{ 0.f, 0.f, 0.f };
    }
    // Sums over d = 0 and d = 1 per axis.
    f3 x0{ 0.f, 0.f, 0.f }, x1 = x0, y0 = x0, y1 = x0, z0 = x0, z1 = x0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const f3 t = v[b];
        f3& xs = (b >> 2) & 1 ? x1 : x0;
        f3& ys = (b >> 1) & 1 ? y1 : y0;
        f3& zs = b & 1 ? z1 : z0;
        xs.x += t.x; xs.y += t.y; xs.z += t.z;
        ys.x += t.x; ys.y += t.y; ys.z += t.z;
        zs.x += t.x; zs.y += t.y; zs.z += t.z;
    }
    const float sg = colour == 0 ? 1.f : -1.f;
    BccSums r;
    r.sum = f3{ x0.x + x1.x, x0.y + x1.y, x0.z + x1.z };
    r.ax = f3{ sg * (x0.x - x1.x), sg * (x0.y - x1.y), sg * (x0.z - x1.z) };
    r.ay = f3{ sg * (y0.x - y1.x), sg * (y0.y - y1.y), sg * (y0.z - y1.z) };
    r.az = f3{ sg * (z0.x - z1.x), sg * (z0.y - z1.y), sg * (z0.z - z1.z) };
    return r;
}

__device__ __forceinline__ f3 bcc_field(const Lad& L, const BccSums& b)
{
    return field_from_sums(L, b.sum, b.ax, b.ay, b.az);
}

// Hyper-kagome (src/unit_cells.f90:32-77): twelve atoms per cubic cell, four
// nearest neighbours each at sqrt(2)/4 of the edge. The bond graph has
// triangles, so the sweep takes three colours of four atom types each. A
// site's storage rank is colour*4 + its type's place in the colour; site
// (rank r, cell c) sits at r*C + c of the component arrays, C = lx*ly*lz,
// c = (iz*ly + iy)*lx + ix, so each colour is one block of half = 4C sites
// and the per-colour planes of the two-colour cells carry over. The host
// builds the table from the atom positions (hk_build_table) into constant
// memory.
constexpr int kHkAtoms = 12;
constexpr int kHkBonds = 4;
constexpr int kHkPerColour = 4;

struct HkBond {
    int32_t rank;            // storage rank of the neighbour's atom type
    int32_t dx, dy, dz;      // cell step to the neighbour, each in {-1, 0, 1}
    float r[3];              // unit bond vector, site -> neighbour
    double rd[3];
};

struct HkTable {
    int32_t type_of_rank[kHkAtoms];
    int32_t rank_of_type[kHkAtoms];
    HkBond bond[kHkAtoms][kHkBonds];    // by storage rank
    double frac[kHkAtoms][3];           // atom position in its cell (units of the edge), by type
};

__constant__ HkTable c_hk;

struct HkSite { uint32_t ix, iy, iz, rank, cell; };

// Cell and storage rank of index q of colour `colour` (q < half).
__device__ __forceinline__ HkSite hk_site(const Rep& R, uint32_t q, int colour)
{
    const uint32_t row = fdiv(q, R.div_hx);
    const uint32_t ix = q - row * (uint32_t)R.hx;
    const uint32_t r2 = fdiv(row, R.div_ly);
    const uint32_t iy = row - r2 * (uint32_t)R.ly;
    const uint32_t j = fdiv(r2, R.div_lz);
    const uint32_t iz = r2 - j * (uint32_t)R.lz;
    const uint32_t C = (uint32_t)R.plane * (uint32_t)R.lz;
    return HkSite{ ix, iy, iz, (uint32_t)colour * kHkPerColour + j, q - j * C };
}

// Index rank*C + cell of bond B's neighbour (always in range: the periodic
// image across an open face) and whether that bond exists.
__device__ __forceinline__ bool hk_nbr(const Rep& R, const HkSite& t, const HkBond& B, uint32_t& idx)
{
    int c[3] = { (int)t.ix + B.dx, (int)t.iy + B.dy, (int)t.iz + B.dz };
    const int n[3] = { R.lx, R.ly, R.lz };
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] < 0) { c[a] += n[a]; ok = ok && ((R.per >> a) & 1); }
        else if (c[a] >= n[a]) { c[a] -= n[a]; ok = ok && ((R.per >> a) & 1); }
    }
    const uint32_t C = (uint32_t)R.plane * (uint32_t)R.lz;
    idx = (uint32_t)B.rank * C + ((uint32_t)c[2] * (uint32_t)R.ly + (uint32_t)c[1]) * (uint32_t)R.lx + (uint32_t)c[0];
    return ok;
}

// Cell-index steps of one site to the cell below and above it on each axis
// (the periodic image at a face) and whether those cells exist.
struct HkSteps { int32_t dn[3], up[3]; bool okdn[3], okup[3]; };

__device__ __forceinline__ HkSteps hk_steps(const Rep& R, const HkSite& t)
{
    const uint32_t c[3] = { t.ix, t.iy, t.iz };
    const int32_t n[3] = { R.lx, R.ly, R.lz };
    const int32_t stride[3] = { 1, R.lx, R.lx * R.ly };
    HkSteps s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool per = (R.per >> a) & 1;
        s.dn[a] = c[a] == 0u ? (n[a] - 1) * stride[a] : -stride[a];
        s.up[a] = c[a] + 1u == (uint32_t)n[a] ? -(n[a] - 1) * stride[a] : stride[a];
        s.okdn[a] = c[a] > 0u || per;
        s.okup[a] = c[a] + 1u < (uint32_t)n[a] || per;
    }
    return s;
}

// The field sum over the four bonds of storage rank `rank` (wave-uniform in
// the common case, so the table reads are scalar).
__device__ __forceinline__ f3 hk_field_rank(const Lad& L, const Rep& R, const Planes& all, const HkSite& t,
                                            const HkSteps& st, uint32_t rank)
{
    const uint32_t C = (uint32_t)R.plane * (uint32_t)R.lz;
    f3 sum{ 0.f, 0.f, 0.f }, ax = sum, ay = sum, az = sum;
#pragma unroll
    for (int b = 0; b < kHkBonds; ++b) {
        const HkBond& B = c_hk.bond[rank][b];
        const int32_t d[3] = { B.dx, B.dy, B.dz };
        int32_t step = 0;
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (d[a] < 0) { step += st.dn[a]; ok = ok && st.okdn[a]; }
            else if (d[a] > 0) { step += st.up[a]; ok = ok && st.okup[a]; }
        }
        f3 v = ld3(all, (uint32_t)B.rank * C + (uint32_t)((int32_t)t.cell + step));
        if (!ok) v = f3{ 0.f, 0.f, 0.f };
        sum.x += v.x; sum.y += v.y; sum.z += v.z;
        ax.x += B.r[0] * v.x; ax.y += B.r[0] * v.y; ax.z += B.r[0] * v.z;
        ay.x += B.r[1] * v.x; ay.y += B.r[1] * v.y; ay.z += B.r[1] * v.z;
        az.x += B.r[2] * v.x; az.y += B.r[2] * v.y; az.z += B.r[2] * v.z;
    }
    return field_from_sums(L, sum, ax, ay, az);
}

// h = J sum_b S_b - sum_a A_a x d_a + Bz z with A_a = sum_b r_a(b) S_b, i.e.
// d_b = sum_a r_a(b) d_a per bond (field_from_sums, d_a unscaled). `all`:
// the colour-0 planes, indexed over the whole replica. A wavefront lies in
// one atom type's block except where it straddles two (one in C/64 at most);
// then each lane reads its own rank's bonds.
__device__ __forceinline__ f3 hk_field(const Lad& L, const Rep& R, const Planes& all, const HkSite& t)
{
    const HkSteps st = hk_steps(R, t);
    const uint32_t r0 = __builtin_amdgcn_readfirstlane(t.rank);
    if (__ballot(t.rank != r0) == 0ull) return hk_field_rank(L, R, all, t, st, r0);
    return hk_field_rank(L, R, all, t, st, t.rank);
}

// FP64 energy of site t (spin s) over its bonds to higher storage ranks
// (every bond once: a bond never joins two atoms of one type):
// -J S.T + d_b.(S x T), d_b = sum_a r_a(b) d_a.
__device__ __forceinline__ double hk_bond_energy(const Lad& L, const Rep& R, const Planes& all, const HkSite& t,
                                                 double sx, double sy, double sz)
{
    double e = 0.0;
    for (int b = 0; b < kHkBonds; ++b) {
        const HkBond& B = c_hk.bond[t.rank][b];
        if ((uint32_t)B.rank < t.rank) continue;
        uint32_t i;
        if (!hk_nbr(R, t, B, i)) continue;
        const f3 tf = ld3(all, i);
        const double tx = tf.x, ty = tf.y, tz = tf.z;
        const double dbx = B.rd[0] * L.dd[0] + B.rd[1] * L.dd[3] + B.rd[2] * L.dd[6];
        const double dby = B.rd[0] * L.dd[1] + B.rd[1] * L.dd[4] + B.rd[2] * L.dd[7];
        const double dbz = B.rd[0] * L.dd[2] + B.rd[1] * L.dd[5] + B.rd[2] * L.dd[8];
        const double cx = sy * tz - sz * ty, cy = sz * tx - sx * tz, cz = sx * ty - sy * tx;
        e += -L.Jd * (sx * tx + sy * ty + sz * tz) + (dbx * cx + dby * cy + dbz * cz);
    }
    return e;
}

__device__ __forceinline__ f3 local_field(const Rep& R, const Lad& L, const Planes& oth, const Nbr& nb)
{
    f3 v[6];
    load_nbrs(R, oth, nb, v);
    return field_from(L, v);
}

// dE = -(S' - S).h - K (S'_z - S_z)(S'_z + S_z): the difference is formed
// first so small moves keep their relative accuracy.
__device__ __forceinline__ float delta_e(const f3& s, const f3& sn, const f3& h, float K)
{
    const float ddx = sn.x - s.x, ddy = sn.y - s.y, ddz = sn.z - s.z;
    return -(ddx * h.x + ddy * h.y + ddz * h.z) - K * ddz * (sn.z + s.z);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Wavefront sums of four values at once: the first level hands two of the
// four values to the partner lane (l ^ 32) and keeps the other two, the
// second hands one of those (l ^ 16), and the last four levels reduce one
// value over 16 lanes: 7 exchanges and 7 FP64 adds instead of 4 x 6. Sum k
// ends in lanes 16k..16k+15; write_wave_sum4 stores it from lane 16k.
__device__ __forceinline__ double wave_sum4(const double (&v)[4])
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const bool h32 = lane & 32u, h16 = lane & 16u;
    const double k0 = h32 ? v[2] : v[0], k1 = h32 ? v[3] : v[1];
    const double g0 = h32 ? v[0] : v[2], g1 = h32 ? v[1] : v[3];
    const double s0 = k0 + __shfl_xor(g0, 32, kWave);
    const double s1 = k1 + __shfl_xor(g1, 32, kWave);
    double t = (h16 ? s1 : s0) + __shfl_xor(h16 ? s0 : s1, 16, kWave);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
    return t;
}

// o[k] = sum k of wave_sum4 (its result t in every lane).
__device__ __forceinline__ void write_wave_sum4(double* __restrict__ o, double t)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if ((lane & 15u) == 0u) o[lane >> 4] = t;
}


// ---------------------------------------------------------------------------
// Two spin buffers of equal layout. Half-sweep launches update the current
// buffer in place; a fused sweep reads the current buffer and writes every
// site to the other one, which then becomes current. The current buffer is
// the parity of a device counter of fused sweeps, so launches captured in a
// graph stay valid on every replay.
struct SpinBuf {
    float* base;
    int64_t total;          // floats per buffer
    const uint32_t* ctr;    // device counter of fused sweeps
};

__device__ __forceinline__ float* spin_buffer(const SpinBuf& b, uint32_t fused_done)
{
    return b.base + (int64_t)(fused_done & 1u) * b.total;
}

// Site number (ix*ly + iy)*lz + iz of the Philox counter, in 24-bit
// multiplies when the replica has fewer than 2^24 sites (on bcc this is the
// cell number; the site is 2*cell + atom).
__device__ __forceinline__ uint32_t site_number(const Rep& R, uint32_t ix, uint32_t iy, uint32_t iz)
{
    return R.n < (1 << 24) ? __umul24(__umul24(ix, (uint32_t)R.ly) + iy, (uint32_t)R.lz) + iz
                           : (ix * (uint32_t)R.ly + iy) * (uint32_t)R.lz + iz;
}

// One Metropolis proposal for a site with spin s, neighbour spins nbr and
// local field hf. Proposal: u uniform in a ball of radius r (cos(theta) =
// 2U1-1, phi = 2 pi U2, radius r U3^(1/3)), S' = (S+u)/|S+u|; accept iff
// U4 < exp(-beta dE). One Philox block keyed by (seed, slot) with counter
// (site, sweep) supplies the four uniforms. Returns the spin after the
// decision.
//
// kLearned: ladders with the learned flag take r from the network of
// mlp.hpp evaluated on the six neighbour spins and beta*h (never the site's
// own spin), so r is the same for S -> S' and S' -> S. Every lane of the
// wavefront must call this (lanes past the end of a replica on a clamped
// valid site, active = false) because the network exchanges data across
// lanes.
template <bool kLearned>
__device__ __forceinline__ f3 metropolis_site(const Lad& L, float beta, const f3& s, const f3 (&nbr)[6], const f3& hf,
                                             uint32_t site, uint32_t sweep, uint32_t seed_lo, uint32_t seed_hi,
                                             uint32_t slot, const MlpDev* __restrict__ mlp, bool active, bool& accepted)
{
    // log2(radius) - 8: the radius enters the proposal as
    // r U3^(1/3) = 2^(log2(x)/3 + log2(r) - 8) for U3 = x 2^-24.
    float lr8 = L.lr8;
    if constexpr (kLearned) {
        if (L.learned) {      // uniform per workgroup (one replica)
            const float x[kMlpIn] = { nbr[0].x, nbr[0].y, nbr[0].z, nbr[1].x, nbr[1].y, nbr[1].z,
                                      nbr[2].x, nbr[2].y, nbr[2].z, nbr[3].x, nbr[3].y, nbr[3].z,
                                      nbr[4].x, nbr[4].y, nbr[4].z, nbr[5].x, nbr[5].y, nbr[5].z,
                                      beta * hf.x, beta * hf.y, beta * hf.z };
            const float f = mlp_forward<false>(*mlp, x, nullptr);
            lr8 = fmaf(clamp_log_radius(f), 1.4426950408889634f, -8.f);
        }
    }
    const u32x4 rnd = philox4x32_10(u32x4{ site, sweep, kStreamMetropolis, seed_hi }, seed_lo, slot);
    const float ct = s24f(rnd.x);
    const float uphi = u24f(rnd.y);
    const float uacc = u24f(rnd.w);
    const float st = __builtin_amdgcn_sqrtf(fmaxf(0.f, (1.f - ct) * (1.f + ct)));
    const float sp = __builtin_amdgcn_sinf(uphi);   // sin(2 pi uphi)
    const float cp = __builtin_amdgcn_cosf(uphi);   // cos(2 pi uphi)
    const float rr = __builtin_amdgcn_exp2f(fmaf(__builtin_amdgcn_logf((float)(rnd.z >> 8)), 1.f / 3.f, lr8));
    const f3 t{ s.x + rr * st * cp, s.y + rr * st * sp, s.z + rr * ct };
    const float n2 = t.x * t.x + t.y * t.y + t.z * t.z;
    const float inv = __builtin_amdgcn_rsqf(n2);
    const f3 sn{ t.x * inv, t.y * inv, t.z * inv };
    const float de = delta_e(s, sn, hf, L.Kz);
    const float w = __builtin_amdgcn_exp2f(-beta * de * 1.4426950408889634f);
    accepted = active && n2 > 1e-30f && uacc < w;
    return accepted ? sn : s;
}

// Demagnetising field of the site at colour-major index t (colour*half + q),
// mu0 H from the FP64 per-site field in the spin layout, added to the local
// field h (src/energyMin.f90:113-116,163-166).
__device__ __forceinline__ void add_demag(const Lad& L, const Rep& R, const double* __restrict__ demag,
                                          uint32_t t, f3& hf)
{
    const double* d = demag + R.off + t;
    hf.x += (float)(L.mu0 * d[0]);
    hf.y += (float)(L.mu0 * d[R.n]);
    hf.z += (float)(L.mu0 * d[2 * R.n]);
}

// Energy of a colour-1 site after the sweep and of its on-site partner: its
// six bonds, -S.(h - Bz z) with h formed from the updated colour-0
// neighbours (each bond has exactly one colour-1 end), plus the on-site
// terms of itself and of the colour-0 site c0 at the same half index.
__device__ __forceinline__ float pair_energy(const Lad& L, const f3& fs, const f3& hf, const f3& c0)
{
    return -(fs.x * hf.x + fs.y * hf.y + fs.z * (hf.z - L.Bz))
           - L.zsf * L.Bz * (fs.z + c0.z) - L.Kz * (fs.z * fs.z + c0.z * c0.z);
}

// The same energy in FP64 from FP32 spins (the exact energy of the swap
// exponent): the site's bonds -J S.sum_j T_j + sum_a d_a.(S x A_a), with
// sum and the signed sums A_a of the neighbour spins formed in FP64, plus the
// on-site terms of the site and of its colour-0 partner c0. d_a.(S x A_a)
// keeps only the nonzero components of d_a (one for bulk DMI, one per
// in-plane axis for the reference's interfacial form).
__device__ __forceinline__ double pair_energy_exact(const Lad& L, const f3& s, const double (&sum)[3],
                                                    const double (&A)[3][3], const f3& c0)
{
    const double sx = s.x, sy = s.y, sz = s.z;
    double e = -L.Jd * (sx * sum[0] + sy * sum[1] + sz * sum[2]);
    if (L.dmi == ISING_DMI_BULK) {
        e += L.dd[0] * (sy * A[0][2] - sz * A[0][1]) + L.dd[4] * (sz * A[1][0] - sx * A[1][2])
             + L.dd[8] * (sx * A[2][1] - sy * A[2][0]);
    } else if (L.dmi == ISING_DMI_INTERFACIAL_REF) {
        e += L.dd[1] * (sz * A[0][0] - sx * A[0][2]) + L.dd[3] * (sy * A[1][2] - sz * A[1][1]);
    }
    const double cz = c0.z;
    return e - L.zs * L.Bzd * (sz + cz) - L.Kzd * (sz * sz + cz * cz);
}

// FP64 neighbour sums of a simple-cubic site (v: +x, -x, +y, -y, +z, -z,
// zero across open faces): sum_j T_j and A_a = T_{+a} - T_{-a}.
__device__ __forceinline__ void sc_sums64(const f3 (&v)[6], double (&sum)[3], double (&A)[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const f3 p = v[2 * a], m = v[2 * a + 1];
        A[a][0] = (double)p.x - (double)m.x;
        A[a][1] = (double)p.y - (double)m.y;
        A[a][2] = (double)p.z - (double)m.z;
    }
    sum[0] = (((double)v[0].x + (double)v[1].x) + ((double)v[2].x + (double)v[3].x)) + ((double)v[4].x + (double)v[5].x);
    sum[1] = (((double)v[0].y + (double)v[1].y) + ((double)v[2].y + (double)v[3].y)) + ((double)v[4].y + (double)v[5].y);
    sum[2] = (((double)v[0].z + (double)v[1].z) + ((double)v[2].z + (double)v[3].z)) + ((double)v[4].z + (double)v[5].z);
}

// FP64 neighbour sums of a bcc site: the eight bonds of bcc_sums, reloaded.
__device__ __forceinline__ void bcc_sums64(const Rep& R, const Planes& oth, uint32_t q, uint32_t ix, uint32_t iy,
                                           uint32_t iz, int colour, double (&sum)[3], double (&A)[3][3])
{
    const BccSteps st = bcc_steps(R, ix, iy, iz, colour);
    const double sg = colour == 0 ? 1.0 : -1.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) { sum[c] = 0.0; A[0][c] = 0.0; A[1][c] = 0.0; A[2][c] = 0.0; }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (!bcc_ok(st, b)) continue;
        const f3 t = ld3(oth, bcc_index(st, q, b));
        const double v[3] = { t.x, t.y, t.z };
        const double gx = ((b >> 2) & 1) ? -sg : sg, gy = ((b >> 1) & 1) ? -sg : sg, gz = (b & 1) ? -sg : sg;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sum[c] += v[c];
            A[0][c] += gx * v[c];
            A[1][c] += gy * v[c];
            A[2][c] += gz * v[c];
        }
    }
}

// ---------------------------------------------------------------------------
// Half-sweep: one Metropolis proposal for every site of one colour of every
// replica, one thread per site, in place in the current buffer (accepted
// moves are stored).
//
// kThree: the variant launched in contexts with a three-colour (hyper-kagome)
// replica: it carries that cell's branch and skips colours a replica does
// not have.
// kMeasure (colour-1 launches only): per-site pair_energy and spins in FP32,
// summed in FP64 per wavefront into wave_part[wave][4] = (E, Mx, My, Mz);
// with kExact the per-site energies are FP64 (pair_energy_exact), the exact
// energies the replica exchange takes.
// The learned measuring variant is held to 96 registers (five wavefronts per
// SIMD, as the plain learned variant): left alone it takes 98, which the
// 8-register allocation rounds to 104, four wavefronts and a 4% slower launch;
// the FP64 (kExact) one would spill at 96.
template <bool kMeasure, bool kLearned, bool kExact = false, bool kThree = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kLearned && kMeasure && !kExact ? 5 : 1)))
void k_half_sweep(
    SpinBuf sb, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const float* __restrict__ beta_slot, const int32_t* __restrict__ block_rep,
    const uint32_t* __restrict__ sweep_ctr, uint32_t sweep_off, int colour,
    uint32_t seed_lo, uint32_t seed_hi, unsigned long long* __restrict__ acc_count,
    double* __restrict__ wave_part, uint32_t tune, const MlpDev* __restrict__ mlp,
    const double* __restrict__ demag)
{
    const uint32_t blk = logical_block(blockIdx.x, gridDim.x, tune);
    const int slot = block_rep[blk];
    const Rep R = reps[slot];
    if constexpr (kThree)
        if (colour >= R.ncol) return;   // the third colour's launch, on a two-colour replica (uniform)
    const Lad L = lads[R.ladder];
    const float beta = beta_slot[slot];
    const uint32_t sweep = *sweep_ctr + sweep_off;
    float* spins = spin_buffer(sb, *sb.ctr);
    const uint32_t q0 = (blk - (uint32_t)R.first_block) * kBlock + threadIdx.x;
    const bool active = q0 < R.half;
    const uint32_t q = active ? q0 : R.half - 1u;
    const bool nt = tune & kTuneNtOwn;

    const Planes own = planes(spins, R, colour);
    const Planes oth = planes(spins, R, colour ^ 1);
    const uint32_t row = fdiv(q, R.div_hx);
    const uint32_t h = fmodq(q, row, R.div_hx);
    const uint32_t iz = fdiv(row, R.div_ly);
    const uint32_t iy = fmodq(row, iz, R.div_ly);
    const f3 s = nt ? f3{ ldf_nt(own.x, q), ldf_nt(own.y, q), ldf_nt(own.z, q) } : ld3(own, q);
    f3 nbr[6];
    f3 hf;
    uint32_t site;
    if (R.cell == ISING_CELL_BCC) {      // uniform per workgroup; never a learned ladder
        hf = bcc_field(L, bcc_sums(R, oth, q, h, iy, iz, colour));
#pragma unroll
        for (int k = 0; k < 6; ++k) nbr[k] = f3{ 0.f, 0.f, 0.f };
        site = 2u * site_number(R, h, iy, iz) + (uint32_t)colour;
    } else if (kThree && R.cell == ISING_CELL_HYPERKAGOME) {    // uniform; never learned, never measured here
        const HkSite t = hk_site(R, q, colour);
        hf = hk_field(L, R, planes(spins, R, 0), t);
#pragma unroll
        for (int k = 0; k < 6; ++k) nbr[k] = f3{ 0.f, 0.f, 0.f };
        site = (uint32_t)kHkAtoms * site_number(R, t.ix, t.iy, t.iz) + (uint32_t)c_hk.type_of_rank[t.rank];
    } else {
        const uint32_t p = (iy + iz + (uint32_t)colour) & 1u;
        const uint32_t ix = 2u * h + p;
        load_nbrs(R, oth, neighbours<false>(R, q, h, iy, iz, p), nbr);
        hf = field_from(L, nbr);
        site = site_number(R, ix, iy, iz);
    }
    // uniform per workgroup; a context with the learned proposal has no
    // ladder with the demagnetisation term (the API refuses the mix)
    if constexpr (!kLearned)
        if (L.demag) add_demag(L, R, demag, (uint32_t)colour * R.half + q, hf);
    bool accepted;
    const f3 fs = metropolis_site<kLearned>(L, beta, s, nbr, hf, site, sweep, seed_lo, seed_hi,
                                            (uint32_t)slot, mlp, active, accepted);
    if (accepted) {
        if (nt) {
            stf_nt((float*)own.x, q, fs.x); stf_nt((float*)own.y, q, fs.y); stf_nt((float*)own.z, q, fs.z);
        } else {
            stf((float*)own.x, q, fs.x); stf((float*)own.y, q, fs.y); stf((float*)own.z, q, fs.z);
        }
    }
    if constexpr (kMeasure) {
        double me = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        if (active) {
            const f3 c0 = ld3(planes(spins, R, 0), q);
            if constexpr (kExact) {
                double sum[3], A[3][3];
                if (R.cell == ISING_CELL_BCC) bcc_sums64(R, oth, q, h, iy, iz, colour, sum, A);
                else sc_sums64(nbr, sum, A);
                me = pair_energy_exact(L, fs, sum, A, c0);
                mx = (double)fs.x + (double)c0.x; my = (double)fs.y + (double)c0.y; mz = (double)fs.z + (double)c0.z;
            } else {
                me = pair_energy(L, fs, hf, c0);
                mx = fs.x + c0.x; my = fs.y + c0.y; mz = fs.z + c0.z;
            }
        }
        const double v[4] = { me, mx, my, mz };
        write_wave_sum4(wave_part + 4 * ((size_t)blk * (kBlock / kWave) + threadIdx.x / kWave), wave_sum4(v));
    }
    // Acceptance count: each wavefront owns one counter (stream-ordered
    // launches never overlap, so a plain read-modify-write is race free).
    const unsigned long long ball = __ballot(accepted);
    if ((threadIdx.x & (kWave - 1)) == 0 && ball) {
        unsigned long long* c = acc_count + (size_t)blk * (kBlock / kWave) + threadIdx.x / kWave;
        *c += (unsigned long long)__popcll(ball);
    }
}

// ---------------------------------------------------------------------------
// Fused sweep: both colours of every site in one launch, marching along z.
// One workgroup per (replica, chunk of lc z planes [z0, z1)). Thread t owns
// the in-plane half indices j = t + m T (m < kFusedPasses) of every plane,
// i.e. the pair of columns ix = 2h, 2h + 1 at row iy (j = iy hx + h): on each
// plane one of the two sites is colour 0 and the other colour 1. Step s
// (z0-1 <= s <= z1) updates colour 0 on plane s, then colour 1 on plane s-1:
//   * the z neighbours of both updates are the other colour on the planes
//     above and below in the same column, which the owning thread keeps in
//     registers (colour 1 before the sweep on planes s-1 .. s+1, colour 0
//     after its update on planes s-2 .. s), and so is one x neighbour (the
//     other site of the column pair);
//   * the other x neighbour and both y neighbours come from LDS (16 bytes a
//     site, one access each): colour 1 of plane s before the sweep (for the
//     colour-0 update) and colour 0 of plane s-1 after its update (for the
//     colour-1 update), each double-buffered by plane parity, so a step has
//     one barrier; a neighbour across an open face reads a slot that holds
//     zeros;
//   * the planes the next step needs are loaded one step ahead, through buffer
//     descriptors whose scalar offsets select colour, component and plane (the
//     per-thread vector offset never changes).
// The plain and E/M variants rename their register windows (the z loop is
// unrolled by four); the FP64 swap-energy variant copies them after each
// step, which keeps it within the 128 registers that 16 wavefronts per CU
// allow.
// Colour-0 planes z0-1 and z1 are halo: recomputed from the same Philox
// counters as by their owners and not written. Reads come from the source
// buffer and every owned site is written to the other one, so neighbouring
// chunks never race. The neighbour sums, arithmetic and counters are those
// of the half sweeps, so the chains are bit-identical.
// kMeasure: per-thread FP64 sums of pair_energy and spins over the owned
// colour-1 sites, then per wavefront into wave_part; with kExact the
// per-site energies are pair_energy_exact of the same six neighbours, in the
// half sweeps' bond order (the energies a replica exchange takes).
constexpr int kFusedMaxThreads = 1024;
constexpr int kFusedPasses = 2;          // column pairs per thread on a plane
constexpr int kFusedMaxPairs = kFusedMaxThreads * kFusedPasses;   // column pairs of one workgroup's window
constexpr int kSlabHalo = 4;             // window rows of a y slab beyond the rows it owns
constexpr int kBufferFormat = 0x00020000; // buffer descriptor word 3: 32-bit data
// A vector offset at or past a descriptor's byte range: loads through it
// return zero and read no memory (every fused-sweep descriptor spans less
// than 2^31 bytes).
constexpr uint32_t kBufferOob = 0x80000000u;
// Cache policy of the source reads: non-temporal (nt). Each source plane is
// read once per sweep (twice at a chunk edge) while the sweep writes as many
// bytes to the other buffer; streaming the reads made the fused sweep 8%
// faster at 64^3 x 64 than the default policy (non-temporal writes did not
// help).
constexpr int kSweepLoadPolicy = 2;

__device__ __forceinline__ f3 bld3(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t s0, uint32_t sN)
{
    return f3{ __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)s0, kSweepLoadPolicy)),
               __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)(s0 + sN), kSweepLoadPolicy)),
               __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)(s0 + 2u * sN),
                                                                     kSweepLoadPolicy)) };
}
__device__ __forceinline__ void bst3(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t s0, uint32_t sN, const f3& v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.x), r, (int)voff, (int)s0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.y), r, (int)voff, (int)(s0 + sN), 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.z), r, (int)voff, (int)(s0 + 2u * sN), 0);
}
__device__ __forceinline__ f3 flip_sign(const f3& v, uint32_t mask)
{
    return f3{ __uint_as_float(__float_as_uint(v.x) ^ mask), __uint_as_float(__float_as_uint(v.y) ^ mask),
               __uint_as_float(__float_as_uint(v.z) ^ mask) };
}

template <bool kMeasure, bool kExact = false, bool kSlab = false>
__global__ __launch_bounds__(kFusedMaxThreads) void k_sweep_fused(
    SpinBuf sb, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const float* __restrict__ beta_slot, const int32_t* __restrict__ fblock_rep,
    const uint32_t* __restrict__ sweep_ctr, uint32_t sweep_off, int lc, uint32_t pstride,
    uint32_t seed_lo, uint32_t seed_hi, unsigned long long* __restrict__ acc_count,
    double* __restrict__ wave_part, uint32_t tune)
{
    extern __shared__ float4 lds[];   // [colour 2][parity 2][pstride] of (x, y, z, unused); slot PW of each is zero
    const uint32_t blk = logical_block(blockIdx.x, gridDim.x, tune);
    const int slot = fblock_rep[blk];
    const Rep R = reps[slot];
    const Lad L = lads[R.ladder];
    const float beta = beta_slot[slot];
    const uint32_t sweep = *sweep_ctr + sweep_off;
    const uint32_t buf = *sb.ctr + sweep_off;       // every sweep of the period is fused
    const uint32_t P = (uint32_t)R.plane;
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    const uint32_t hx = (uint32_t)R.hx, ly = (uint32_t)R.ly;
    const int lz = R.lz;
    const bool pz = (R.per >> 2) & 1;
    const bool px = R.per & 1, py = (R.per >> 1) & 1;
    // workgroup bi of the slot: z chunk bi / nslab, y slab bi % nslab
    const int bi = (int)blk - R.f_first_block;
    const int nslab = kSlab ? R.f_nslab : 1;
    const int zc = kSlab ? bi / nslab : bi;
    const int z0 = zc * lc;
    const int z1 = min(z0 + lc, lz);
    // The window: LDS rows of one plane, PW column pairs, slot PW zero. A
    // slab owning rows [y0, y0 + yown) reads rows y0-2 .. y0+hy+1 (global row
    // of window row r: y0 - 2 + r, periodic image or none beyond an open
    // face), updates colour 0 on rows y0-1 .. y0+hy as well (the halo rows
    // its colour-1 updates need, recomputed from the same counters as by
    // their owners) and writes only its own rows.
    const uint32_t PW = kSlab ? (uint32_t)R.f_rows * hx : P;
    bool slab = false;
    int Y0 = 0, yown = (int)ly;
    if constexpr (kSlab) {
        slab = nslab > 1;
        const int y0 = (bi - zc * nslab) * R.f_hy;
        yown = min(R.f_hy, (int)ly - y0);
        Y0 = slab ? y0 - 2 : 0;
    }
    const bool small = R.n < (1 << 24);
    // the replica's three component blocks in the source and destination
    // buffers; byte offsets: component c at c*sN, colour 1 at sH, plane z at z*sP
    const uint32_t sN = (uint32_t)R.n * 4u, sH = R.half * 4u, sP = P * 4u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(spin_buffer(sb, buf) + R.off), 0, (int)(3u * sN), kBufferFormat);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(spin_buffer(sb, buf + 1u) + R.off), 0, (int)(3u * sN), kBufferFormat);

    // Plane z of the lattice (periodic image), or -1 beyond an open face.
    auto wrapz = [&](int z) -> int {
        if (z < 0) return pz ? z + lz : -1;
        if (z >= lz) return pz ? z - lz : -1;
        return z;
    };
    auto lplane = [&](int c, int q) -> float4* { return lds + (size_t)(c * 2 + q) * pstride; };
    if (tid < 4u) lds[(size_t)tid * pstride + PW] = float4{ 0.f, 0.f, 0.f, 0.f };    // the zero slots

    // Per owned column pair: its index, row and the LDS indices of its
    // in-plane neighbours (the +x one of an odd site and the -x one of an even
    // site, +-y; the zero slot across an open face). A lane past the end of
    // the plane takes the last column pair: it reads, updates and writes that
    // pair exactly as its owner does (same inputs, same counters), so its
    // writes repeat the owner's values and the kernel needs no per-lane
    // branches; only the acceptance counts and the sums skip it.
    // kSlab: vo / so are the byte offsets of the pair in the spin planes for
    // loads (past the range for rows beyond an open face, which read zeros)
    // and stores (past the range for rows the slab does not own, whose
    // stores are dropped); own: counted and summed; okr: a lattice row.
    uint32_t jj[kFusedPasses], iyv[kFusedPasses], xnp[kFusedPasses], xnm[kFusedPasses], ynp[kFusedPasses],
        ynm[kFusedPasses], sb0[kFusedPasses], sb1[kFusedPasses];
    bool act[kFusedPasses];
    [[maybe_unused]] uint32_t vo[kFusedPasses], so[kFusedPasses];
    [[maybe_unused]] bool own[kFusedPasses], okr[kFusedPasses];
#pragma unroll
    for (int m = 0; m < kFusedPasses; ++m) {
        const uint32_t k = (uint32_t)m * T + tid;
        act[m] = k < PW;
        const uint32_t j = act[m] ? k : PW - 1u;
        jj[m] = j;
        const uint32_t r = fdiv(j, R.div_hx);       // window row (the lattice row on whole planes)
        const uint32_t h = fmodq(j, r, R.div_hx);
        uint32_t iy = r;
        if constexpr (kSlab) {
            vo[m] = j * 4u; so[m] = j * 4u; own[m] = act[m]; okr[m] = true;
            if (slab) {
                int gy = Y0 + (int)r;
                bool ok = true;
                if (gy < 0) { gy += (int)ly; ok = py; }
                else if (gy >= (int)ly) { gy -= (int)ly; ok = py; }
                iy = (uint32_t)gy;
                const bool mine = (int)r >= 2 && (int)r < 2 + yown;
                vo[m] = ok ? (iy * hx + h) * 4u : kBufferOob;
                so[m] = mine ? (iy * hx + h) * 4u : kBufferOob;
                own[m] = mine && act[m];
                okr[m] = ok;
            }
        }
        iyv[m] = iy;
        const bool xpw = h == hx - 1u, xmw = h == 0u;
        xnp[m] = xpw ? (px ? j + 1u - hx : PW) : j + 1u;
        xnm[m] = xmw ? (px ? j - 1u + hx : PW) : j - 1u;
        if (slab) {                  // uniform; window rows 0 and W-1 have no neighbour beyond
            ynp[m] = j + hx < PW ? j + hx : PW;
            ynm[m] = r > 0u ? j - hx : PW;
        } else {
            const bool ypw = iy + 1u == ly, ymw = iy == 0u;
            ynp[m] = ypw ? (py ? j + hx - P : P) : j + hx;
            ynm[m] = ymw ? (py ? j - hx + P : P) : j - hx;
        }
        // site number (ix*ly + iy)*lz of ix = 2h and 2h + 1, without the plane
        sb0[m] = small ? __umul24(__umul24(2u * h, ly) + iy, (uint32_t)lz) : ((2u * h) * ly + iy) * (uint32_t)lz;
        sb1[m] = small ? __umul24(__umul24(2u * h + 1u, ly) + iy, (uint32_t)lz) : ((2u * h + 1u) * ly + iy) * (uint32_t)lz;
    }
    // Colour `colour` of plane z in the source buffer. A plane beyond an open
    // face, or one the chunk never reads (colour 1 above z1 + 1, colour 0
    // above z1), is read through an offset past the descriptor's range, which
    // returns zeros without a branch or a memory access.
    auto load = [&](int colour, int z, int m) -> f3 {
        const int zz = z > z1 + colour ? -1 : wrapz(z);      // uniform
        return bld3(rs, (kSlab ? vo[m] : jj[m] * 4u) | (zz >= 0 ? 0u : kBufferOob),
                    zz >= 0 ? (colour ? sH : 0u) + (uint32_t)zz * sP : 0u, sN);
    };
    auto store_off = [&](int m) -> uint32_t { return kSlab ? so[m] : jj[m] * 4u; };
    auto counted = [&](int m) -> bool { return kSlab ? own[m] : act[m]; };
    // one 16-byte LDS access per site: a single address per neighbour
    auto put = [&](float4* b, int m, const f3& v) { b[jj[m]] = float4{ v.x, v.y, v.z, 0.f }; };
    auto get = [&](const float4* b, uint32_t i) -> f3 {
        const float4 t = b[i];
        return f3{ t.x, t.y, t.z };
    };

    // Register windows: colour 1 before the sweep on planes s-1, s, s+1 and
    // the plane s+2 loaded during step s; colour 0 after its update on planes
    // s-2, s-1, before it (then after it) on plane s, and plane s+1 loaded
    // during step s. Four slots each. The plain and E/M sweeps rotate them by
    // renaming (a loop body of four steps whose window arguments shift by one
    // slot, so no step copies a window: 109 and 117 registers); the FP64
    // swap-energy variant copies them after every step, which keeps it under
    // the 128 registers that 16 wavefronts per CU allow (renamed, it spills).
    f3 A[4][kFusedPasses], B[4][kFusedPasses];
#pragma unroll
    for (int m = 0; m < kFusedPasses; ++m) {
        A[0][m] = load(1, z0 - 2, m);
        A[1][m] = load(1, z0 - 1, m);
        A[2][m] = load(1, z0, m);
        B[0][m] = f3{ 0.f, 0.f, 0.f };
        B[1][m] = f3{ 0.f, 0.f, 0.f };
        B[2][m] = load(0, z0 - 1, m);
    }

    double me = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    uint32_t n_acc = 0;        // wavefront-uniform
    using Win = f3[kFusedPasses];
    auto step = [&](int s, Win& am1, Win& a0, Win& ap1, Win& na, Win& bm2, Win& bm1, Win& b0,
                    Win& nb) __attribute__((always_inline)) {
        const int zs = wrapz(s);
        const int q = s & 1;
        float4* l1 = lplane(1, q);         // colour 1 of plane s, before the sweep
        float4* l0 = lplane(0, q);         // colour 0 of plane s, after its update
        const float4* l0p = lplane(0, q ^ 1);   // colour 0 of plane s-1, after its update
#pragma unroll
        for (int m = 0; m < kFusedPasses; ++m)
            put(l1, m, a0[m]);
        __syncthreads();
        // the planes step s+1 adds to the windows
#pragma unroll
        for (int m = 0; m < kFusedPasses; ++m) {
            na[m] = load(1, s + 2, m);
            nb[m] = load(0, s + 1, m);
        }
        // colour 0 on plane s
        const bool owned0 = s >= z0 && s < z1;
#pragma unroll
        for (int m = 0; m < kFusedPasses; ++m) {
            f3 fs = f3{ 0.f, 0.f, 0.f };
            bool accepted = false;
            if (zs >= 0) {             // uniform per workgroup
                const uint32_t p = (iyv[m] + (uint32_t)zs) & 1u;
                // +x and -x: the own pair's colour-1 site and the LDS one;
                // their sum, and their difference with the sign of +x - -x
                const f3 own = a0[m], oth = get(l1, p ? xnp[m] : xnm[m]);
                const f3 yp = get(l1, ynp[m]), ym = get(l1, ynm[m]);
                const f3 &zp = ap1[m], &zm = am1[m];
                const f3 sum{ (own.x + oth.x) + (yp.x + ym.x) + (zp.x + zm.x),
                              (own.y + oth.y) + (yp.y + ym.y) + (zp.y + zm.y),
                              (own.z + oth.z) + (yp.z + ym.z) + (zp.z + zm.z) };
                const f3 dx = flip_sign(f3{ own.x - oth.x, own.y - oth.y, own.z - oth.z }, p << 31);
                const f3 dy{ yp.x - ym.x, yp.y - ym.y, yp.z - ym.z };
                const f3 dz{ zp.x - zm.x, zp.y - zm.y, zp.z - zm.z };
                const f3 hf = field_from_sums(L, sum, dx, dy, dz);
                const f3 nbr[6] = {};      // (the learned proposal's input; never fused)
                fs = metropolis_site<false>(L, beta, b0[m], nbr, hf, (p ? sb1[m] : sb0[m]) + (uint32_t)zs,
                                            sweep, seed_lo, seed_hi, (uint32_t)slot, nullptr, true, accepted);
                if (owned0) bst3(rd, store_off(m), (uint32_t)zs * sP, sN, fs);
            }
            if (owned0) n_acc += (uint32_t)__popcll(__ballot(accepted && counted(m)));
            if constexpr (kSlab)        // a row beyond an open face stays zero as a neighbour
                put(l0, m, okr[m] ? fs : f3{ 0.f, 0.f, 0.f });
            else
                put(l0, m, fs);
            b0[m] = fs;                 // colour 0 of plane s, after its update
        }
        // colour 1 on plane s-1 (owned planes only; l0p was filled in step s-1)
        const int zb = s - 1;
        if (zb >= z0 && zb < z1) {     // uniform
#pragma unroll
            for (int m = 0; m < kFusedPasses; ++m) {
                const uint32_t p = (iyv[m] + (uint32_t)zb + 1u) & 1u;
                const f3 own = bm1[m], oth = get(l0p, p ? xnp[m] : xnm[m]);
                const f3 yp = get(l0p, ynp[m]), ym = get(l0p, ynm[m]);
                const f3 &zp = b0[m], &zm = bm2[m];
                const f3 sum{ (own.x + oth.x) + (yp.x + ym.x) + (zp.x + zm.x),
                              (own.y + oth.y) + (yp.y + ym.y) + (zp.y + zm.y),
                              (own.z + oth.z) + (yp.z + ym.z) + (zp.z + zm.z) };
                const f3 dx = flip_sign(f3{ own.x - oth.x, own.y - oth.y, own.z - oth.z }, p << 31);
                const f3 dy{ yp.x - ym.x, yp.y - ym.y, yp.z - ym.z };
                const f3 dz{ zp.x - zm.x, zp.y - zm.y, zp.z - zm.z };
                const f3 hf = field_from_sums(L, sum, dx, dy, dz);
                const f3 nbr[6] = {};
                bool accepted;
                const f3 fs = metropolis_site<false>(L, beta, am1[m], nbr, hf, (p ? sb1[m] : sb0[m]) + (uint32_t)zb,
                                                     sweep, seed_lo, seed_hi, (uint32_t)slot, nullptr, true, accepted);
                bst3(rd, store_off(m), sH + (uint32_t)zb * sP, sN, fs);
                if (counted(m)) {
                    if constexpr (kMeasure) {
                        const f3 c0 = bm1[m];
                        if constexpr (kExact) {
                            const f3 v[6] = { p ? oth : own, p ? own : oth, yp, ym, zp, zm };
                            double sum[3], A[3][3];
                            sc_sums64(v, sum, A);
                            me += pair_energy_exact(L, fs, sum, A, c0);
                            mx += (double)fs.x + (double)c0.x;
                            my += (double)fs.y + (double)c0.y;
                            mz += (double)fs.z + (double)c0.z;
                        } else {
                            me += (double)pair_energy(L, fs, hf, c0);
                            mx += (double)(fs.x + c0.x); my += (double)(fs.y + c0.y); mz += (double)(fs.z + c0.z);
                        }
                    }
                }
                n_acc += (uint32_t)__popcll(__ballot(accepted && counted(m)));
            }
        }
    };
    if constexpr (!kExact) {
        for (int s = z0 - 1; s <= z1; s += 4) {
            step(s, A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3]);
            if (s + 1 > z1) break;
            step(s + 1, A[1], A[2], A[3], A[0], B[1], B[2], B[3], B[0]);
            if (s + 2 > z1) break;
            step(s + 2, A[2], A[3], A[0], A[1], B[2], B[3], B[0], B[1]);
            if (s + 3 > z1) break;
            step(s + 3, A[3], A[0], A[1], A[2], B[3], B[0], B[1], B[2]);
        }
    } else {
        for (int s = z0 - 1; s <= z1; ++s) {
            step(s, A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3]);
#pragma unroll
            for (int m = 0; m < kFusedPasses; ++m) {
                A[0][m] = A[1][m]; A[1][m] = A[2][m]; A[2][m] = A[3][m];
                B[0][m] = B[1][m]; B[1][m] = B[2][m]; B[2][m] = B[3][m];
            }
        }
    }

    const uint32_t wpb = T / kWave, wave = tid / kWave;
    if constexpr (kMeasure) {
        const double v[4] = { me, mx, my, mz };
        write_wave_sum4(wave_part + 4 * ((size_t)blk * wpb + wave), wave_sum4(v));
    }
    if ((tid & (kWave - 1)) == 0 && n_acc) acc_count[(size_t)blk * wpb + wave] += n_acc;
}

// ---------------------------------------------------------------------------
// Training records from the current configurations: per replica slot,
// `per` sites drawn with the collect stream; record = the six neighbour
// spins, beta*h, S, h, beta, K (mlp.hpp layout).
__global__ __launch_bounds__(kBlock) void k_mlp_collect(
    const float* __restrict__ spins, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const float* __restrict__ beta_slot, int R, int per, uint32_t call, uint32_t seed_lo, uint32_t seed_hi,
    float* __restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= R * per) return;
    const int slot = e / per, k = e - slot * per;
    const Rep Rr = reps[slot];
    const Lad L = lads[Rr.ladder];
    const float beta = beta_slot[slot];
    const u32x4 r = philox4x32_10(u32x4{ (uint32_t)k, call, kStreamCollect, seed_hi }, seed_lo, (uint32_t)slot);
    const uint32_t q = r.x % Rr.half;
    const int colour = (int)(r.y & 1u);
    const uint32_t row = fdiv(q, Rr.div_hx);
    const uint32_t h = q - row * (uint32_t)Rr.hx;
    const uint32_t iz = fdiv(row, Rr.div_ly);
    const uint32_t iy = row - iz * (uint32_t)Rr.ly;
    const uint32_t p = (iy + iz + (uint32_t)colour) & 1u;
    const Planes own = planes(spins, Rr, colour);
    const Planes oth = planes(spins, Rr, colour ^ 1);
    f3 nbr[6];
    load_nbrs(Rr, oth, neighbours<false>(Rr, q, h, iy, iz, p), nbr);
    const f3 hf = field_from(L, nbr);
    const f3 s = ld3(own, q);
    float* o = out + (size_t)e * kRecord;
#pragma unroll
    for (int j = 0; j < 6; ++j) { o[3 * j] = nbr[j].x; o[3 * j + 1] = nbr[j].y; o[3 * j + 2] = nbr[j].z; }
    o[18] = beta * hf.x; o[19] = beta * hf.y; o[20] = beta * hf.z;
    o[21] = s.x; o[22] = s.y; o[23] = s.z;
    o[24] = hf.x; o[25] = hf.y; o[26] = hf.z;
    o[27] = beta; o[28] = L.Kz; o[29] = 0.f; o[30] = 0.f; o[31] = 0.f;
}

// Sum of n counters over a workgroup of kBlock threads; valid in thread 0.
__device__ __forceinline__ unsigned long long block_sum_u64(const unsigned long long* __restrict__ a, int n,
                                                            unsigned long long* lds /* [kBlock] */)
{
    unsigned long long v = 0;
    for (int j = threadIdx.x; j < n; j += kBlock) v += a[j];
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
        __syncthreads();
    }
    return lds[0];
}

// Per-slot acceptance total: counters retired from earlier fused layouts,
// plus the per-wavefront counters of the half-sweep grid and of the fused
// grid (wpb_f wavefronts per fused workgroup).
__global__ __launch_bounds__(kBlock) void k_acc_finalize(
    const Rep* __restrict__ reps, const unsigned long long* __restrict__ acc_wave,
    const unsigned long long* __restrict__ acc_wave_f, int wpb_f,
    const unsigned long long* __restrict__ acc_base, unsigned long long* __restrict__ acc)
{
    __shared__ unsigned long long lds[kBlock];
    const Rep R = reps[blockIdx.x];
    const unsigned long long a = block_sum_u64(acc_wave + (size_t)R.first_block * (kBlock / kWave),
                                               R.nblocks * (kBlock / kWave), lds);
    __syncthreads();
    const unsigned long long b = block_sum_u64(acc_wave_f + (size_t)R.f_first_block * wpb_f, R.f_nblocks * wpb_f, lds);
    if (threadIdx.x == 0) acc[blockIdx.x] = acc_base[blockIdx.x] + a + b;
}

// Moves the fused-grid counters of each slot into acc_base (before the fused
// layout changes); the caller then clears acc_wave_f.
__global__ __launch_bounds__(kBlock) void k_acc_fold(
    const Rep* __restrict__ reps, const unsigned long long* __restrict__ acc_wave_f, int wpb_f,
    unsigned long long* __restrict__ acc_base)
{
    __shared__ unsigned long long lds[kBlock];
    const Rep R = reps[blockIdx.x];
    const unsigned long long b = block_sum_u64(acc_wave_f + (size_t)R.f_first_block * wpb_f, R.f_nblocks * wpb_f, lds);
    if (threadIdx.x == 0) acc_base[blockIdx.x] += b;
}

// ---------------------------------------------------------------------------
// Measurement series of one run: base pointers and capacities in a device
// descriptor, written before the run by a one-thread kernel, so launches
// captured in a graph stay valid when the series buffers are reallocated or
// placed elsewhere. e: energies [cap_e][R]; m: magnetisation [cap_e][R][3];
// qr / qi: FP64 and integer Q [cap_q][R][zmax], indexed by temperature
// position.
struct SeriesDesc {
    double* e;
    double* m;
    double* qr;
    long long* qi;
    uint32_t cap_e, cap_q;
    uint32_t zrec;      // charge entries per (record, temperature): Zmax, or 1 for one layer
};

// One record of a series, written by the kernel that finishes the value:
// entry *ctr + local. ctr advances once per period (k_advance), local counts
// this series' records earlier in the period, so one captured period replays
// correctly. desc == nullptr: nothing is recorded.
struct SeriesRec {
    const SeriesDesc* desc;
    const uint32_t* ctr;
    uint32_t local;
};

// Temperature position of `slot` (the inverse of slot_of_temp), found by the
// whole workgroup; -1 if absent. Every thread must call it.
__device__ __forceinline__ int temp_of_slot(const int32_t* __restrict__ slot_of_temp, int R, int slot)
{
    __shared__ int tpos;
    if (threadIdx.x == 0) tpos = -1;
    __syncthreads();
    for (int t = threadIdx.x; t < R; t += blockDim.x)
        if (slot_of_temp[t] == slot) tpos = t;
    __syncthreads();
    return tpos;
}

// ---------------------------------------------------------------------------
// FP64 reductions.


// Sum of 4 values over a workgroup of kBlock threads; result valid in thread 0.
__device__ __forceinline__ void block_sum4(double v[4], double* lds /* [4][kBlock/kWave] */)
{
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) lds[k * (kBlock / kWave) + w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double a = 0.0;
            for (int j = 0; j < kBlock / kWave; ++j) a += lds[k * (kBlock / kWave) + j];
            v[k] = a;
        }
    }
}

// Per-site energy of the sites at half index q of every colour (each bond
// once: the +x, +y, +z bonds on SC, the bonds from atom 0 on bcc, the bonds
// to higher storage ranks on hyper-kagome) plus on-site terms, and the
// spins; per-workgroup partial sums.
__global__ __launch_bounds__(kBlock) void k_measure(
    SpinBuf sb, uint32_t sweep_off, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const Lad* __restrict__ lad_override, const int32_t* __restrict__ block_rep, int block_base,
    double* __restrict__ partial, const double* __restrict__ demag)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    const float* spins = spin_buffer(sb, *sb.ctr + sweep_off);
    const int b = blockIdx.x + block_base;
    const int slot = block_rep[b];
    const Rep R = reps[slot];
    const Lad& L = lad_override ? *lad_override : lads[R.ladder];
    const uint32_t q = (uint32_t)(b - R.first_block) * kBlock + threadIdx.x;
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    if (q < R.half && R.cell == ISING_CELL_BCC) {
        // Every bond once, from its atom-0 end (colour 0):
        // -J S_i.S_j + d_b.(S_i x S_j), d_b = sum_a sgn_a(b) d_a (d_a scaled
        // by 1/sqrt 3); on-site terms and spins of both atoms of the cell.
        const uint32_t row = fdiv(q, R.div_hx);
        const uint32_t ix = q - row * (uint32_t)R.hx;
        const uint32_t iz = fdiv(row, R.div_ly);
        const uint32_t iy = row - iz * (uint32_t)R.ly;
        const Planes p0 = planes(spins, R, 0), p1 = planes(spins, R, 1);
        const BccSteps st = bcc_steps(R, ix, iy, iz, 0);
        const f3 af = ld3(p0, q);
        const double sx = af.x, sy = af.y, sz = af.z;
        double e = 0.0;
        for (int b = 0; b < 8; ++b) {
            if (!bcc_ok(st, b)) continue;
            const f3 tf = ld3(p1, bcc_index(st, q, b));
            const double tx = tf.x, ty = tf.y, tz = tf.z;
            const double gx = ((b >> 2) & 1) ? -1.0 : 1.0, gy = ((b >> 1) & 1) ? -1.0 : 1.0, gz = (b & 1) ? -1.0 : 1.0;
            const double dbx = gx * L.dd[0] + gy * L.dd[3] + gz * L.dd[6];
            const double dby = gx * L.dd[1] + gy * L.dd[4] + gz * L.dd[7];
            const double dbz = gx * L.dd[2] + gy * L.dd[5] + gz * L.dd[8];
            const double cx = sy * tz - sz * ty, cy = sz * tx - sx * tz, cz = sx * ty - sy * tx;
            e += -L.Jd * (sx * tx + sy * ty + sz * tz) + (dbx * cx + dby * cy + dbz * cz);
        }
        e -= L.zs * L.Bzd * sz + L.Kzd * sz * sz;
        const f3 bf = ld3(p1, q);
        const double ux = bf.x, uy = bf.y, uz = bf.z;
        e -= L.zs * L.Bzd * uz + L.Kzd * uz * uz;
        if (L.demag) {      // -mu0/2 S.H of both atoms (src/energyMin.f90:223-226)
            const double* d = demag + R.off + q;
            e -= 0.5 * L.mu0 * (sx * d[0] + sy * d[R.n] + sz * d[2 * R.n]);
            d += R.half;
            e -= 0.5 * L.mu0 * (ux * d[0] + uy * d[R.n] + uz * d[2 * R.n]);
        }
        v[0] = e; v[1] = sx + ux; v[2] = sy + uy; v[3] = sz + uz;
    } else if (q < R.half && R.cell == ISING_CELL_HYPERKAGOME) {
        // The sites at index q of the three colours: each one's bonds to
        // higher storage ranks, its on-site terms and spin.
        const Planes all = planes(spins, R, 0);
        for (int colour = 0; colour < R.ncol; ++colour) {
            const HkSite t = hk_site(R, q, colour);
            const f3 sf = ld3(planes(spins, R, colour), q);
            const double sx = sf.x, sy = sf.y, sz = sf.z;
            double e = hk_bond_energy(L, R, all, t, sx, sy, sz);
            e -= L.zs * L.Bzd * sz + L.Kzd * sz * sz;
            if (L.demag) {
                const double* d = demag + R.off + (uint32_t)colour * R.half + q;
                e -= 0.5 * L.mu0 * (sx * d[0] + sy * d[R.n] + sz * d[2 * R.n]);
            }
            v[0] += e; v[1] += sx; v[2] += sy; v[3] += sz;
        }
    } else if (q < R.half) {
        const uint32_t row = fdiv(q, R.div_hx);
        const uint32_t h = q - row * (uint32_t)R.hx;
        const uint32_t iz = fdiv(row, R.div_ly);
        const uint32_t iy = row - iz * (uint32_t)R.ly;
        for (int colour = 0; colour < 2; ++colour) {
            const Planes own = planes(spins, R, colour);
            const Planes oth = planes(spins, R, colour ^ 1);
            const uint32_t p = (iy + iz + (uint32_t)colour) & 1u;
            const Nbr nb = neighbours<false>(R, q, h, iy, iz, p);
            const f3 sf = ld3(own, q);
            const double sx = sf.x, sy = sf.y, sz = sf.z;
            double e = 0.0;
            const int plus[3] = { 0, 2, 4 };
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int k = plus[a];
                if (!nb.ok[k]) continue;
                const f3 tf = ld3(oth, nb.i[k]);
                const double tx = tf.x, ty = tf.y, tz = tf.z;
                // d_a.(S x T): one component of S x T for bulk DMI (d_a = D e_a),
                // one for the reference's interfacial form (d_x = -D e_y,
                // d_y = D e_x, d_z = 0); the other terms of the dot are zeros
                double dm = 0.0;
                if (L.dmi == ISING_DMI_BULK)
                    dm = L.dd[4 * a] * (a == 0 ? sy * tz - sz * ty : a == 1 ? sz * tx - sx * tz : sx * ty - sy * tx);
                else if (L.dmi == ISING_DMI_INTERFACIAL_REF)
                    dm = a == 0 ? L.dd[1] * (sz * tx - sx * tz) : a == 1 ? L.dd[3] * (sy * tz - sz * ty) : 0.0;
                e += -L.Jd * (sx * tx + sy * ty + sz * tz) + dm;
            }
            e -= L.zs * L.Bzd * sz + L.Kzd * sz * sz;
            if (L.demag) {
                const double* d = demag + R.off + (uint32_t)colour * R.half + q;
                e -= 0.5 * L.mu0 * (sx * d[0] + sy * d[R.n] + sz * d[2 * R.n]);
            }
            v[0] += e; v[1] += sx; v[2] += sy; v[3] += sz;
        }
    }
    block_sum4(v, lds);
    if (threadIdx.x == 0) {
        double* o = partial + 4 * (int64_t)b;
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
    }
}

// One workgroup per slot: fixed-order FP64 sum of the per-wavefront
// measurement partials of the fused colour-1 half-sweep (deterministic).
// wpb_f = 0: partials of the half-sweep grid; else of the fused grid with
// wpb_f wavefronts per workgroup.
__device__ __forceinline__ void wave_finalize_slot(
    int slot, double* lds, const Rep* __restrict__ reps, const double* __restrict__ wave_part, int wpb_f,
    double* __restrict__ energy, double* __restrict__ mag, const int32_t* __restrict__ slot_of_temp, int nrep,
    const SeriesRec& rec)
{
    const Rep R = reps[slot];
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    const int nw = wpb_f ? R.f_nblocks * wpb_f : R.nblocks * (kBlock / kWave);
    const double* base = wave_part + 4 * (wpb_f ? (size_t)R.f_first_block * wpb_f
                                               : (size_t)R.first_block * (kBlock / kWave));
    for (int j = threadIdx.x; j < nw; j += kBlock) {
        v[0] += base[4 * j]; v[1] += base[4 * j + 1]; v[2] += base[4 * j + 2]; v[3] += base[4 * j + 3];
    }
    block_sum4(v, lds);
    const double inv = 1.0 / (double)R.n;
    if (threadIdx.x == 0) {
        energy[slot] = v[0];
        mag[3 * slot] = v[1] * inv; mag[3 * slot + 1] = v[2] * inv; mag[3 * slot + 2] = v[3] * inv;
    }
    if (rec.desc) {       // uniform
        const int t = temp_of_slot(slot_of_temp, nrep, slot);
        const uint32_t m = *rec.ctr + rec.local;
        const SeriesDesc& D = *rec.desc;
        if (threadIdx.x == 0 && m < D.cap_e && t >= 0) {
            const int64_t e = (int64_t)m * nrep + t;
            D.e[e] = v[0];
            D.m[3 * e] = v[1] * inv; D.m[3 * e + 1] = v[2] * inv; D.m[3 * e + 2] = v[3] * inv;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_wave_finalize(
    const Rep* __restrict__ reps, const double* __restrict__ wave_part, int wpb_f,
    double* __restrict__ energy, double* __restrict__ mag, const int32_t* __restrict__ slot_of_temp, int nrep,
    SeriesRec rec)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    wave_finalize_slot(blockIdx.x, lds, reps, wave_part, wpb_f, energy, mag, slot_of_temp, nrep, rec);
}

// One workgroup per slot: fixed-order sum of its partials (deterministic);
// with rec, the sums are also written to the series at the slot's
// temperature (measurements of three-colour sweeps).
__global__ __launch_bounds__(kBlock) void k_measure_finalize(
    const Rep* __restrict__ reps, const double* __restrict__ partial, int slot_base,
    double* __restrict__ energy, double* __restrict__ mag, const int32_t* __restrict__ slot_of_temp, int nrep,
    SeriesRec rec)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    const int slot = blockIdx.x + slot_base;
    const Rep R = reps[slot];
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (int j = threadIdx.x; j < R.nblocks; j += kBlock) {
        const double* p = partial + 4 * (int64_t)(R.first_block + j);
        v[0] += p[0]; v[1] += p[1]; v[2] += p[2]; v[3] += p[3];
    }
    block_sum4(v, lds);
    const double inv = 1.0 / (double)R.n;
    if (threadIdx.x == 0) {
        energy[slot] = v[0];
        mag[3 * slot] = v[1] * inv; mag[3 * slot + 1] = v[2] * inv; mag[3 * slot + 2] = v[3] * inv;
    }
    if (rec.desc) {       // uniform
        const int t = temp_of_slot(slot_of_temp, nrep, slot);
        const uint32_t m = *rec.ctr + rec.local;
        const SeriesDesc& D = *rec.desc;
        if (threadIdx.x == 0 && m < D.cap_e && t >= 0) {
            const int64_t e = (int64_t)m * nrep + t;
            D.e[e] = v[0];
            D.m[3 * e] = v[1] * inv; D.m[3 * e + 1] = v[2] * inv; D.m[3 * e + 2] = v[3] * inv;
        }
    }
}

// ---------------------------------------------------------------------------
// Topological charge. Every operation of a triangle is FP64 in a fixed order
// with contraction off, so it repeats bit for bit on any IEEE FP64 unit.
__device__ __forceinline__ void renorm3(const f3& v, double o[3])
{
#pragma clang fp contract(off)
    const double x = v.x, y = v.y, z = v.z;
    const double nn = sqrt(x * x + y * y + z * z);
    o[0] = x / nn; o[1] = y / nn; o[2] = z / nn;
}

// Solid angle of the triangle (a, b, c) of renormalised spins.
__device__ __forceinline__ double triangle_sigma(const double a[3], const double b[3], const double c[3])
{
#pragma clang fp contract(off)
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

// Spin of the plaquette corner at cell (ix, iy, iz): the site on SC; on
// multi-atom cells the cell's firstAtomInMeshCell, which is the LAST atom of
// the unit cell (makeChainMesh adds a cell's atoms in order and
// addAtomToChainCell puts each at the head of the list,
// src/chainMesh.f90:663-671,873-881): the body-centre atom on bcc, atom 12 on
// hyper-kagome. derivativeList steps to the same atom of the neighbouring
// cell (src/chainMesh.f90:931-1031).
__device__ __forceinline__ f3 spin_at(const float* __restrict__ spins, const Rep& R, uint32_t ix, uint32_t iy, uint32_t iz)
{
    if (R.cell == ISING_CELL_BCC)
        return ld3(planes(spins, R, 1), (iz * (uint32_t)R.ly + iy) * (uint32_t)R.hx + ix);
    if (R.cell == ISING_CELL_HYPERKAGOME)
        return ld3(planes(spins, R, 0), (uint32_t)c_hk.rank_of_type[kHkAtoms - 1] * ((uint32_t)R.plane * (uint32_t)R.lz)
                                            + (iz * (uint32_t)R.ly + iy) * (uint32_t)R.lx + ix);
    const uint32_t k = (ix + iy + iz) & 1u;
    const uint32_t idx = (iz * (uint32_t)R.ly + iy) * (uint32_t)R.hx + (ix >> 1);
    return ld3(planes(spins, R, (int)k), idx);
}

// layer < 0: every layer of every replica (grid = qblock table); layer >= 0:
// that layer only (grid = one-layer table, q_blocks_layer blocks per replica).
__global__ __launch_bounds__(kBlock) void k_winding(
    SpinBuf sb, uint32_t sweep_off, const Rep* __restrict__ reps, const int32_t* __restrict__ table, int layer,
    double* __restrict__ qpartial)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    const float* spins = spin_buffer(sb, *sb.ctr + sweep_off);
    const int slot = table[blockIdx.x];
    const Rep R = reps[slot];
    uint32_t z, chunk;
    if (layer < 0) {
        const uint32_t bi = (uint32_t)(blockIdx.x - R.q_first_block);
        z = bi / (uint32_t)R.q_blocks_layer;
        chunk = bi - z * (uint32_t)R.q_blocks_layer;
    } else {
        z = (uint32_t)layer;
        chunk = (uint32_t)(blockIdx.x - R.l_first_block);
    }
    const uint32_t pl = chunk * kBlock + threadIdx.x;
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    const uint32_t lx = (uint32_t)R.lx, ly = (uint32_t)R.ly;
    if (pl < lx * ly && z < (uint32_t)R.lz) {
        const uint32_t j = fdiv(pl, R.div_lx);
        const uint32_t i = pl - j * lx;
        const bool px = R.per & 1, py = (R.per >> 1) & 1;
        const bool okx = (i + 1u < lx) || px, oky = (j + 1u < ly) || py;
        if (okx && oky) {
            const uint32_t i2 = (i + 1u == lx) ? 0u : i + 1u;
            const uint32_t j2 = (j + 1u == ly) ? 0u : j + 1u;
            double s1[3], s2[3], s3[3], s4[3];
            renorm3(spin_at(spins, R, i, j, z), s1);
            renorm3(spin_at(spins, R, i2, j, z), s2);
            renorm3(spin_at(spins, R, i, j2, z), s3);
            renorm3(spin_at(spins, R, i2, j2, z), s4);
            v[0] = triangle_sigma(s1, s2, s4) + triangle_sigma(s1, s4, s3);
        }
    }
    block_sum4(v, lds);
    if (threadIdx.x == 0) qpartial[blockIdx.x] = v[0];
}

// Charge density of layer z of one slot: out[j*lx + i] = (sigma(s1,s2,s4) +
// sigma(s1,s4,s3)) / (4 pi) per plaquette, the terms k_winding sums (0 where
// a corner lies across an open face).
__global__ __launch_bounds__(kBlock) void k_winding_density(
    const float* __restrict__ spins, const Rep* __restrict__ reps, int slot, int z, double* __restrict__ out)
{
    const Rep R = reps[slot];
    const uint32_t lx = (uint32_t)R.lx, ly = (uint32_t)R.ly;
    const uint32_t pl = blockIdx.x * kBlock + threadIdx.x;
    if (pl >= lx * ly) return;
    const uint32_t j = fdiv(pl, R.div_lx);
    const uint32_t i = pl - j * lx;
    const bool px = R.per & 1, py = (R.per >> 1) & 1;
    double d = 0.0;
    if (((i + 1u < lx) || px) && ((j + 1u < ly) || py)) {
        const uint32_t i2 = (i + 1u == lx) ? 0u : i + 1u;
        const uint32_t j2 = (j + 1u == ly) ? 0u : j + 1u;
        double s1[3], s2[3], s3[3], s4[3];
        renorm3(spin_at(spins, R, i, j, (uint32_t)z), s1);
        renorm3(spin_at(spins, R, i2, j, (uint32_t)z), s2);
        renorm3(spin_at(spins, R, i, j2, (uint32_t)z), s3);
        renorm3(spin_at(spins, R, i2, j2, (uint32_t)z), s4);
        d = (triangle_sigma(s1, s2, s4) + triangle_sigma(s1, s4, s3)) / (4.0 * M_PI);
    }
    out[pl] = d;
}

// ---------------------------------------------------------------------------
// Skyrmion count of one layer (compute_skyrmion_distribution ->
// calculate_skyrmion_number -> calculate_winding_density_interpolated,
// src/reciprocalSpaceProcesses.f90:930-983, 627-812, 365-431), with the
// intended semantics of each step:
//   k_sk_coarse      corner spins of layer z (map_spins_to_std_grid), FP64;
//   k_sk_interp_x/_y band-limited interpolation onto the (sx nx) x (sy ny)
//                    grid, separable, kernel tables D_N(r/s + m) from the
//                    host, sums in ascending m without contraction;
//   k_sk_density     per fine plaquette (sigma(s1,s2,s3) + sigma(s1,s3,s4))
//                    / (4 pi), periodic, renormalised spins;
//   k_sk_filter      centred 5x5 Gaussian, periodic;
//   k_sk_absmax      max |f| (order-free: the bit patterns of non-negative
//                    doubles order like their values);
//   k_sk_root        each pixel's steepest-ascent root over its 8
//                    neighbours (|f|, ties to the earlier pixel of the
//                    reference's x-major scan);
//   k_sk_level       first threshold index that masks the pixel;
//   k_sk_accum       region charges in fixed point, threshold by threshold
//                    (64-bit integer atomics: exact and order-free);
//   k_sk_count       regions whose charge is within 0.05 of a particle
//                    number add its sign once.
constexpr double kSkFixed = 1099511627776.0;    // 2^40
constexpr double kSkFixedInv = 1.0 / 1099511627776.0;

__global__ __launch_bounds__(kBlock) void k_sk_coarse(const float* __restrict__ spins, const Rep* __restrict__ reps,
                                                      int slot, int z, double* __restrict__ S)
{
    const Rep R = reps[slot];
    const uint32_t nx = (uint32_t)R.lx, ny = (uint32_t)R.ly;
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= nx * ny) return;
    const uint32_t j = p / nx, i = p - j * nx;
    const f3 v = spin_at(spins, R, i, j, (uint32_t)z);
    S[p] = v.x; S[nx * ny + p] = v.y; S[2 * nx * ny + p] = v.z;
}

// T[c][j][I] = sum_m Kx[r][m] S[c][j][(n0 - m) mod nx], I = n0 sx + r.
__global__ __launch_bounds__(kBlock) void k_sk_interp_x(const double* __restrict__ S, int nx, int ny, int sx,
                                                        const double* __restrict__ Kx, double* __restrict__ T)
{
#pragma clang fp contract(off)
    const int NX = nx * sx;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (int64_t)3 * ny * NX) return;
    const int I = (int)(e % NX);
    const int64_t row = e / NX;                 // c*ny + j
    const int n0 = I / sx, r = I - n0 * sx;
    const double* k = Kx + (int64_t)r * nx;
    const double* in = S + row * nx;
    double acc = 0.0;
    int src = n0;                               // (n0 - m) mod nx
    for (int m = 0; m < nx; ++m) {
        acc += k[m] * in[src];
        src = src == 0 ? nx - 1 : src - 1;
    }
    T[e] = acc;
}

// F[c][J][I] = sum_m Ky[r][m] T[c][(m0 - m) mod ny][I], J = m0 sy + r.
__global__ __launch_bounds__(kBlock) void k_sk_interp_y(const double* __restrict__ T, int ny, int NX, int sy,
                                                        const double* __restrict__ Ky, double* __restrict__ F)
{
#pragma clang fp contract(off)
    const int NY = ny * sy;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (int64_t)3 * NY * NX) return;
    const int I = (int)(e % NX);
    const int64_t cj = e / NX;
    const int J = (int)(cj % NY), c = (int)(cj / NY);
    const int m0 = J / sy, r = J - m0 * sy;
    const double* k = Ky + (int64_t)r * ny;
    const double* in = T + (int64_t)c * ny * NX + I;
    double acc = 0.0;
    int src = m0;
    for (int m = 0; m < ny; ++m) {
        acc += k[m] * in[(int64_t)src * NX];
        src = src == 0 ? ny - 1 : src - 1;
    }
    F[e] = acc;
}

__device__ __forceinline__ bool sk_unit(const double* __restrict__ F, int64_t plane, int64_t idx, double o[3])
{
#pragma clang fp contract(off)
    const double x = F[idx], y = F[plane + idx], z = F[2 * plane + idx];
    const double nn = sqrt(x * x + y * y + z * z);
    if (!(nn > 0.0)) return false;
    o[0] = x / nn; o[1] = y / nn; o[2] = z / nn;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_sk_density(const double* __restrict__ F, int NX, int NY,
                                                       double* __restrict__ d)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t plane = (int64_t)NX * NY;
    if (p >= plane) return;
    const int J = (int)(p / NX), I = (int)(p - (int64_t)J * NX);
    const int I2 = I + 1 == NX ? 0 : I + 1, J2 = J + 1 == NY ? 0 : J + 1;
    double s1[3], s2[3], s3[3], s4[3];
    double v = 0.0;
    if (sk_unit(F, plane, p, s1) && sk_unit(F, plane, (int64_t)J * NX + I2, s2) &&
        sk_unit(F, plane, (int64_t)J2 * NX + I2, s3) && sk_unit(F, plane, (int64_t)J2 * NX + I, s4))
        v = (triangle_sigma(s1, s2, s3) + triangle_sigma(s1, s3, s4)) / (4.0 * M_PI);
    d[p] = v;
}

// f(I, J) = sum_{a,b = -2..2} d(I - a, J - b) G[a+2][b+2], a outer.
__global__ __launch_bounds__(kBlock) void k_sk_filter(const double* __restrict__ d, int NX, int NY,
                                                      const double* __restrict__ G, double* __restrict__ f)
{
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= (int64_t)NX * NY) return;
    const int J = (int)(p / NX), I = (int)(p - (int64_t)J * NX);
    double acc = 0.0;
    for (int a = -2; a <= 2; ++a) {
        int x = I - a;
        x = x < 0 ? x + NX : (x >= NX ? x - NX : x);
        for (int b = -2; b <= 2; ++b) {
            int y = J - b;
            y = y < 0 ? y + NY : (y >= NY ? y - NY : y);
            acc += d[(int64_t)y * NX + x] * G[(a + 2) * 5 + (b + 2)];
        }
    }
    f[p] = acc;
}

__global__ __launch_bounds__(kBlock) void k_sk_absmax(const double* __restrict__ f, int64_t n,
                                                      unsigned long long* __restrict__ mx)
{
    unsigned long long v = 0ull;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlock)
        v = max(v, (unsigned long long)__double_as_longlong(fabs(f[p])));
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, (unsigned long long)__shfl_xor((long long)v, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicMax(mx, v);
}

// Pixel key (|f|, earlier scan rank wins ties); scan rank = I*NY + J.
__device__ __forceinline__ bool sk_above(double v, int64_t r, double bv, int64_t br)
{
    return v > bv || (v == bv && r < br);
}

// Root of every pixel the lowest threshold masks (lvl <= T; the others get
// -1): repeated steps to the highest of the 8 neighbours while one is above
// the current pixel. The key order is strict, so every path ends, and it
// never leaves the mask (|f| does not decrease along it).
__global__ __launch_bounds__(kBlock) void k_sk_root(const double* __restrict__ f, int NX, int NY,
                                                    const uint8_t* __restrict__ lvl, int T, int32_t* __restrict__ root)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= (int64_t)NX * NY) return;
    if (lvl[p] > T) { root[p] = -1; return; }
    int J = (int)(p / NX), I = (int)(p - (int64_t)J * NX);
    double v = fabs(f[p]);
    for (;;) {
        double bv = v;
        int64_t br = (int64_t)I * NY + J;
        int bi = I, bj = J;
        for (int a = -1; a <= 1; ++a) {
            const int x = I + a < 0 ? NX - 1 : (I + a >= NX ? 0 : I + a);
            for (int b = -1; b <= 1; ++b) {
                if (a == 0 && b == 0) continue;
                const int y = J + b < 0 ? NY - 1 : (J + b >= NY ? 0 : J + b);
                const double w = fabs(f[(int64_t)y * NX + x]);
                const int64_t r = (int64_t)x * NY + y;
                if (sk_above(w, r, bv, br)) { bv = w; br = r; bi = x; bj = y; }
            }
        }
        if (bi == I && bj == J) break;
        I = bi; J = bj; v = bv;
    }
    root[p] = (int32_t)((int64_t)J * NX + I);
}

// First threshold index t (1-based) with |f| > q_t max|f|, q_t = max_t -
// (t/T)(max_t - min_t); T + 1 when none masks the pixel.
__global__ __launch_bounds__(kBlock) void k_sk_level(const double* __restrict__ f, int64_t n,
                                                     const unsigned long long* __restrict__ mxbits, int T,
                                                     double min_t, double max_t, uint8_t* __restrict__ lvl)
{
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const double mx = __longlong_as_double((long long)*mxbits);
    const double v = fabs(f[p]);
    int t = 1;
    for (; t <= T; ++t) {
        const double q = max_t - ((double)t / (double)T) * (max_t - min_t);
        if (v > q * mx) break;
    }
    lvl[p] = (uint8_t)t;
}

__global__ __launch_bounds__(kBlock) void k_sk_accum(const double* __restrict__ f, int64_t n,
                                                     const int32_t* __restrict__ root, const uint8_t* __restrict__ lvl,
                                                     int t, unsigned long long* __restrict__ acc)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n || lvl[p] != t) return;
    atomicAdd(acc + root[p], (unsigned long long)(long long)llrint(f[p] * kSkFixed));
}

__global__ __launch_bounds__(kBlock) void k_sk_count(int64_t n, const int32_t* __restrict__ root,
                                                     const uint8_t* __restrict__ lvl,
                                                     const unsigned long long* __restrict__ acc, int t, int n_part,
                                                     int32_t* __restrict__ counts)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n || root[p] != (int32_t)p || lvl[p] > t) return;
    const double ch = (double)(long long)acc[p] * kSkFixedInv;
    for (int j = 1; j <= n_part; ++j)
        if (fabs(fabs(ch) - (double)j) < 0.05) atomicAdd(counts + (j - 1), ch >= 0.0 ? 1 : -1);
}

// One workgroup per (slot, layer): fixed-order sum, Q = sum / (4 pi). In
// one-layer mode (layer >= 0) the other layers read 0.
// Block b finishes the charge of one (slot, layer): every layer of every
// slot (layer < 0, zmax blocks per slot) or the one layer (one block per
// slot).
__device__ __forceinline__ void winding_finalize_block(
    int b, double* lds, const Rep* __restrict__ reps, const double* __restrict__ qpartial, int zmax, int layer,
    double* __restrict__ qraw, long long* __restrict__ qint, const int32_t* __restrict__ slot_of_temp, int nrep,
    const SeriesRec& rec)
{
    const int zl = layer < 0 ? zmax : 1;
    const int slot = b / zl;
    const int z = layer < 0 ? b - slot * zl : layer;
    const Rep R = reps[slot];
    double q = 0.0;
    if (z < R.lz) {      // uniform
        double v[4] = { 0.0, 0.0, 0.0, 0.0 };
        const int base = layer < 0 ? R.q_first_block + z * R.q_blocks_layer : R.l_first_block;
        for (int j = threadIdx.x; j < R.q_blocks_layer; j += kBlock) v[0] += qpartial[base + j];
        block_sum4(v, lds);
        q = v[0] / (4.0 * 3.14159265358979323846);
    }
    if (threadIdx.x == 0) {
        qraw[slot * zmax + z] = q;
        qint[slot * zmax + z] = llrint(q);
    }
    if (rec.desc) {       // uniform
        const int t = temp_of_slot(slot_of_temp, nrep, slot);
        const uint32_t m = *rec.ctr + rec.local;
        const SeriesDesc& D = *rec.desc;
        if (threadIdx.x == 0 && m < D.cap_q && t >= 0) {
            const int64_t e = ((int64_t)m * nrep + t) * D.zrec + (D.zrec == 1u ? 0 : z);
            D.qr[e] = q;
            D.qi[e] = llrint(q);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_winding_finalize(
    const Rep* __restrict__ reps, const double* __restrict__ qpartial, int zmax, int layer,
    double* __restrict__ qraw, long long* __restrict__ qint, const int32_t* __restrict__ slot_of_temp, int nrep,
    SeriesRec rec)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    winding_finalize_block(blockIdx.x, lds, reps, qpartial, zmax, layer, qraw, qint, slot_of_temp, nrep, rec);
}

// Both finalizes of a sweep that records E/M and the charge in one launch:
// blocks [0, nrep*zmax) finish the charges, the next nrep blocks E and M.
__global__ __launch_bounds__(kBlock) void k_finalize_q_and_e(
    const Rep* __restrict__ reps, const double* __restrict__ qpartial, int zmax, int layer,
    double* __restrict__ qraw, long long* __restrict__ qint, SeriesRec rq,
    const double* __restrict__ wave_part, int wpb_f, double* __restrict__ energy, double* __restrict__ mag,
    SeriesRec rm, const int32_t* __restrict__ slot_of_temp, int nrep)
{
    __shared__ double lds[4 * (kBlock / kWave)];
    const int nq = nrep * (layer < 0 ? zmax : 1);
    if ((int)blockIdx.x < nq)
        winding_finalize_block(blockIdx.x, lds, reps, qpartial, zmax, layer, qraw, qint, slot_of_temp, nrep, rq);
    else
        wave_finalize_slot((int)blockIdx.x - nq, lds, reps, wave_part, wpb_f, energy, mag, slot_of_temp, nrep, rm);
}

// ---------------------------------------------------------------------------
// Replica exchange: one round over every ladder. Pairs (t, t+1) with
// t = round mod 2 inside each ladder; accept iff log u < (b_t - b_{t+1})(E_t - E_{t+1})
// in FP64. Labels move (which slot sits at which temperature), spins stay.
__global__ __launch_bounds__(kBlock) void k_swap(
    const double* __restrict__ energy, const double* __restrict__ beta_temp,
    int32_t* __restrict__ slot_of_temp, float* __restrict__ beta_slot,
    const int32_t* __restrict__ lad_of_pos, const int32_t* __restrict__ lad_start,
    const int32_t* __restrict__ lad_n, int R, uint32_t* __restrict__ round_ctr,
    uint32_t seed_lo, uint32_t seed_hi,
    unsigned long long* __restrict__ attempts, unsigned long long* __restrict__ accepts,
    const double* __restrict__ mag, SeriesRec rec)
{
    const uint32_t round = *round_ctr;
    __syncthreads();
    for (int t = threadIdx.x; t < R; t += blockDim.x) {
        const int l = lad_of_pos[t];
        const int tl = t - lad_start[l];
        if (((uint32_t)tl & 1u) != (round & 1u) || tl + 1 >= lad_n[l]) continue;
        const int a = slot_of_temp[t], b = slot_of_temp[t + 1];
        const double delta = (beta_temp[t] - beta_temp[t + 1]) * (energy[a] - energy[b]);
        const u32x4 r = philox4x32_10(u32x4{ (uint32_t)tl, round, kStreamSwap, seed_hi }, seed_lo, (uint32_t)l);
        const uint64_t m = ((uint64_t)(r.x >> 5) << 26) | (uint64_t)(r.y >> 6);
        const double u = ((double)m + 1.0) * 0x1p-53;
        attempts[t] += 1ull;
        if (log(u) < delta) {
            slot_of_temp[t] = b;
            slot_of_temp[t + 1] = a;
            beta_slot[b] = (float)beta_temp[t];
            beta_slot[a] = (float)beta_temp[t + 1];
            accepts[t] += 1ull;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *round_ctr = round + 1u;
    // E and M after the round, at each temperature (configurations stay put)
    const uint32_t m = rec.desc ? *rec.ctr + rec.local : 0u;
    if (rec.desc && m < rec.desc->cap_e) {
        const SeriesDesc& D = *rec.desc;
        for (int t = threadIdx.x; t < R; t += blockDim.x) {
            const int s = slot_of_temp[t];
            const int64_t e = (int64_t)m * R + t;
            D.e[e] = energy[s];
            D.m[3 * e] = mag[3 * s]; D.m[3 * e + 1] = mag[3 * s + 1]; D.m[3 * e + 2] = mag[3 * s + 2];
        }
    }
}

// Plain copy (16-byte vectors, grid-stride): the bandwidth and traffic probe.
__global__ __launch_bounds__(kBlock) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

// Advances the sweep counter by `sweeps`, the E/M and Q record counters by
// the records of the period and the fused-sweep counter by `fused`.
__global__ void k_advance(uint32_t* __restrict__ ctr, uint32_t sweeps, uint32_t meas, uint32_t q, uint32_t fused)
{
    if (threadIdx.x == 0) { ctr[0] += sweeps; ctr[1] += meas; ctr[2] += q; ctr[4] += fused; }
}
__global__ void k_set_u32(uint32_t* __restrict__ p, int n, uint32_t v)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}
// Writes the series descriptor of the next run (a kernel rather than a copy:
// it needs no host staging buffer and may be recorded in a graph).
__global__ void k_set_series_desc(SeriesDesc* __restrict__ d, SeriesDesc v)
{
    if (threadIdx.x == 0) *d = v;
}

// ---------------------------------------------------------------------------
// Initial states (the sites at half index q of every colour).
__global__ __launch_bounds__(kBlock) void k_init(
    float* __restrict__ spins, const Rep* __restrict__ reps, const int32_t* __restrict__ block_rep,
    int block_base, int mode, double frac, uint32_t seed_lo, uint32_t seed_hi)
{
    const int b = blockIdx.x + block_base;
    const int slot = block_rep[b];
    const Rep R = reps[slot];
    const uint32_t q = (uint32_t)(b - R.first_block) * kBlock + threadIdx.x;
    if (q >= R.half) return;
    const uint32_t row = fdiv(q, R.div_hx);
    const uint32_t h = q - row * (uint32_t)R.hx;
    const uint32_t iz = fdiv(row, R.div_ly);
    const uint32_t iy = row - iz * (uint32_t)R.ly;
    const bool bcc = R.cell == ISING_CELL_BCC, hk = R.cell == ISING_CELL_HYPERKAGOME;
    for (int colour = 0; colour < R.ncol; ++colour) {
        const Planes own = planes(spins, R, colour);
        uint32_t site;
        if (hk) {
            const HkSite t = hk_site(R, q, colour);
            site = (uint32_t)kHkAtoms * ((t.ix * (uint32_t)R.ly + t.iy) * (uint32_t)R.lz + t.iz)
                   + (uint32_t)c_hk.type_of_rank[t.rank];
        } else {
            const uint32_t p = (iy + iz + (uint32_t)colour) & 1u;
            const uint32_t ix = bcc ? h : 2u * h + p;
            const uint32_t cellsite = (ix * (uint32_t)R.ly + iy) * (uint32_t)R.lz + iz;
            site = bcc ? 2u * cellsite + (uint32_t)colour : cellsite;
        }
        const u32x4 r = philox4x32_10(u32x4{ site, 0xFFFFFFFFu, kStreamInit, seed_hi }, seed_lo, (uint32_t)slot);
        float x, y, z;
        if (mode == 0) {
            z = s24f(r.x);
            const float u = u24f(r.y);
            const float st = sqrtf(fmaxf(0.f, (1.f - z) * (1.f + z)));
            x = st * __builtin_amdgcn_cosf(u);
            y = st * __builtin_amdgcn_sinf(u);
            const float inv = 1.0f / sqrtf(x * x + y * y + z * z);
            x *= inv; y *= inv; z *= inv;
        } else {
            x = 0.f; y = 0.f;
            z = ((double)(r.z >> 8) * 0x1p-24 < frac) ? -1.f : 1.f;
        }
        stf((float*)own.x, q, x); stf((float*)own.y, q, y); stf((float*)own.z, q, z);
    }
}

// Single-site dE test hook: same local field and dE code as the sweep.
__global__ __launch_bounds__(kBlock) void k_delta_energy(
    const float* __restrict__ spins, const Rep* __restrict__ reps, const Lad* __restrict__ lads, int slot,
    const int32_t* __restrict__ sites, const float* __restrict__ props, int count, double* __restrict__ out,
    const double* __restrict__ demag)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const Rep R = reps[slot];
    const Lad L = lads[R.ladder];
    const f3 sn{ props[3 * e], props[3 * e + 1], props[3 * e + 2] };
    if (R.cell == ISING_CELL_HYPERKAGOME) {
        const uint32_t site = (uint32_t)sites[e];
        const uint32_t type = site % (uint32_t)kHkAtoms, cs = site / (uint32_t)kHkAtoms;
        const uint32_t iz = cs % (uint32_t)R.lz, iy = (cs / (uint32_t)R.lz) % (uint32_t)R.ly;
        const uint32_t ix = cs / ((uint32_t)R.lz * (uint32_t)R.ly);
        const uint32_t rank = (uint32_t)c_hk.rank_of_type[type];
        const uint32_t cell = (iz * (uint32_t)R.ly + iy) * (uint32_t)R.lx + ix;
        const uint32_t t = rank * ((uint32_t)R.plane * (uint32_t)R.lz) + cell;   // colour*half + q
        const Planes all = planes(spins, R, 0);
        const f3 s = ld3(all, t);
        f3 hf = hk_field(L, R, all, HkSite{ ix, iy, iz, rank, cell });
        if (L.demag) add_demag(L, R, demag, t, hf);
        out[e] = (double)delta_e(s, sn, hf, L.Kz);
        return;
    }
    const bool bcc = R.cell == ISING_CELL_BCC;
    const uint32_t site = (uint32_t)sites[e];
    const uint32_t cellsite = bcc ? site >> 1 : site;
    const uint32_t iz = cellsite % (uint32_t)R.lz;
    const uint32_t iy = (cellsite / (uint32_t)R.lz) % (uint32_t)R.ly;
    const uint32_t ix = cellsite / ((uint32_t)R.lz * (uint32_t)R.ly);
    const uint32_t colour = bcc ? site & 1u : (ix + iy + iz) & 1u;
    const uint32_t row = iz * (uint32_t)R.ly + iy;
    const uint32_t h = bcc ? ix : ix >> 1;
    const uint32_t q = row * (uint32_t)R.hx + h;
    const Planes own = planes(spins, R, (int)colour);
    const Planes oth = planes(spins, R, (int)(colour ^ 1u));
    const f3 s = ld3(own, q);
    f3 hf = bcc ? bcc_field(L, bcc_sums(R, oth, q, ix, iy, iz, (int)colour))
                : local_field(R, L, oth, neighbours<false>(R, q, h, iy, iz, ix & 1u));
    if (L.demag) add_demag(L, R, demag, colour * R.half + q, hf);
    out[e] = (double)delta_e(s, sn, hf, L.Kz);
}

// ---------------------------------------------------------------------------
// Demagnetising field, calculate_demagnetisation_field
// (src/reciprocalSpaceProcesses.f90:202-281), FP64 throughout:
//   k_demag_deposit    cloud-in-cell deposit onto one mesh point per cell
//                      centre (interpolate_to_fft_array, :76-129);
//   k_dft_lines<kR2cX> forward real-to-half-complex transform along x;
//   k_dft_lines<kC2cY>, <kC2cZ>  complex transforms along y and z (both
//                      directions);
//   k_demag_project    H(k) = -(k.M) k / k^2, k = 0 removed (:237-270);
//   k_dft_lines<kC2rX> half-complex-to-real inverse along x, reading only
//                      the real parts of the DC and Nyquist terms (FFTW's
//                      c2r convention), scaled by 1/(lx ly lz) (:271-275);
//   k_demag_gather     the same weights back to the atoms
//                      (interpolate_fft_to_atoms, :131-178).
// Mesh arrays per replica: real [3][lz][ly][lx] (cell (k*ly + j)*lx + i, x
// fastest as in the spin layout, so the deposit and the gather read
// neighbouring cells and atoms with neighbouring lanes), complex
// [3][lz][ly][lx/2+1]; the half axis is x, as in the reference. The transforms
// (k_dft_lines) take every line of one axis through LDS: one Cooley-Tukey
// step L = P Q per line, with exact FP64 twiddles e^{-2 pi i m/L} from a
// table per length; any length up to kDftMaxLen. Every kernel runs on grid
// (x blocks, replica slot); a slot is skipped unless it is fully periodic
// and selected (demag_slot).
struct DemagRep {
    int64_t moff;        // doubles: real mesh of this slot
    int64_t koff;        // complex elements: half-complex mesh of this slot
    int32_t tw[3];       // first entry of the twiddle table per axis length
    int32_t hx;          // lx/2 + 1
    int32_t fp[3];       // P of L = P Q per axis
    int32_t tl[3];       // lines per workgroup per axis
    int32_t tz;          // z lines per workgroup of k_demag_zproject
};

constexpr int kDftMaxLen = 1024;     // longest axis of a lattice with the field
constexpr int kDftMaxLines = 64;     // lines per workgroup
constexpr int kDftLdsTarget = 2560;  // double2 of LDS per workgroup, lengths up to 512

// only >= 0: that slot; kDemagLadders: slots whose ladder has the term on;
// kDemagAll: every fully periodic slot.
constexpr int kDemagLadders = -1;
constexpr int kDemagAll = -2;

__device__ __forceinline__ bool demag_slot(const Rep& R, const Lad* __restrict__ lads, int only)
{
    if (R.per != 7 || R.lx > kDftMaxLen || R.ly > kDftMaxLen || R.lz > kDftMaxLen) return false;
    if (only >= 0) return (int)blockIdx.y == only;
    return only == kDemagAll || lads[R.ladder].demag;
}

// One step of getNeighboringCells (src/chainMesh.f90:98-117): n -> 0, -1 -> n-1.
__host__ __device__ inline int wrap_cell(int i, int n) { return i == n ? 0 : (i == -1 ? n - 1 : i); }

// Weight factor of one axis (host): d = atom - centre in lattice units,
// minimum image d - n nint(d/n) (nint half away from zero;
// distance_points_vec_bravais, src/chainMesh.f90:268-270), then
// max(1 - |d|, 0).
inline double cic_axis(int atom_cell, double frac, int centre_cell, int n)
{
    double d = ((double)atom_cell + frac) - ((double)centre_cell + 0.5);
    d -= (double)n * std::round(d / (double)n);
    return std::fmax(1.0 - std::fabs(d), 0.0);
}

// Colour-major index colour*half + q of atom a of cell (i, j, k).
__device__ __forceinline__ uint32_t atom_index(const Rep& R, int i, int j, int k, int a)
{
    const uint32_t row = (uint32_t)k * (uint32_t)R.ly + (uint32_t)j;
    if (R.cell == ISING_CELL_HYPERKAGOME)
        return (uint32_t)c_hk.rank_of_type[a] * ((uint32_t)R.plane * (uint32_t)R.lz) + row * (uint32_t)R.lx + (uint32_t)i;
    if (R.cell == ISING_CELL_BCC) return (uint32_t)a * R.half + row * (uint32_t)R.hx + (uint32_t)i;
    return (uint32_t)((i + j + k) & 1) * R.half + row * (uint32_t)R.hx + ((uint32_t)i >> 1);
}

// Atoms per cell.
__device__ __forceinline__ int atoms_per_cell(const Rep& R)
{
    return R.cell == ISING_CELL_HYPERKAGOME ? kHkAtoms : (R.cell == ISING_CELL_BCC ? 2 : 1);
}

// Weight factors per slot, atom type and axis: cic[((slot * kHkAtoms + a) * 3
// + axis) * 3 + di + 1] = cic_axis(c + di, frac, c, n) for an atom of type a
// in cell c + di (wrapped) and the mesh point of cell c. The minimum image
// takes the wrap back out, so the factor is the same for every cell c; the
// host fills the table with the same formula (every operand a multiple of
// 1/4: exact). Atom positions: the corner on SC, corner and body centre on
// bcc, the table's fractions on hyper-kagome (down to -3/4: such an atom
// still deposits only on the 27 cells around its own, as the reference's
// loop does). The gather's factor for the mesh point of cell c + di seen
// from an atom in cell c is the entry of -di.
constexpr int kCicPerSlot = kHkAtoms * 9;

__global__ __launch_bounds__(kBlock) void k_demag_deposit(
    const float* __restrict__ spins, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, const double* __restrict__ cic, int only, double* __restrict__ mesh)
{
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const int nx = R.lx, ny = R.ly, nz = R.lz;
    const int64_t Nc = (int64_t)nx * ny * nz;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= Nc) return;
    const uint32_t row = (uint32_t)cell / (uint32_t)nx;     // cells < 2^31
    const int i = (int)((uint32_t)cell - row * (uint32_t)nx), j = (int)(row % (uint32_t)ny), k = (int)(row / (uint32_t)ny);
    const int na = atoms_per_cell(R);
    const float* sp = spins + R.off;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll 1
    for (int a = 0; a < na; ++a) {
        // the weight is a product of one factor per axis (uniform loads)
        const double* w = cic + ((size_t)blockIdx.y * kHkAtoms + a) * 9;
        double fx[3], fy[3], fz[3];
        for (int d = 0; d < 3; ++d) {
            fx[d] = w[d];
            fy[d] = w[3 + d];
            fz[d] = w[6 + d];
        }
#pragma unroll 1
        for (int di = 0; di < 3; ++di) {
            if (fx[di] == 0.0) continue;
            const int ni = wrap_cell(i + di - 1, nx);
            for (int dj = 0; dj < 3; ++dj) {
                if (fy[dj] == 0.0) continue;
                const int nj = wrap_cell(j + dj - 1, ny);
                for (int dk = 0; dk < 3; ++dk) {
                    const double w = fx[di] * fy[dj] * fz[dk];
                    if (w == 0.0) continue;
                    const uint32_t t = atom_index(R, ni, nj, wrap_cell(k + dk - 1, nz), a);
                    m0 += w * (double)sp[t];
                    m1 += w * (double)sp[R.n + t];
                    m2 += w * (double)sp[2 * R.n + t];
                }
            }
        }
    }
    double* o = mesh + dreps[blockIdx.y].moff;
    o[cell] = m0;
    o[Nc + cell] = m1;
    o[2 * Nc + cell] = m2;
}

// Line transforms of the field mesh, one kernel per direction:
//   kR2cX  X[c][k][j][kx] = sum_n x[c][k][j][n] e^{-2 pi i kx n/lx}, kx <= lx/2;
//   kC2cY, kC2cZ  complex transforms along y and z of [3][lz][ly][hx], sign -1
//          forward (e^{-2 pi i k n/len}), +1 inverse (unnormalised);
//   kC2rX  x[c][k][j][i] = Re sum_kx Y_kx e^{2 pi i kx i/lx} / (lx ly lz), Y the
//          Hermitian completion of the lx/2+1 stored values with the DC and
//          Nyquist terms taken real (FFTW's c2r convention).
// A workgroup holds T lines of one slot in LDS ([n][T+1]) with the slot's
// twiddle table W^j = e^{-2 pi i j/L}, and transforms them by one
// Cooley-Tukey step L = P Q (n = Q n1 + n2, k = k1 + P k2):
//   Y[n2][k1] = W^{n2 k1} sum_n1 x[Q n1 + n2] W^{Q (n1 k1 mod P)},
//   X[k1 + P k2] = sum_n2 Y[n2][k1] W^{P (n2 k2 mod Q)},
// P + Q + 1 complex products per output (a direct sum takes L). The inverse
// reads the table conjugated. Lines: x (c, row), contiguous; y (c, kz, kx),
// element stride hx; z (c, ky, kx), stride ly hx. Loads and stores run with
// consecutive lanes on consecutive addresses (elements for x, lines for y
// and z).
//
// Two kernels share this: k_dft_lines_r (P, Q <= kDftRadix) gives each thread
// one (line, n2) of the first step and one (line, k1) of the second, with its
// P or Q inputs in registers: one LDS read per input, twiddles read as
// wavefront-uniform broadcasts, the result written back in place. k_dft_lines
// (any P, Q) loops over outputs with a second LDS buffer.
enum DftMode { kR2cX = 0, kC2cY = 1, kC2cZ = 2, kC2rX = 3 };
constexpr int kDftRadix = 16;

template <int kMode> using DftIn = std::conditional_t<kMode == kR2cX, double, double2>;
template <int kMode> constexpr int kDftAxis = kMode == kC2cY ? 1 : (kMode == kC2cZ ? 2 : 0);
template <int kMode> using DftOut = std::conditional_t<kMode == kC2rX, double, double2>;

// Geometry of a workgroup's lines: length, factors L = P Q, lines per
// workgroup T, lines present nt, element stride.
struct DftLines {
    int L, P, Q, T, nt;
    int64_t stride;
};

// Lines of this workgroup: sets g, fills the twiddles w[L] and the first
// element of every line (s_in in the source mesh, s_out in the destination),
// then loads the lines into a[n * (T + 1) + t]. False (for the whole
// workgroup) when it has no line.
template <int kMode>
__device__ __forceinline__ bool dft_load(const Rep& R, const DemagRep& D, const DftIn<kMode>* __restrict__ in,
                                         const double2* __restrict__ tw, double2* a, double2* w, int64_t* s_in,
                                         int64_t* s_out, DftLines& g)
{
    constexpr int kAxis = kDftAxis<kMode>;
    constexpr bool kContig = kAxis == 0;
    g.L = kAxis == 0 ? R.lx : (kAxis == 1 ? R.ly : R.lz);
    g.P = D.fp[kAxis];
    g.Q = g.L / g.P;
    g.T = D.tl[kAxis];
    const int L = g.L, T = g.T, TP = T + 1;
    const int64_t rows = (int64_t)R.ly * R.lz;                 // x lines per component
    const int64_t krow = (int64_t)R.ly * D.hx;                 // modes per kz plane
    const int64_t lines = kAxis == 0 ? 3 * rows : 3 * (int64_t)D.hx * (kAxis == 1 ? R.lz : R.ly);
    const int64_t first = (int64_t)blockIdx.x * T;
    if (first >= lines) return false;
    g.nt = (int)min((int64_t)T, lines - first);
    g.stride = kAxis == 0 ? 1 : (kAxis == 1 ? (int64_t)D.hx : krow);
    for (int i = threadIdx.x; i < L; i += kBlock) w[i] = tw[D.tw[kAxis] + i];
    if ((int)threadIdx.x < g.nt) {
        const int64_t l = first + threadIdx.x;
        if (kAxis == 0) {         // l = c rows + row
            const int64_t kbase = D.koff + l * D.hx, mbase = D.moff + l * R.lx;
            s_in[threadIdx.x] = kMode == kR2cX ? mbase : kbase;
            s_out[threadIdx.x] = kMode == kR2cX ? kbase : mbase;
        } else {                  // y: l = (c lz + kz) hx + kx; z: l = c krow + (ky hx + kx)
            const int64_t inner = kAxis == 1 ? D.hx : krow;
            const int64_t outer = l / inner, r = l - outer * inner;
            const int64_t kbase = D.koff + outer * (kAxis == 1 ? krow : krow * R.lz) + r;
            s_in[threadIdx.x] = kbase;
            s_out[threadIdx.x] = kbase;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L * T; i += kBlock) {
        int t, n;
        if (kContig) { t = i / L; n = i - t * L; }
        else { n = i / T; t = i - n * T; }
        double2 v = double2{ 0.0, 0.0 };
        if (t < g.nt) {
            const int64_t base = s_in[t];
            if constexpr (kMode == kR2cX) {
                v = double2{ in[base + n * g.stride], 0.0 };
            } else if constexpr (kMode == kC2rX) {
                if (n < D.hx) {
                    v = in[base + n * g.stride];
                    if (n == 0 || 2 * n == L) v.y = 0.0;
                } else {
                    const double2 u = in[base + (int64_t)(L - n) * g.stride];
                    v = double2{ u.x, -u.y };
                }
            } else {
                v = in[base + n * g.stride];
            }
        }
        a[n * TP + t] = v;
    }
    __syncthreads();
    return true;
}

// v * conj^s(u): the table holds e^{-...}; sg = -1 reads it conjugated.
__device__ __forceinline__ void cmac(double& re, double& im, const double2& v, const double2& u, double sg)
{
    const double ui = sg * u.y;
    re += v.x * u.x - v.y * ui;
    im += v.x * ui + v.y * u.x;
}

template <int kMode>
__device__ __forceinline__ void dft_store(DftOut<kMode>* __restrict__ out, int64_t at, double re, double im,
                                          double inv_cells)
{
    if constexpr (kMode == kC2rX) out[at] = re * inv_cells;
    else out[at] = double2{ re, im };
}

template <int kMode>
__device__ __forceinline__ double dft_sign(int sign) { return (kMode == kC2rX || sign > 0) ? -1.0 : 1.0; }

template <int kMode>
__global__ __launch_bounds__(kBlock) void k_dft_lines(
    const DftIn<kMode>* __restrict__ in, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, int only, const double2* __restrict__ tw, int sign,
    DftOut<kMode>* __restrict__ out)
{
    constexpr bool kContig = kMode == kR2cX || kMode == kC2rX;
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const DemagRep D = dreps[blockIdx.y];
    extern __shared__ double2 dft_lds[];
    __shared__ int64_t s_in[kDftMaxLines], s_out[kDftMaxLines];
    const int len = kDftAxis<kMode> == 0 ? R.lx : (kDftAxis<kMode> == 1 ? R.ly : R.lz);
    double2* a = dft_lds;                                       // [L][TP] input lines
    double2* b = a + len * (D.tl[kDftAxis<kMode>] + 1);          // [L][TP] after the first step
    double2* w = b + len * (D.tl[kDftAxis<kMode>] + 1);          // [L] twiddles
    DftLines g;
    if (!dft_load<kMode>(R, D, in, tw, a, w, s_in, s_out, g)) return;
    const int L = g.L, P = g.P, Q = g.Q, T = g.T, TP = T + 1;
    const double sg = dft_sign<kMode>(sign);
    for (int i = threadIdx.x; i < L * T; i += kBlock) {
        const int r = i / T, t = i - r * T;
        const int n2 = r / P, k1 = r - n2 * P;
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int n1 = 0; n1 < P; ++n1) {
            cmac(re, im, a[(Q * n1 + n2) * TP + t], w[m * Q], sg);
            m += k1;
            if (m >= P) m -= P;
        }
        const double2 u = w[n2 * k1];
        const double ui = sg * u.y;
        b[(n2 * P + k1) * TP + t] = double2{ re * u.x - im * ui, re * ui + im * u.x };
    }
    __syncthreads();
    const int kout = kMode == kR2cX ? D.hx : L;
    const double inv_cells = 1.0 / ((double)R.lx * R.ly * R.lz);
    for (int i = threadIdx.x; i < kout * T; i += kBlock) {
        int t, k;
        if (kContig) { t = i / kout; k = i - t * kout; }
        else { k = i / T; t = i - k * T; }
        if (t >= g.nt) continue;
        const int k2 = k / P, k1 = k - k2 * P;
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int n2 = 0; n2 < Q; ++n2) {
            cmac(re, im, b[(n2 * P + k1) * TP + t], w[m * P], sg);
            m += k2;
            if (m >= Q) m -= Q;
        }
        dft_store<kMode>(out, s_out[t] + k * g.stride, re, im, inv_cells);
    }
}

// Bit reversal of n < N = 2^m.
template <int N>
__host__ __device__ constexpr int bitrev(int n)
{
    int r = 0;
    for (int b = 1; b < N; b <<= 1) {
        r = (r << 1) | (n & 1);
        n >>= 1;
    }
    return r;
}

// In-register DFT of N = 2^m points, natural order in and out:
// x[k] <- sum_n x[n] W^{(n k mod N) s}, W^j = w[j] (conjugated for sg = -1),
// by radix-2 decimation in time.
template <int N>
__device__ __forceinline__ void fft_pow2(double2 (&x)[kDftRadix], const double2* w, int s, double sg)
{
    double2 y[N];
#pragma unroll
    for (int n = 0; n < N; ++n) y[bitrev<N>(n)] = x[n];
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int j = 0; j < len / 2; ++j) {
            double2 u = double2{ 1.0, 0.0 };
            if (j > 0) {
                u = w[j * (N / len) * s];
                u.y *= sg;
            }
#pragma unroll
            for (int b = 0; b < N; b += len) {
                const double2 p = y[b + j], q = y[b + j + len / 2];
                const double2 t = j > 0 ? double2{ q.x * u.x - q.y * u.y, q.x * u.y + q.y * u.x } : q;
                y[b + j] = double2{ p.x + t.x, p.y + t.y };
                y[b + j + len / 2] = double2{ p.x - t.x, p.y - t.y };
            }
        }
    }
#pragma unroll
    for (int n = 0; n < N; ++n) x[n] = y[n];
}

// x[0..n) <- its n-point DFT (twiddle stride s) when n is 1, 2, 4 or 8; false
// for other n (n is uniform per workgroup). Sixteen points in registers would
// double the registers of the kernel.
__device__ __forceinline__ bool fft_reg(double2 (&x)[kDftRadix], int n, const double2* w, int s, double sg)
{
    switch (n) {
    case 1: return true;
    case 2: fft_pow2<2>(x, w, s, sg); return true;
    case 4: fft_pow2<4>(x, w, s, sg); return true;
    case 8: fft_pow2<8>(x, w, s, sg); return true;
    default: return false;
    }
}

// P, Q <= kDftRadix, Q T <= kBlock and P T <= kBlock: one item per thread
// in each step, in place in one LDS buffer; power-of-two factors run as
// in-register radix-2 FFTs, other factors as direct sums.
template <int kMode>
__global__ __launch_bounds__(kBlock) void k_dft_lines_r(
    const DftIn<kMode>* __restrict__ in, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, int only, const double2* __restrict__ tw, int sign,
    DftOut<kMode>* __restrict__ out)
{
    constexpr bool kContig = kMode == kR2cX || kMode == kC2rX;
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const DemagRep D = dreps[blockIdx.y];
    extern __shared__ double2 dft_lds[];
    __shared__ int64_t s_in[kDftMaxLines], s_out[kDftMaxLines];
    const int len = kDftAxis<kMode> == 0 ? R.lx : (kDftAxis<kMode> == 1 ? R.ly : R.lz);
    double2* a = dft_lds;                                       // [L][TP] lines
    double2* w = a + len * (D.tl[kDftAxis<kMode>] + 1);          // [L] twiddles
    DftLines g;
    if (!dft_load<kMode>(R, D, in, tw, a, w, s_in, s_out, g)) return;
    const int P = g.P, Q = g.Q, T = g.T, TP = T + 1;
    const double sg = dft_sign<kMode>(sign);

    // first step: thread (n2, t) takes x[Q n1 + n2], n1 < P
    const int t1 = threadIdx.x % T, n2 = threadIdx.x / T;
    const bool on1 = n2 < Q;
    double2 x[kDftRadix];
#pragma unroll
    for (int n1 = 0; n1 < kDftRadix; ++n1)
        if (on1 && n1 < P) x[n1] = a[(Q * n1 + n2) * TP + t1];
    __syncthreads();
    if (on1 && fft_reg(x, P, w, Q, sg)) {
#pragma unroll
        for (int k1 = 0; k1 < kDftRadix; ++k1) {
            if (k1 < P) {
                const double2 u = w[n2 * k1];
                const double ui = sg * u.y;
                a[(n2 * P + k1) * TP + t1] = double2{ x[k1].x * u.x - x[k1].y * ui, x[k1].x * ui + x[k1].y * u.x };
            }
        }
    } else if (on1) {
        for (int k1 = 0; k1 < P; ++k1) {
            double re = 0.0, im = 0.0;
            int m = 0;
#pragma unroll
            for (int n1 = 0; n1 < kDftRadix; ++n1) {
                if (n1 < P) {
                    cmac(re, im, x[n1], w[m * Q], sg);
                    m += k1;
                    if (m >= P) m -= P;
                }
            }
            const double2 u = w[n2 * k1];
            const double ui = sg * u.y;
            a[(n2 * P + k1) * TP + t1] = double2{ re * u.x - im * ui, re * ui + im * u.x };
        }
    }
    __syncthreads();

    // second step: thread (k1, t) takes Y[n2][k1], n2 < Q, and writes
    // X[k1 + P k2]; k1 fastest across lanes on contiguous (x) lines
    int t2, k1;
    if (kContig) { k1 = threadIdx.x % P; t2 = threadIdx.x / P; }
    else { t2 = threadIdx.x % T; k1 = threadIdx.x / T; }
    if (k1 >= P || t2 >= g.nt) return;
#pragma unroll
    for (int j = 0; j < kDftRadix; ++j)
        if (j < Q) x[j] = a[(j * P + k1) * TP + t2];
    const int kout = kMode == kR2cX ? D.hx : g.L;
    const double inv_cells = 1.0 / ((double)R.lx * R.ly * R.lz);
    const int64_t base = s_out[t2];
    if (fft_reg(x, Q, w, P, sg)) {
#pragma unroll
        for (int k2 = 0; k2 < kDftRadix; ++k2) {
            const int k = k1 + P * k2;
            if (k2 < Q && k < kout) dft_store<kMode>(out, base + k * g.stride, x[k2].x, x[k2].y, inv_cells);
        }
        return;
    }
    for (int k2 = 0; k2 < Q; ++k2) {
        const int k = k1 + P * k2;
        if (k >= kout) break;
        double re = 0.0, im = 0.0;
        int m = 0;
#pragma unroll
        for (int j = 0; j < kDftRadix; ++j) {
            if (j < Q) {
                cmac(re, im, x[j], w[m * P], sg);
                m += k2;
                if (m >= Q) m -= Q;
            }
        }
        dft_store<kMode>(out, base + k * g.stride, re, im, inv_cells);
    }
}

// H(k) = -(k.M(k)) k / k^2 on the half-complex mesh, in place; wave indices
// wx = kx, wy and wz folded to (-n/2, n/2]; k = 2 pi w / n per axis (cubic
// cells in lattice units, so the reference's cutoff |k_a| > pi/a never
// removes a mode).
__device__ __forceinline__ void project_mode(int kx, int ky, int kz, int nx, int ny, int nz, double2& a, double2& b,
                                             double2& c)
{
    if (kx == 0 && ky == 0 && kz == 0) {
        a = b = c = double2{ 0.0, 0.0 };
        return;
    }
    const int wy = ky > ny / 2 ? ky - ny : ky, wz = kz > nz / 2 ? kz - nz : kz;
    const double two_pi = 6.283185307179586476925;
    const double k0 = two_pi * ((double)kx / (double)nx);
    const double k1 = two_pi * ((double)wy / (double)ny);
    const double k2 = two_pi * ((double)wz / (double)nz);
    const double inv = 1.0 / (k0 * k0 + k1 * k1 + k2 * k2);
    const double dr = (k0 * a.x + k1 * b.x + k2 * c.x) * inv;
    const double di = (k0 * a.y + k1 * b.y + k2 * c.y) * inv;
    a = double2{ -dr * k0, -di * k0 };
    b = double2{ -dr * k1, -di * k1 };
    c = double2{ -dr * k2, -di * k2 };
}

__global__ __launch_bounds__(kBlock) void k_demag_project(
    double2* __restrict__ F, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, int only)
{
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const DemagRep D = dreps[blockIdx.y];
    const int nx = R.lx, ny = R.ly, nz = R.lz;
    const int64_t Nk = (int64_t)D.hx * ny * nz;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= Nk) return;
    const int kx = (int)(e % D.hx), ky = (int)((e / D.hx) % ny), kz = (int)(e / ((int64_t)D.hx * ny));
    double2* f = F + D.koff + e;
    project_mode(kx, ky, kz, nx, ny, nz, f[0], f[Nk], f[2 * Nk]);
}

// One transform of nc components x T lines in LDS, in place and in natural
// order: a[(c L + n) TP + t] -> X[k] at (c L + k) TP + t, by the two steps of
// k_dft_lines_r (P, Q <= kDftRadix, nc Q T <= kBlock). Every thread of the
// workgroup calls it (it holds barriers).
__device__ __forceinline__ void dft_lines_lds(double2* a, const double2* w, int L, int P, int Q, int T, int nc,
                                              double sg)
{
    const int TP = T + 1;
    double2 x[kDftRadix];
    // first step: item (c, n2, t) takes x[Q n1 + n2], n1 < P
    {
        const int t = threadIdx.x % T, r = threadIdx.x / T;
        const int c = r / Q, n2 = r - c * Q;
        const bool on = c < nc;
        double2* line = a + (size_t)c * L * TP + t;
#pragma unroll
        for (int n1 = 0; n1 < kDftRadix; ++n1)
            if (on && n1 < P) x[n1] = line[(Q * n1 + n2) * TP];
        __syncthreads();
        if (on && fft_reg(x, P, w, Q, sg)) {
#pragma unroll
            for (int k1 = 0; k1 < kDftRadix; ++k1) {
                if (k1 < P) {
                    const double2 u = w[n2 * k1];
                    const double ui = sg * u.y;
                    line[(n2 * P + k1) * TP] = double2{ x[k1].x * u.x - x[k1].y * ui, x[k1].x * ui + x[k1].y * u.x };
                }
            }
        } else if (on) {
            for (int k1 = 0; k1 < P; ++k1) {
                double re = 0.0, im = 0.0;
                int m = 0;
#pragma unroll
                for (int n1 = 0; n1 < kDftRadix; ++n1) {
                    if (n1 < P) {
                        cmac(re, im, x[n1], w[m * Q], sg);
                        m += k1;
                        if (m >= P) m -= P;
                    }
                }
                const double2 u = w[n2 * k1];
                const double ui = sg * u.y;
                line[(n2 * P + k1) * TP] = double2{ re * u.x - im * ui, re * ui + im * u.x };
            }
        }
        __syncthreads();
    }
    // second step: item (c, k1, t) takes Y[n2][k1], n2 < Q, writes X[k1 + P k2]
    {
        const int t = threadIdx.x % T, r = threadIdx.x / T;
        const int c = r / P, k1 = r - c * P;
        const bool on = c < nc;
        double2* line = a + (size_t)c * L * TP + t;
#pragma unroll
        for (int j = 0; j < kDftRadix; ++j)
            if (on && j < Q) x[j] = line[(j * P + k1) * TP];
        __syncthreads();
        if (on && fft_reg(x, Q, w, P, sg)) {
#pragma unroll
            for (int k2 = 0; k2 < kDftRadix; ++k2)
                if (k2 < Q) line[(k1 + P * k2) * TP] = x[k2];
        } else if (on) {
            for (int k2 = 0; k2 < Q; ++k2) {
                double re = 0.0, im = 0.0;
                int m = 0;
#pragma unroll
                for (int j = 0; j < kDftRadix; ++j) {
                    if (j < Q) {
                        cmac(re, im, x[j], w[m * P], sg);
                        m += k2;
                        if (m >= Q) m -= Q;
                    }
                }
                line[(k1 + P * k2) * TP] = double2{ re, im };
            }
        }
        __syncthreads();
    }
}

// The forward z transform, the projection and the inverse z transform in one
// pass: T z lines (ky, kx) of all three components in LDS (D.tz lines per
// workgroup, 3 Q T <= kBlock), the complex mesh overwritten in place with H.
__global__ __launch_bounds__(kBlock) void k_demag_zproject(
    double2* __restrict__ F, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, int only, const double2* __restrict__ tw)
{
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const DemagRep D = dreps[blockIdx.y];
    const int L = R.lz, P = D.fp[2], Q = L / P, T = D.tz, TP = T + 1;
    const int64_t krow = (int64_t)R.ly * D.hx, Nk = krow * R.lz;
    const int64_t first = (int64_t)blockIdx.x * T;
    if (first >= krow) return;    // uniform per workgroup
    const int nt = (int)min((int64_t)T, krow - first);
    extern __shared__ double2 dft_lds[];
    double2* a = dft_lds;                     // [3][L][TP]
    double2* w = a + 3 * L * TP;              // [L]
    for (int i = threadIdx.x; i < L; i += kBlock) w[i] = tw[D.tw[2] + i];
    double2* f = F + D.koff + first;
    for (int i = threadIdx.x; i < 3 * L * T; i += kBlock) {
        const int t = i % T, r = i / T;       // r = c L + n
        const int c = r / L, n = r - c * L;
        a[r * TP + t] = t < nt ? f[c * Nk + n * krow + t] : double2{ 0.0, 0.0 };
    }
    __syncthreads();
    dft_lines_lds(a, w, L, P, Q, T, 3, 1.0);
    for (int i = threadIdx.x; i < L * T; i += kBlock) {
        const int t = i % T, kz = i / T;
        if (t >= nt) continue;
        const int64_t r = first + t;
        const int ky = (int)(r / D.hx), kx = (int)(r - (int64_t)ky * D.hx);
        project_mode(kx, ky, kz, R.lx, R.ly, R.lz, a[kz * TP + t], a[(L + kz) * TP + t], a[(2 * L + kz) * TP + t]);
    }
    __syncthreads();
    dft_lines_lds(a, w, L, P, Q, T, 3, -1.0);
    for (int i = threadIdx.x; i < 3 * L * T; i += kBlock) {
        const int t = i % T, r = i / T;
        const int c = r / L, n = r - c * L;
        if (t < nt) f[c * Nk + n * krow + t] = a[r * TP + t];
    }
}

// Sum of the mesh field over the 27 cells around the atom's cell (ix, iy,
// iz) with the factors cw of its type.
__device__ __forceinline__ void gather_sum(const Rep& R, const double* __restrict__ m, const double* __restrict__ cw,
                                           int ix, int iy, int iz, double (&h)[3])
{
    const int nx = R.lx, ny = R.ly, nz = R.lz;
    const int64_t Nc = (int64_t)nx * ny * nz;
    h[0] = h[1] = h[2] = 0.0;
    for (int di = -1; di <= 1; ++di) {
        const double wx = cw[1 - di];
        if (wx == 0.0) continue;
        const int ci = wrap_cell(ix + di, nx);
        for (int dj = -1; dj <= 1; ++dj) {
            const double wy = cw[3 + 1 - dj];
            if (wy == 0.0) continue;
            const int cj = wrap_cell(iy + dj, ny);
            for (int dk = -1; dk <= 1; ++dk) {
                const double w = wx * wy * cw[6 + 1 - dk];
                if (w == 0.0) continue;
                const int64_t cell = ((int64_t)wrap_cell(iz + dk, nz) * ny + cj) * nx + ci;
                h[0] += w * m[cell];
                h[1] += w * m[Nc + cell];
                h[2] += w * m[2 * Nc + cell];
            }
        }
    }
}

// H of every atom: the mesh field of the 27 cells around the atom's cell with
// the deposit's weights; written in the spin layout (colour-major t).
__global__ __launch_bounds__(kBlock) void k_demag_gather(
    const double* __restrict__ mesh, const Rep* __restrict__ reps, const Lad* __restrict__ lads,
    const DemagRep* __restrict__ dreps, const double* __restrict__ cic, int only, double* __restrict__ H)
{
    const Rep R = reps[blockIdx.y];
    if (!demag_slot(R, lads, only)) return;
    const int64_t tt = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (tt >= R.n) return;
    uint32_t t = (uint32_t)tt;    // colour-major index of the atom
    int ix, iy, iz, type;
    if (R.cell == ISING_CELL_SC) {
        // lanes in x-fastest site order, so that each wavefront reads whole
        // mesh rows (in colour-major order they would take every other cell)
        const uint32_t row = t / (uint32_t)R.lx;
        ix = (int)(t - row * (uint32_t)R.lx);
        iz = (int)(row / (uint32_t)R.ly);
        iy = (int)(row - (uint32_t)iz * (uint32_t)R.ly);
        type = 0;
        t = atom_index(R, ix, iy, iz, 0);
    } else if (R.cell == ISING_CELL_HYPERKAGOME) {
        const uint32_t C = (uint32_t)R.plane * (uint32_t)R.lz;
        const uint32_t rank = t / C, cell = t - rank * C;
        const uint32_t row = cell / (uint32_t)R.lx;
        ix = (int)(cell - row * (uint32_t)R.lx);
        iz = (int)(row / (uint32_t)R.ly);
        iy = (int)(row - (uint32_t)iz * (uint32_t)R.ly);
        type = c_hk.type_of_rank[rank];
    } else {                      // bcc: atom = colour, rows of lx cells
        const uint32_t colour = t >= R.half ? 1u : 0u;
        const uint32_t q = t - colour * R.half;
        const uint32_t row = q / (uint32_t)R.hx;
        ix = (int)(q - row * (uint32_t)R.hx);
        iz = (int)(row / (uint32_t)R.ly);
        iy = (int)(row - (uint32_t)iz * (uint32_t)R.ly);
        type = (int)colour;
    }
    const double* m = mesh + dreps[blockIdx.y].moff;
    const double* base = cic + (size_t)blockIdx.y * kHkAtoms * 9;
    // the atom type is the same on the whole wavefront except where it
    // straddles two colours or ranks: there the factors are read per lane
    const int t0 = __builtin_amdgcn_readfirstlane(type);
    double h[3];
    if (__builtin_amdgcn_ballot_w64(type != t0) == 0) gather_sum(R, m, base + t0 * 9, ix, iy, iz, h);
    else gather_sum(R, m, base + type * 9, ix, iy, iz, h);
    double* o = H + R.off + t;
    o[0] = h[0];
    o[R.n] = h[1];
    o[2 * R.n] = h[2];
}

} // namespace ising

// ===========================================================================
// Host side.
using namespace ising;

struct LadHost {
    ising_lattice lat;
    ising_hamiltonian ham;
    int n;
    int start;       // first temperature position / slot
    double cone;
    std::vector<double> betas;
};

// What a captured schedule period depends on besides the context's fixed
// buffers: the schedule, the launch tuning (kernel shapes and the fused
// layout), the kernel variants (learned, demagnetisation) and the field
// buffer the half sweeps read.
struct PtGraphKey {
    int32_t swap_every = -1, measure_every = -1, winding_every = -1, winding_layer = -1;
    int64_t period = 0;
    uint32_t tune = 0;
    int32_t learned = 0, demag = 0, records = 0;
    const void* demag_buf = nullptr;
    bool operator==(const PtGraphKey& o) const
    {
        return swap_every == o.swap_every && measure_every == o.measure_every && winding_every == o.winding_every &&
               winding_layer == o.winding_layer && period == o.period && tune == o.tune && learned == o.learned &&
               demag == o.demag && records == o.records && demag_buf == o.demag_buf;
    }
};

struct ising_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t seed = 0;
    int R = 0, zmax = 0, n_blocks = 0, n_qblocks = 0, n_lblocks = 0;
    std::vector<LadHost> ladders;
    std::vector<Rep> reps;
    std::vector<int32_t> block_rep, qblock_rep, lblock_rep;
    int64_t total_floats = 0;    // floats per spin buffer (two buffers, sweep parity)
    uint32_t host_sweep = 0;
    uint32_t host_fused = 0;     // fused sweeps enqueued: parity selects the current spin buffer
    uint32_t tune = 0;           // sweep tuning flags (kTune*, fused chunk in bits 8..15)
    // fused sweep (k_sweep_fused): simple-cubic contexts whose planes fit one workgroup, or y slabs of them
    bool fused_ok = false;
    bool fused_slab = false;     // some replica's planes are cut into y slabs (the kSlab kernels)
    int fused_threads = 0, fused_wpb = 0, n_fblocks = 0;
    size_t fblock_cap = 0, fwave_cap = 0;    // capacities of the block table and per-wavefront arrays
    uint32_t fused_pstride = 0;
    size_t fused_lds = 0;
    std::vector<int32_t> fblock_rep;
    bool any_learned = false;    // some ladder takes its cone radius from the network
    std::vector<int> learned;    // per ladder
    // demagnetising field (allocated on first use)
    bool any_demag = false;      // some ladder has the demagnetisation term on
    int max_ncol = 2;            // colour launches per sweep: 3 with a hyper-kagome ladder
    int demag_interval = 1;      // sweeps between field recomputations in ising_sweep
    bool demag_ready = false;
    std::vector<DemagRep> dreps;
    int64_t dmax_cells = 0, dmax_modes = 0, dmax_sites = 0;
    int64_t dft_tiles[3] = { 0, 0, 0 };   // workgroups per slot of the line transforms, per axis
    int64_t dft_tiles_zp = 0;             // the same for k_demag_zproject (when dft_reg[2])
    size_t dft_lds_zp = 0;
    bool dft_reg[3] = { true, true, true };   // k_dft_lines_r for that axis (every P, Q <= kDftRadix)
    size_t dft_lds[3] = { 0, 0, 0 };      // their dynamic LDS bytes
    double* d_demag = nullptr;   // FP64 field per site, spin layout (offsets of d_spins)
    double* d_dmesh = nullptr;   // real mesh [3][cells] per slot
    double2* d_dk1 = nullptr;    // half-complex meshes [3][modes] per slot (ping-pong)
    double2* d_dk2 = nullptr;
    double2* d_tw = nullptr;     // twiddle tables e^{-2 pi i m/n}
    double* d_cic = nullptr;     // weight factors, kCicPerSlot per slot
    DemagRep* d_dreps = nullptr;

    // learned proposal
    MlpDev* d_mlp = nullptr;
    float* d_mlp_prm = nullptr;  // FP32 master parameters
    float* d_mlp_m1 = nullptr;   // Adam moments
    float* d_mlp_m2 = nullptr;
    float* d_mlp_grad = nullptr; // [kMlpParams + 1], last = loss sum
    float* d_wave_grad = nullptr;
    int wave_grad_waves = 0;
    float* d_grad_part = nullptr;   // [kMlpReduceGroups][kMlpParams + 1]
    float* d_recs = nullptr;     // training records [cap][kRecord]
    int64_t n_recs = 0, cap_recs = 0;
    uint32_t collect_calls = 0;
    int adam_step = 0;

    float* d_spins = nullptr;
    Rep* d_reps = nullptr;
    Lad* d_lads = nullptr;
    Lad* d_lad_override = nullptr;
    int32_t* d_block_rep = nullptr;
    int32_t* d_qblock_rep = nullptr;
    int32_t* d_lblock_rep = nullptr;
    float* d_beta_slot = nullptr;
    double* d_beta_temp = nullptr;
    int32_t* d_slot_of_temp = nullptr;
    int32_t* d_lad_of_pos = nullptr;
    int32_t* d_lad_start = nullptr;
    int32_t* d_lad_n = nullptr;
    double* d_partial = nullptr;
    double* d_energy = nullptr;
    double* d_mag = nullptr;
    double* d_qpartial = nullptr;
    double* d_qraw = nullptr;
    long long* d_qint = nullptr;
    uint32_t* d_ctr = nullptr;  // [0] sweep, [1] meas, [2] winding meas, [3] swap round, [4] fused sweeps
    unsigned long long* d_acc = nullptr;       // [R] per-slot totals (filled on request)
    unsigned long long* d_acc_wave = nullptr;  // [n_blocks * waves per block]
    double* d_wave_part = nullptr;             // [n_blocks * waves per block][4]
    int32_t* d_fblock_rep = nullptr;           // [fblock_cap]
    unsigned long long* d_acc_wave_f = nullptr;  // [fwave_cap]: one counter per fused wavefront
    unsigned long long* d_acc_base = nullptr;    // [R] counters of earlier fused layouts
    double* d_wave_part_f = nullptr;             // [fwave_cap][4]
    unsigned long long* d_swap_att = nullptr;
    unsigned long long* d_swap_acc = nullptr;
    int32_t* d_sites = nullptr; float* d_props = nullptr; double* d_de = nullptr; int de_cap = 0;
    // ising_pt_run: series buffers (grown, never shrunk), their descriptor,
    // and the instantiated graph of the last schedule period (replaced when
    // its key changes, released with the context)
    double* ser_e = nullptr; double* ser_m = nullptr; double* ser_qr = nullptr; long long* ser_qi = nullptr;
    size_t ser_cap_e = 0, ser_cap_q = 0;         // entries of ser_e and of ser_qr
    SeriesDesc* d_series = nullptr;
    unsigned char* h_stage = nullptr;            // pinned host staging of the series copy-out
    size_t h_stage_cap = 0;
    hipGraph_t pt_graph = nullptr;
    hipGraphExec_t pt_exec = nullptr;
    PtGraphKey pt_key;
    // checkpoint (ising_save_state)
    float* ck_spins = nullptr; int32_t* ck_sot = nullptr; float* ck_bslot = nullptr; double* ck_btemp = nullptr;
    uint32_t* ck_ctr = nullptr; uint32_t ck_host_sweep = 0, ck_host_fused = 0; bool ck_valid = false;
};

// Series staging of ising_pt_run_static (see static_scratch_setup below).
static __device__ __align__(256) unsigned char scratch[ISING_STATIC_SCRATCH_BYTES];

namespace {

thread_local std::string g_create_err;

int set_err(ising_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                               \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_err((ctx), ISING_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(ctx)                                                                \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return set_err((ctx), ISING_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(ising_ctx* c, T** p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, e == hipErrorOutOfMemory ? ISING_ERR_OUT_OF_MEMORY : ISING_ERR_HIP,
                       "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    return ISING_OK;
}

// Per-axis DMI vectors d_a (bulk D e_a; reference interfacial d_x = -D e_y,
// d_y = D e_x). On bcc a bond b with unit vector sgn(b)/sqrt 3 carries
// d_b = sum_a sgn_a(b) d_a / sqrt 3, so the stored vectors are scaled by
// 1/sqrt 3 and the kernels combine them with the signs.
Lad make_lad(const ising_hamiltonian& h, double cone, int cell)
{
    Lad L;
    memset(&L, 0, sizeof L);
    double d[9] = { 0 };
    if (h.dmi_type == ISING_DMI_BULK) { d[0] = h.D; d[4] = h.D; d[8] = h.D; }
    else if (h.dmi_type == ISING_DMI_INTERFACIAL_REF) { d[1] = -h.D; d[3] = h.D; }
    if (cell == ISING_CELL_BCC)
        for (double& v : d) v /= std::sqrt(3.0);
    L.J = (float)h.J; L.Bz = (float)h.Bz; L.Kz = (float)h.Kz; L.r = (float)cone;
    L.lr8 = (float)(std::log2(cone) - 8.0);
    L.zsf = (float)h.zeeman_total_scale;
    for (int k = 0; k < 9; ++k) { L.d[k] = (float)d[k]; L.dd[k] = d[k]; }
    L.dmi = h.dmi_type;
    L.Jd = h.J; L.Bzd = h.Bz; L.Kzd = h.Kz; L.zs = h.zeeman_total_scale;
    L.demag = h.demag ? 1 : 0;
    L.mu0 = h.demag ? ISING_MU0_REF : 0.0;
    return L;
}

int atoms_of(int cell) { return cell == ISING_CELL_HYPERKAGOME ? kHkAtoms : (cell == ISING_CELL_BCC ? 2 : 1); }

// The hyper-kagome table. Atom positions of hyper_kagome_unit_cell
// (src/unit_cells.f90:32-77) in units of the edge a = 6; each atom's bonds
// are the first shell of initAtomShells (src/chainMesh.f90:495-591): the
// atoms of the cells within two cells sorted by distance, the shell being the
// nearest one and each next one at most 0.1 a (makeChainMesh's threshold,
// :783) further than the one before. Colours: the fewest for which the atom
// types' bond graph has a proper colouring, the lexicographically first one
// (each type the smallest colour no earlier bonded type holds, backtracking).
// Storage ranks run colour by colour, types in order within a colour.
static const double kHkFrac[kHkAtoms][3] = {
    { -0.5, 0.0, 0.5 },   { -0.25, 0.75, 0.5 }, { -0.5, 0.75, 0.25 }, { -0.25, 0.25, 0.0 },
    { -0.5, 0.25, 0.75 }, { -0.25, 0.5, 0.75 }, { -0.75, 0.5, 0.25 }, { 0.0, 0.5, 0.5 },
    { 0.0, 0.25, 0.25 },  { -0.75, 0.75, 0.0 }, { 0.0, 0.0, 0.0 },    { -0.75, 0.0, 0.75 } };

static bool hk_colour_from(int t, int k, const bool (&adj)[kHkAtoms][kHkAtoms], int (&col)[kHkAtoms])
{
    if (t == kHkAtoms) return true;
    for (int cc = 0; cc < k; ++cc) {
        bool ok = true;
        for (int u = 0; u < t && ok; ++u) ok = !(adj[t][u] && col[u] == cc);
        if (!ok) continue;
        col[t] = cc;
        if (hk_colour_from(t + 1, k, adj, col)) return true;
    }
    return false;
}

static bool hk_build_table(HkTable& T)
{
    struct Cand { double d; int u, s[3]; };
    const double a = 6.0, thr = 0.1 * a;
    bool adj[kHkAtoms][kHkAtoms] = {};
    std::vector<Cand> shell[kHkAtoms];
    for (int t = 0; t < kHkAtoms; ++t) {
        std::vector<Cand> v;
        for (int u = 0; u < kHkAtoms; ++u)
            for (int sx = -2; sx <= 2; ++sx)
                for (int sy = -2; sy <= 2; ++sy)
                    for (int sz = -2; sz <= 2; ++sz) {
                        if (u == t && sx == 0 && sy == 0 && sz == 0) continue;
                        const double d[3] = { (kHkFrac[u][0] + sx - kHkFrac[t][0]) * a,
                                              (kHkFrac[u][1] + sy - kHkFrac[t][1]) * a,
                                              (kHkFrac[u][2] + sz - kHkFrac[t][2]) * a };
                        v.push_back(Cand{ std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), u, { sx, sy, sz } });
                    }
        std::stable_sort(v.begin(), v.end(), [](const Cand& x, const Cand& y) { return x.d < y.d; });
        shell[t].push_back(v[0]);
        for (size_t j = 1; j < v.size() && v[j].d - v[j - 1].d <= thr; ++j) shell[t].push_back(v[j]);
        if ((int)shell[t].size() != kHkBonds) return false;
        for (const Cand& b : shell[t]) {
            if (b.u == t || std::abs(b.s[0]) > 1 || std::abs(b.s[1]) > 1 || std::abs(b.s[2]) > 1) return false;
            adj[t][b.u] = adj[b.u][t] = true;
        }
    }
    int col[kHkAtoms] = {};
    int k = 1;
    while (k <= kHkAtoms && !hk_colour_from(0, k, adj, col)) ++k;
    if (k * kHkPerColour != kHkAtoms) return false;     // three colours of four types
    int fill[kHkAtoms] = {};
    for (int t = 0; t < kHkAtoms; ++t) {
        if (fill[col[t]] == kHkPerColour) return false;
        const int r = col[t] * kHkPerColour + fill[col[t]]++;
        T.rank_of_type[t] = r;
        T.type_of_rank[r] = t;
    }
    for (int t = 0; t < kHkAtoms; ++t) {
        for (int x = 0; x < 3; ++x) T.frac[t][x] = kHkFrac[t][x];
        for (int b = 0; b < kHkBonds; ++b) {
            const Cand& C = shell[t][b];
            HkBond& B = T.bond[T.rank_of_type[t]][b];
            B.rank = T.rank_of_type[C.u];
            B.dx = C.s[0]; B.dy = C.s[1]; B.dz = C.s[2];
            for (int x = 0; x < 3; ++x) {
                B.rd[x] = (kHkFrac[C.u][x] + C.s[x] - kHkFrac[t][x]) * a / C.d;
                B.r[x] = (float)B.rd[x];
            }
        }
    }
    return true;
}

// The table, built once per process (nullptr if the construction failed).
const HkTable* hk_table()
{
    static HkTable T;
    static const bool ok = hk_build_table(T);
    return ok ? &T : nullptr;
}

int check_lattice(ising_ctx* c, const ising_lattice& l)
{
    const int n[3] = { l.lx, l.ly, l.lz };
    if (l.cell != ISING_CELL_SC && l.cell != ISING_CELL_BCC && l.cell != ISING_CELL_HYPERKAGOME)
        return set_err(c, ISING_ERR_BAD_DIMS, "unknown unit cell %d", l.cell);
    const bool sc = l.cell == ISING_CELL_SC;
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return set_err(c, ISING_ERR_BAD_DIMS, "axis %d has %d cells", a, n[a]);
        // bcc and hyper-kagome colour by atom, for any length; a periodic axis
        // of one cell would bond an atom to the same neighbour twice (bcc) or
        // bring atoms of other shells to nearest-neighbour distance
        // (hyper-kagome)
        if (!sc && l.periodic[a] && n[a] < 2)
            return set_err(c, ISING_ERR_BAD_DIMS, "periodic %s axis %d needs >= 2 cells (got %d)",
                           l.cell == ISING_CELL_BCC ? "bcc" : "hyper-kagome", a, n[a]);
        if (sc && l.periodic[a] && (n[a] < 4 || (n[a] & 1)))
            return set_err(c, ISING_ERR_BAD_DIMS,
                           "periodic axis %d needs an even length >= 4 for the two-colour sweep (got %d)", a, n[a]);
    }
    if (sc && (l.lx & 1)) return set_err(c, ISING_ERR_BAD_DIMS, "lx must be even (got %d)", l.lx);
    const int64_t N = (int64_t)l.lx * l.ly * l.lz * atoms_of(l.cell);
    if (N >= (1ll << 31)) return set_err(c, ISING_ERR_BAD_DIMS, "lattice of %lld sites too large", (long long)N);
    return ISING_OK;
}

int check_ham(ising_ctx* c, const ising_hamiltonian& h)
{
    if (h.dmi_type < ISING_DMI_NONE || h.dmi_type > ISING_DMI_INTERFACIAL_REF)
        return set_err(c, ISING_ERR_INVALID_ARG, "unknown dmi_type %d", h.dmi_type);
    const double v[5] = { h.J, h.D, h.Bz, h.Kz, h.zeeman_total_scale };
    for (double x : v)
        if (!std::isfinite(x)) return set_err(c, ISING_ERR_INVALID_ARG, "non-finite Hamiltonian parameter");
    if (h.demag != 0 && h.demag != 1) return set_err(c, ISING_ERR_INVALID_ARG, "demag must be 0 or 1 (got %d)", h.demag);
    return ISING_OK;
}

// The demagnetising field is periodic (a discrete Fourier transform of the
// whole lattice, src/chainMesh.f90:717-728).
int check_demag_lattice(ising_ctx* c, const ising_lattice& l)
{
    if (!l.periodic[0] || !l.periodic[1] || !l.periodic[2])
        return set_err(c, ISING_ERR_UNSUPPORTED, "the demagnetising field needs a fully periodic lattice");
    if (l.lx > kDftMaxLen || l.ly > kDftMaxLen || l.lz > kDftMaxLen)
        return set_err(c, ISING_ERR_UNSUPPORTED, "the demagnetising field takes at most %d cells per axis",
                       kDftMaxLen);
    return ISING_OK;
}

int upload_lads(ising_ctx* c)
{
    std::vector<Lad> l;
    c->any_demag = false;
    for (size_t k = 0; k < c->ladders.size(); ++k) {
        c->any_demag = c->any_demag || c->ladders[k].ham.demag != 0;
        l.push_back(make_lad(c->ladders[k].ham, c->ladders[k].cone, c->ladders[k].lat.cell));
        l.back().learned = k < c->learned.size() ? c->learned[k] : 0;
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_lads, l.data(), l.size() * sizeof(Lad), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

// Index colour*half + q in a replica's component arrays of the site with
// reference index `site` (SC: (ix*ly + iy)*lz + iz; bcc and hyper-kagome:
// that cell number times the atoms per cell plus the atom).
int64_t dev_index(const Rep& r, int64_t site)
{
    if (r.cell == ISING_CELL_HYPERKAGOME) {
        const int type = (int)(site % kHkAtoms);
        int64_t t = site / kHkAtoms;
        const int64_t iz = t % r.lz; t /= r.lz;
        const int64_t iy = t % r.ly;
        const int64_t ix = t / r.ly;
        return (int64_t)hk_table()->rank_of_type[type] * ((int64_t)r.lx * r.ly * r.lz) + (iz * r.ly + iy) * r.lx + ix;
    }
    const bool bcc = r.cell == ISING_CELL_BCC;
    int64_t t = bcc ? site >> 1 : site;
    const int64_t iz = t % r.lz; t /= r.lz;
    const int64_t iy = t % r.ly;
    const int64_t ix = t / r.ly;
    const int64_t k = bcc ? (site & 1) : ((ix + iy + iz) & 1);
    return k * r.half + (iz * r.ly + iy) * r.hx + (bcc ? ix : ix >> 1);
}

int check_replica(ising_ctx* c, int r)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (r < 0 || r >= c->R) return set_err(c, ISING_ERR_INVALID_ARG, "replica %d out of range [0,%d)", r, c->R);
    return ISING_OK;
}

// ---- launch helpers --------------------------------------------------------

constexpr int kCounters = 5;

SpinBuf spin_buf(const ising_ctx* c) { return SpinBuf{ c->d_spins, c->total_floats, c->d_ctr + 4 }; }

// The buffer that holds the state after the sweeps the host has enqueued.
float* cur_spins(const ising_ctx* c) { return c->d_spins + (int64_t)(c->host_fused & 1u) * c->total_floats; }

// z planes per fused-sweep workgroup (0: two half-sweep launches per sweep).
int fused_lc(const ising_ctx* c) { return (int)((c->tune >> 8) & 0xFFu); }
int fused_hy(const ising_ctx* c) { return (int)((c->tune & kTuneSlabMask) >> kTuneSlabShift); }

// Sweeps run as fused launches (which swap the spin buffers): the fused flag
// is set, no ladder has the demagnetisation term, which only the in-place
// half sweeps carry, and no ladder takes its radius from the network (the
// network's registers do not fit beside the fused sweep's windows under the
// 128 per thread of a 1024-thread workgroup, so the learned sampler keeps
// the half sweeps under every tuning).
bool fused_active(const ising_ctx* c) { return fused_lc(c) > 0 && !c->any_demag && !c->any_learned; }

// t0/t1: start/stop events in the dispatch packet (timed launches only;
// graph-captured launches take the plain path).
template <bool kMeasure, bool kLearned, bool kExact = false, bool kThree = false>
void launch_half_kernel(ising_ctx* c, uint32_t off, int colour, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr)
{
    if (t0) {
        hipExtLaunchKernelGGL((k_half_sweep<kMeasure, kLearned, kExact, kThree>), dim3(c->n_blocks), dim3(kBlock), 0,
                              c->stream, t0, t1, 0u, spin_buf(c), c->d_reps, c->d_lads, (const float*)c->d_beta_slot,
                              (const int32_t*)c->d_block_rep, (const uint32_t*)c->d_ctr, off, colour,
                              (uint32_t)c->seed, (uint32_t)(c->seed >> 32), c->d_acc_wave,
                              kMeasure ? c->d_wave_part : (double*)nullptr, c->tune & 0xFFu, (const MlpDev*)c->d_mlp,
                              (const double*)c->d_demag);
        return;
    }
    hipLaunchKernelGGL((k_half_sweep<kMeasure, kLearned, kExact, kThree>), dim3(c->n_blocks), dim3(kBlock), 0, c->stream,
                       spin_buf(c), c->d_reps, c->d_lads, c->d_beta_slot, c->d_block_rep,
                       (const uint32_t*)c->d_ctr, off, colour, (uint32_t)c->seed, (uint32_t)(c->seed >> 32), c->d_acc_wave,
                       kMeasure ? c->d_wave_part : (double*)nullptr, c->tune & 0xFFu, (const MlpDev*)c->d_mlp,
                       (const double*)c->d_demag);
}

template <bool kMeasure, bool kExact, bool kSlab>
void launch_fused_variant(ising_ctx* c, uint32_t off, hipEvent_t t0, hipEvent_t t1)
{
    if (t0) {
        hipExtLaunchKernelGGL((k_sweep_fused<kMeasure, kExact, kSlab>), dim3(c->n_fblocks), dim3(c->fused_threads),
                              (uint32_t)c->fused_lds, c->stream, t0, t1, 0u, spin_buf(c), c->d_reps, c->d_lads,
                              (const float*)c->d_beta_slot, (const int32_t*)c->d_fblock_rep, (const uint32_t*)c->d_ctr,
                              off, fused_lc(c), c->fused_pstride, (uint32_t)c->seed, (uint32_t)(c->seed >> 32),
                              c->d_acc_wave_f, kMeasure ? c->d_wave_part_f : (double*)nullptr, c->tune & 0xFFu);
        return;
    }
    hipLaunchKernelGGL((k_sweep_fused<kMeasure, kExact, kSlab>), dim3(c->n_fblocks), dim3(c->fused_threads), c->fused_lds,
                       c->stream, spin_buf(c), c->d_reps, c->d_lads, c->d_beta_slot, c->d_fblock_rep,
                       (const uint32_t*)c->d_ctr, off, fused_lc(c), c->fused_pstride, (uint32_t)c->seed, (uint32_t)(c->seed >> 32),
                       c->d_acc_wave_f, kMeasure ? c->d_wave_part_f : (double*)nullptr, c->tune & 0xFFu);
}

// The slab kernels only where some replica's planes are cut into y slabs.
template <bool kMeasure, bool kExact = false>
void launch_fused_kernel(ising_ctx* c, uint32_t off, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr)
{
    if (c->fused_slab) launch_fused_variant<kMeasure, kExact, true>(c, off, t0, t1);
    else launch_fused_variant<kMeasure, kExact, false>(c, off, t0, t1);
}

// Sweep number `off` of a period (relative to the device sweep counter):
// colour 0 then colour 1 of every replica, from the buffer of the current
// parity into the other one; with `measure`, E and M of the new state are
// recorded by the same launches.
constexpr SeriesRec kNoRecord{ nullptr, nullptr, 0u };

// Start/stop events carried by the dispatch packets of the plain
// (non-measuring) sweep launches of an eagerly enqueued schedule
// (hipExtLaunchKernel): the kernels' own execution time inside the workload
// they are part of, without the barrier packets of separate event records
// (ising_time_pt_sweeps).
// Events that only time device work: without the system-scope release an
// event normally carries, recording one does not write back and invalidate
// the caches, which would add several microseconds to the launch it closes.
hipError_t timing_event(hipEvent_t* e) { return hipEventCreateWithFlags(e, hipEventDisableSystemFence); }

struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    int launches_per_pair = 1;   // half sweeps covered by one timed launch (2 for a fused sweep)
    ~LaunchTimer() { for (hipEvent_t e : ev) (void)hipEventDestroy(e); }
    // Events are created before the timed launches are enqueued: creating
    // them between launches lets the device idle while the host catches up.
    hipError_t reserve(size_t n)
    {
        while (ev.size() < n) {
            hipEvent_t e;
            hipError_t r = timing_event(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        return hipSuccess;
    }
    hipError_t next_pair(hipEvent_t& a, hipEvent_t& b)
    {
        hipError_t r = reserve(used + 2);
        if (r != hipSuccess) return r;
        a = ev[used]; b = ev[used + 1];
        used += 2;
        return hipSuccess;
    }
};

// Buffer offset of the state after `done` sweeps of a period.
uint32_t buf_off(const ising_ctx* c, int done) { return fused_active(c) ? (uint32_t)done : 0u; }

// Exact FP64 E and M of every slot for the state after `done` sweeps of the
// period; with rec, also recorded in the series.
int launch_measure_all(ising_ctx* c, int done = 0, const SeriesRec& rec = kNoRecord);

// exact: the colour-1 launch (or the fused sweep) records FP64 per-site
// energies (pair_energy_exact) instead of FP32 ones. finalize = false leaves
// the sums of a measuring sweep to launch_winding_all's combined finalize. With a hyper-kagome
// ladder every sweep is three colour launches (colour 2 leaves two-colour
// replicas alone) and measurements come from the measure kernel, the
// colour-1 pair energies needing a bipartite bond graph.
int launch_sweep(ising_ctx* c, uint32_t off, bool measure = false, const SeriesRec& rec = kNoRecord,
                 bool exact = false, LaunchTimer* timer = nullptr, bool finalize = true)
{
    const bool learned = c->any_learned;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    auto pair = [&]() -> int {
        if (timer && timer->next_pair(t0, t1) != hipSuccess) return set_err(c, ISING_ERR_HIP, "hipEventCreate");
        return ISING_OK;
    };
    int rc;
    if (fused_active(c)) {
        if (measure && exact) {
            launch_fused_kernel<true, true>(c, off);
        } else if (measure) {
            launch_fused_kernel<true>(c, off);
        } else {
            if (timer) timer->launches_per_pair = 2;
            if ((rc = pair())) return rc;
            launch_fused_kernel<false>(c, off, t0, t1);
        }
        LAUNCH_CHECK(c);
        if (measure && finalize) {
            hipLaunchKernelGGL(k_wave_finalize, dim3(c->R), dim3(kBlock), 0, c->stream, c->d_reps,
                               (const double*)c->d_wave_part_f, c->fused_wpb, c->d_energy, c->d_mag,
                               (const int32_t*)c->d_slot_of_temp, c->R, rec);
            LAUNCH_CHECK(c);
        }
        return ISING_OK;
    }
    if (c->max_ncol > 2) {
        // the kernel variant that carries the hyper-kagome branch (kept out
        // of the two-colour contexts' kernels, where it costs ~5% of a sweep)
        for (int k = 0; k < c->max_ncol; ++k) {
            if ((rc = pair())) return rc;
            if (learned) launch_half_kernel<false, true, false, true>(c, off, k, t0, t1);
            else launch_half_kernel<false, false, false, true>(c, off, k, t0, t1);
            LAUNCH_CHECK(c);
        }
        return measure ? launch_measure_all(c, 0, rec) : ISING_OK;
    }
    if ((rc = pair())) return rc;
    if (learned) launch_half_kernel<false, true>(c, off, 0, t0, t1);
    else launch_half_kernel<false, false>(c, off, 0, t0, t1);
    LAUNCH_CHECK(c);
    if (measure) {
        if (exact) {
            if (learned) launch_half_kernel<true, true, true>(c, off, 1);
            else launch_half_kernel<true, false, true>(c, off, 1);
        } else {
            if (learned) launch_half_kernel<true, true>(c, off, 1);
            else launch_half_kernel<true, false>(c, off, 1);
        }
        LAUNCH_CHECK(c);
        if (finalize)
            hipLaunchKernelGGL(k_wave_finalize, dim3(c->R), dim3(kBlock), 0, c->stream, c->d_reps,
                               (const double*)c->d_wave_part, 0, c->d_energy, c->d_mag,
                               (const int32_t*)c->d_slot_of_temp, c->R, rec);
    } else {
        if ((rc = pair())) return rc;
        if (learned) launch_half_kernel<false, true>(c, off, 1, t0, t1);
        else launch_half_kernel<false, false>(c, off, 1, t0, t1);
    }
    LAUNCH_CHECK(c);
    return ISING_OK;
}

// Mesh, twiddle and field buffers of the demagnetising field, allocated on
// first use (the field starts at zero).
int demag_setup(ising_ctx* c)
{
    if (c->demag_ready) return ISING_OK;
    std::vector<double2> tw;
    std::vector<std::pair<int, int>> tables;   // (length, first entry)
    auto table = [&](int n) {
        for (auto& t : tables)
            if (t.first == n) return t.second;
        const int first = (int)tw.size();
        for (int m = 0; m < n; ++m) {
            const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
            tw.push_back(double2{ (double)cosl(a), -(double)sinl(a) });
        }
        tables.emplace_back(n, first);
        return first;
    };
    c->dreps.clear();
    int64_t moff = 0, koff = 0;
    for (const Rep& r : c->reps) {
        DemagRep d{};
        const int64_t cells = (int64_t)r.lx * r.ly * r.lz;
        d.hx = r.lx / 2 + 1;
        const int64_t modes = (int64_t)d.hx * r.ly * r.lz;
        d.moff = moff; d.koff = koff;
        d.tw[0] = table(r.lx); d.tw[1] = table(r.ly); d.tw[2] = table(r.lz);
        // slots the kernels skip (demag_slot) take no part in the launch shapes
        const bool used = r.per == 7 && std::max(r.lx, std::max(r.ly, r.lz)) <= kDftMaxLen;
        for (int a = 0; a < 3 && used; ++a) {
            const int L = a == 0 ? r.lx : (a == 1 ? r.ly : r.lz);
            int p = 1;   // largest divisor of L not above sqrt(L)
            for (int f = 1; f * f <= L; ++f)
                if (L % f == 0) p = f;
            d.fp[a] = p;
            c->dft_reg[a] = c->dft_reg[a] && L / p <= kDftRadix;
        }
        moff += 3 * cells; koff += 3 * modes;
        c->dmax_cells = std::max(c->dmax_cells, cells);
        c->dmax_modes = std::max(c->dmax_modes, modes);
        c->dmax_sites = std::max(c->dmax_sites, r.n);
        c->dreps.push_back(d);
    }
    // weight factors of every slot, from cell 1 (any cell gives the same)
    std::vector<double> cic(c->reps.size() * kCicPerSlot, 0.0);
    for (size_t k = 0; k < c->reps.size(); ++k) {
        const Rep& r = c->reps[k];
        const int na = r.cell == ISING_CELL_HYPERKAGOME ? kHkAtoms : (r.cell == ISING_CELL_BCC ? 2 : 1);
        for (int a = 0; a < na; ++a)
            for (int ax = 0; ax < 3; ++ax) {
                const int n = ax == 0 ? r.lx : (ax == 1 ? r.ly : r.lz);
                const double frac = r.cell == ISING_CELL_HYPERKAGOME ? hk_table()->frac[a][ax] : 0.5 * a;
                const int centre = 1 % n;
                for (int d = 0; d < 3; ++d)
                    cic[(k * kHkAtoms + a) * 9 + ax * 3 + d] = cic_axis(wrap_cell(centre + d - 1, n), frac, centre, n);
            }
    }
    // lines per workgroup, once the kernel of each axis is known: one item per
    // thread (k_dft_lines_r) or two buffers within kDftLdsTarget (k_dft_lines)
    for (size_t k = 0; k < c->reps.size(); ++k) {
        const Rep& r = c->reps[k];
        DemagRep& d = c->dreps[k];
        if (r.per != 7 || std::max(r.lx, std::max(r.ly, r.lz)) > kDftMaxLen) continue;
        const int64_t modes = (int64_t)d.hx * r.ly * r.lz;
        const int64_t lines[3] = { 3 * (int64_t)r.ly * r.lz, 3 * modes / r.ly, 3 * modes / r.lz };
        for (int a = 0; a < 3; ++a) {
            const int L = a == 0 ? r.lx : (a == 1 ? r.ly : r.lz);
            size_t lds;
            if (c->dft_reg[a]) {
                d.tl[a] = std::min(kDftMaxLines, kBlock / (L / d.fp[a]));
                lds = (size_t)(L * (d.tl[a] + 1) + L) * sizeof(double2);
            } else {
                d.tl[a] = std::max(1, std::min(kDftMaxLines, (kDftLdsTarget - L) / (2 * L) - 1));
                lds = (size_t)(2 * L * (d.tl[a] + 1) + L) * sizeof(double2);
            }
            c->dft_tiles[a] = std::max(c->dft_tiles[a], (lines[a] + d.tl[a] - 1) / d.tl[a]);
            c->dft_lds[a] = std::max(c->dft_lds[a], lds);
        }
        if (c->dft_reg[2]) {      // k_demag_zproject: 3 components x Q x tz items
            d.tz = std::min(kDftMaxLines, kBlock / (3 * (r.lz / d.fp[2])));
            const int64_t krow = (int64_t)r.ly * d.hx;
            c->dft_tiles_zp = std::max(c->dft_tiles_zp, (krow + d.tz - 1) / d.tz);
            c->dft_lds_zp = std::max(c->dft_lds_zp, (size_t)(3 * r.lz * (d.tz + 1) + r.lz) * sizeof(double2));
        }
    }
    // A set-up that failed part way keeps what it allocated (the sizes depend
    // only on the context's lattices), so a retry neither leaks nor
    // allocates twice; the context frees them at destruction.
    auto once = [&](auto** p, size_t n) { return *p ? ISING_OK : dalloc(c, p, n); };
    int rc;
    if ((rc = once(&c->d_demag, (size_t)c->total_floats)) || (rc = once(&c->d_dmesh, (size_t)moff)) ||
        (rc = once(&c->d_dk1, (size_t)koff)) || (rc = once(&c->d_dk2, (size_t)koff)) ||
        (rc = once(&c->d_tw, tw.size())) || (rc = once(&c->d_dreps, c->dreps.size())) ||
        (rc = once(&c->d_cic, cic.size())))
        return rc;
    HIP_TRY(c, hipMemcpyAsync(c->d_cic, cic.data(), cic.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // LDS beyond the default 64 KB per workgroup (axes longer than 512)
    const void* fns[2][4] = { { (const void*)k_dft_lines<kR2cX>, (const void*)k_dft_lines<kC2cY>,
                                (const void*)k_dft_lines<kC2cZ>, (const void*)k_dft_lines<kC2rX> },
                              { (const void*)k_dft_lines_r<kR2cX>, (const void*)k_dft_lines_r<kC2cY>,
                                (const void*)k_dft_lines_r<kC2cZ>, (const void*)k_dft_lines_r<kC2rX> } };
    const int fn_axis[4] = { 0, 1, 2, 0 };
    if (c->dft_reg[2] && c->dft_lds_zp > 0)
        HIP_TRY(c, hipFuncSetAttribute((const void*)k_demag_zproject, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)c->dft_lds_zp));
    for (int f = 0; f < 4; ++f)
        if (c->dft_lds[fn_axis[f]] > 0)
            HIP_TRY(c, hipFuncSetAttribute(fns[c->dft_reg[fn_axis[f]] ? 1 : 0][f],
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->dft_lds[fn_axis[f]]));
    HIP_TRY(c, hipMemcpyAsync(c->d_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_dreps, c->dreps.data(), c->dreps.size() * sizeof(DemagRep), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_demag, 0, (size_t)c->total_floats * sizeof(double), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->demag_ready = true;
    return ISING_OK;
}

// Recomputes the field of the selected slots (demag_slot) from the current
// spins, on the context's stream.
int launch_demag(ising_ctx* c, int only)
{
    int rc = demag_setup(c);
    if (rc) return rc;
    const unsigned R = (unsigned)c->R;
    auto grid = [&](int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock), R); };
    hipStream_t s = c->stream;
    const DemagRep* dr = c->d_dreps;
    const double2* tw = c->d_tw;
    hipLaunchKernelGGL(k_demag_deposit, grid(c->dmax_cells), dim3(kBlock), 0, s, (const float*)cur_spins(c),
                       c->d_reps, c->d_lads, dr, (const double*)c->d_cic, only, c->d_dmesh);
    auto lines = [&](int a) { return dim3((unsigned)std::max<int64_t>(1, c->dft_tiles[a]), R); };
    // one line transform; k_dft_lines_r where every slot's factors fit
    auto dft = [&](auto mode, const auto* in, int sign, auto* out) {
        constexpr int m = decltype(mode)::value;
        constexpr int a = kDftAxis<m>;
        if (c->dft_reg[a])
            hipLaunchKernelGGL(k_dft_lines_r<m>, lines(a), dim3(kBlock), c->dft_lds[a], s, in, c->d_reps, c->d_lads,
                               dr, only, tw, sign, out);
        else
            hipLaunchKernelGGL(k_dft_lines<m>, lines(a), dim3(kBlock), c->dft_lds[a], s, in, c->d_reps, c->d_lads,
                               dr, only, tw, sign, out);
    };
    using R2c = std::integral_constant<int, kR2cX>;
    using C2cY = std::integral_constant<int, kC2cY>;
    using C2cZ = std::integral_constant<int, kC2cZ>;
    using C2r = std::integral_constant<int, kC2rX>;
    dft(R2c{}, (const double*)c->d_dmesh, -1, c->d_dk1);
    dft(C2cY{}, (const double2*)c->d_dk1, -1, c->d_dk2);
    if (c->dft_reg[2]) {          // z transforms and projection in one pass, in place
        hipLaunchKernelGGL(k_demag_zproject, dim3((unsigned)std::max<int64_t>(1, c->dft_tiles_zp), R), dim3(kBlock),
                           c->dft_lds_zp, s, c->d_dk2, c->d_reps, c->d_lads, dr, only, tw);
    } else {
        dft(C2cZ{}, (const double2*)c->d_dk2, -1, c->d_dk1);
        hipLaunchKernelGGL(k_demag_project, grid(c->dmax_modes), dim3(kBlock), 0, s, c->d_dk1, c->d_reps,
                           c->d_lads, dr, only);
        dft(C2cZ{}, (const double2*)c->d_dk1, +1, c->d_dk2);
    }
    dft(C2cY{}, (const double2*)c->d_dk2, +1, c->d_dk1);
    dft(C2r{}, (const double2*)c->d_dk1, +1, c->d_dmesh);
    hipLaunchKernelGGL(k_demag_gather, grid(c->dmax_sites), dim3(kBlock), 0, s, (const double*)c->d_dmesh,
                       c->d_reps, c->d_lads, dr, (const double*)c->d_cic, only, c->d_demag);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

int launch_measure_all(ising_ctx* c, int done, const SeriesRec& rec)
{
    hipLaunchKernelGGL(k_measure, dim3(c->n_blocks), dim3(kBlock), 0, c->stream,
                       spin_buf(c), buf_off(c, done), c->d_reps, c->d_lads, (const Lad*)nullptr, c->d_block_rep, 0, c->d_partial,
                       (const double*)c->d_demag);
    LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_measure_finalize, dim3(c->R), dim3(kBlock), 0, c->stream,
                       c->d_reps, c->d_partial, 0, c->d_energy, c->d_mag, (const int32_t*)c->d_slot_of_temp, c->R, rec);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

// with_e: also finish the E/M sums of the measuring sweep just launched (its
// launch_sweep ran with finalize = false), in the same finalize launch.
int launch_winding_all(ising_ctx* c, int layer = -1, int done = 0, const SeriesRec& rec = kNoRecord,
                       bool with_e = false, const SeriesRec& rec_e = kNoRecord)
{
    const uint32_t off = buf_off(c, done);
    const int zl = layer < 0 ? c->zmax : 1;       // finalize blocks per slot
    if (layer < 0) {
        hipLaunchKernelGGL(k_winding, dim3(c->n_qblocks), dim3(kBlock), 0, c->stream,
                           spin_buf(c), off, c->d_reps, (const int32_t*)c->d_qblock_rep, -1, c->d_qpartial);
    } else {
        hipLaunchKernelGGL(k_winding, dim3(c->n_lblocks), dim3(kBlock), 0, c->stream,
                           spin_buf(c), off, c->d_reps, (const int32_t*)c->d_lblock_rep, layer, c->d_qpartial);
    }
    LAUNCH_CHECK(c);
    if (with_e) {
        const bool fz = fused_active(c);
        hipLaunchKernelGGL(k_finalize_q_and_e, dim3(c->R * zl + c->R), dim3(kBlock), 0, c->stream,
                           c->d_reps, (const double*)c->d_qpartial, c->zmax, layer, c->d_qraw, c->d_qint, rec,
                           (const double*)(fz ? c->d_wave_part_f : c->d_wave_part), fz ? c->fused_wpb : 0,
                           c->d_energy, c->d_mag, rec_e, (const int32_t*)c->d_slot_of_temp, c->R);
    } else {
        hipLaunchKernelGGL(k_winding_finalize, dim3(c->R * zl), dim3(kBlock), 0, c->stream,
                           c->d_reps, c->d_qpartial, c->zmax, layer, c->d_qraw, c->d_qint,
                           (const int32_t*)c->d_slot_of_temp, c->R, rec);
    }
    LAUNCH_CHECK(c);
    return ISING_OK;
}

int launch_swap(ising_ctx* c, const SeriesRec& rec = kNoRecord)
{
    hipLaunchKernelGGL(k_swap, dim3(1), dim3(kBlock), 0, c->stream,
                       (const double*)c->d_energy, (const double*)c->d_beta_temp, c->d_slot_of_temp, c->d_beta_slot,
                       c->d_lad_of_pos, c->d_lad_start, c->d_lad_n, c->R, c->d_ctr + 3,
                       (uint32_t)c->seed, (uint32_t)(c->seed >> 32), c->d_swap_att, c->d_swap_acc,
                       (const double*)c->d_mag, rec);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

// Device counters after n sweeps (graph-capturable); advance_host follows
// once they are enqueued.
int launch_advance(ising_ctx* c, uint32_t n, uint32_t meas = 0u, uint32_t q = 0u)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->d_ctr, n, meas, q, fused_active(c) ? n : 0u);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

void advance_host(ising_ctx* c, uint32_t n)
{
    c->host_sweep += n;
    if (fused_active(c)) c->host_fused += n;
}

// Where the records of a run go: the device descriptor (nullptr: records
// are computed, not stored).
struct SeriesDev {
    const SeriesDesc* desc = nullptr;
};

// Sweeps [0, count) of a schedule period, with swaps / records after every
// sweep k+1 that the schedule names. Offsets are relative to the device sweep
// counter, which is advanced by `count` at the end, so the same launch
// sequence can be captured once and replayed.
int enqueue_period(ising_ctx* c, const ising_schedule& s, int count, const SeriesDev& sd, LaunchTimer* timer = nullptr)
{
    int rc;
    uint32_t nm = 0, nq = 0;      // records of each series so far in this period
    for (int k = 0; k < count; ++k) {
        const int done = k + 1;
        const bool do_swap = s.swap_every > 0 && done % s.swap_every == 0;
        const bool do_meas = s.measure_every > 0 && done % s.measure_every == 0;
        const bool do_q = s.winding_every > 0 && done % s.winding_every == 0;
        const SeriesRec rm = (do_meas && sd.desc) ? SeriesRec{ sd.desc, c->d_ctr + 1, nm } : kNoRecord;
        // Measurements ride on the sweep launches. Before a swap, whose
        // exponent takes FP64 per-site energies, the half sweep's colour-1
        // launch or the fused sweep computes them (three-colour sweeps leave
        // them to the measure kernel); the swap kernel then records E and M
        // at the new temperatures.
        const bool exact = do_swap && c->max_ncol == 2;
        // a sweep that records E/M on its launches and also the charge finishes
        // both sums in one launch (before a swap E must be final first)
        const bool merge = do_meas && !do_swap && do_q && c->max_ncol == 2;
        if ((rc = launch_sweep(c, (uint32_t)k, (do_meas && !do_swap) || exact, do_swap ? kNoRecord : rm, exact, timer,
                               !merge)))
            return rc;
        if (do_swap && !exact)
            if ((rc = launch_measure_all(c, done))) return rc;
        if (do_swap)
            if ((rc = launch_swap(c, rm))) return rc;
        if (do_q) {
            const SeriesRec rq = sd.desc ? SeriesRec{ sd.desc, c->d_ctr + 2, nq } : kNoRecord;
            if ((rc = launch_winding_all(c, s.winding_layer - 1, done, rq, merge, rm))) return rc;
        }
        nm += do_meas ? 1u : 0u;
        nq += do_q ? 1u : 0u;
    }
    return launch_advance(c, (uint32_t)count, nm, nq);
}

// Fused-sweep resources: available when every replica is simple cubic and
// each z plane either fits one workgroup (two column pairs per thread, 2048
// sites of one colour) or can be cut into y slabs whose windows of at least
// six rows do. The per-layout arrays (block table, per-wavefront counters and
// sums) are sized by fused_layout.
int fused_setup(ising_ctx* c)
{
    bool all_sc = true, fits = true;
    int64_t nmax = 0;
    for (const Rep& r : c->reps) {
        all_sc = all_sc && r.cell == ISING_CELL_SC;
        fits = fits && (r.plane <= kFusedMaxPairs || r.hx * (kSlabHalo + 2) <= kFusedMaxPairs);
        nmax = std::max(nmax, r.n);
    }
    int lds_max = 0, optin = 0;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, c->device);
    lds_max = std::max(lds_max, optin);
    // 2 colours x 2 parities of (window + one zero slot), 16 bytes a site
    const size_t lds = (size_t)4 * (kFusedMaxPairs + 1) * sizeof(float4);
    // the kernel addresses a replica's spins through 32-bit buffer offsets
    c->fused_ok = all_sc && fits && !c->reps.empty() && lds <= (size_t)lds_max && 12 * nmax < ((int64_t)1 << 31);
    if (c->fused_ok) {
        const void* fns[] = { (const void*)k_sweep_fused<false>, (const void*)k_sweep_fused<true>,
                              (const void*)k_sweep_fused<true, true>, (const void*)k_sweep_fused<false, false, true>,
                              (const void*)k_sweep_fused<true, false, true>, (const void*)k_sweep_fused<true, true, true> };
        for (const void* f : fns)
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
                c->fused_ok = false;
        (void)hipGetLastError();
    }
    return dalloc(c, &c->d_acc_base, (size_t)c->R);
}

// Window of a replica under a requested slab height hy: whole planes (rows
// = hy = ly, one slab) or slabs of hy owned rows in windows of hy + 4 rows.
// hy = 0: whole planes where they fit; otherwise the fewest slabs whose
// windows fit, of equal height (the last may own fewer rows), so that no
// slab's window is mostly halo.
struct SlabShape { int rows, hy, nslab; };

SlabShape slab_shape(const Rep& r, int hy_req)
{
    const bool whole = r.plane <= kFusedMaxPairs && (hy_req == 0 || hy_req >= r.ly);
    if (whole) return { r.ly, r.ly, 1 };
    const int hmax = kFusedMaxPairs / r.hx - kSlabHalo;     // rows the tallest window owns
    int hy = std::min(hy_req > 0 ? hy_req : hmax, hmax);
    if (hy_req == 0 && hy >= 2) hy = (r.ly + (r.ly + hy - 1) / hy - 1) / ((r.ly + hy - 1) / hy);
    if (hy < 2 || hy >= r.ly) return { 0, 0, 0 };
    return { hy + kSlabHalo, hy, (r.ly + hy - 1) / hy };
}

// Chunks of lc z planes (and y slabs of hy rows) per fused workgroup (lc =
// 0: none). Acceptance counters of the previous layout are folded into the
// per-slot base first.
int fused_layout(ising_ctx* c, int lc, int hy_req)
{
    std::vector<SlabShape> shape(c->reps.size());
    int pw_max = 0, fb = 0;
    bool any_slab = false;
    for (size_t k = 0; k < c->reps.size(); ++k) {
        const Rep& r = c->reps[k];
        shape[k] = lc > 0 ? slab_shape(r, hy_req) : SlabShape{ r.ly, r.ly, 1 };
        if (shape[k].nslab == 0)
            return set_err(c, ISING_ERR_UNSUPPORTED, "fused sweep: no y slab of %d rows fits a workgroup on a "
                           "%d x %d plane", hy_req, r.lx, r.ly);
        pw_max = std::max(pw_max, shape[k].rows * r.hx);
        any_slab = any_slab || shape[k].nslab > 1;
        if (lc > 0) fb += (r.lz + lc - 1) / lc * shape[k].nslab;
    }
    const int threads = std::max(kWave, ((pw_max + kFusedPasses - 1) / kFusedPasses + kWave - 1) / kWave * kWave);
    const int wpb = threads / kWave;
    // Larger arrays for the block table and the per-wavefront counters and
    // sums (never shrunk) are allocated before anything changes, so a
    // failure leaves the previous layout in place.
    int rc;
    int32_t* nrep = nullptr;
    unsigned long long* nacc = nullptr;
    double* npart = nullptr;
    const bool grow_b = (size_t)fb > c->fblock_cap, grow_w = (size_t)fb * wpb > c->fwave_cap;
    if ((grow_b && (rc = dalloc(c, &nrep, (size_t)fb))) ||
        (grow_w && ((rc = dalloc(c, &nacc, (size_t)fb * wpb)) || (rc = dalloc(c, &npart, 4 * (size_t)fb * wpb))))) {
        if (nrep) (void)hipFree(nrep);
        if (nacc) (void)hipFree(nacc);
        return rc;
    }
    // the old layout's acceptance counts move to the per-slot base
    hipError_t e = hipSuccess;
    if (c->n_fblocks > 0) {
        hipLaunchKernelGGL(k_acc_fold, dim3(c->R), dim3(kBlock), 0, c->stream, c->d_reps,
                           (const unsigned long long*)c->d_acc_wave_f, c->fused_wpb, c->d_acc_base);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {            // the previous layout stays
        if (nrep) (void)hipFree(nrep);
        if (nacc) (void)hipFree(nacc);
        if (npart) (void)hipFree(npart);
        return set_err(c, ISING_ERR_HIP, "fused layout: %s", hipGetErrorString(e));
    }
    if (grow_b) {
        if (c->d_fblock_rep) (void)hipFree(c->d_fblock_rep);
        c->d_fblock_rep = nrep;
        c->fblock_cap = (size_t)fb;
    }
    if (grow_w) {
        if (c->d_acc_wave_f) (void)hipFree(c->d_acc_wave_f);
        if (c->d_wave_part_f) (void)hipFree(c->d_wave_part_f);
        c->d_acc_wave_f = nacc;
        c->d_wave_part_f = npart;
        c->fwave_cap = (size_t)fb * wpb;
    }
    c->fblock_rep.clear();
    fb = 0;
    for (size_t k = 0; k < c->reps.size(); ++k) {
        Rep& r = c->reps[k];
        r.f_rows = shape[k].rows;
        r.f_hy = shape[k].hy;
        r.f_nslab = shape[k].nslab;
        r.f_first_block = fb;
        r.f_nblocks = lc > 0 ? (r.lz + lc - 1) / lc * r.f_nslab : 0;
        fb += r.f_nblocks;
        for (int b = 0; b < r.f_nblocks; ++b) c->fblock_rep.push_back((int32_t)k);
    }
    c->n_fblocks = fb;
    c->fused_threads = threads;
    c->fused_wpb = wpb;
    c->fused_pstride = (uint32_t)pw_max + 1u;
    c->fused_lds = (size_t)4 * (pw_max + 1) * sizeof(float4);
    c->fused_slab = any_slab;
    if (c->fwave_cap) e = hipMemsetAsync(c->d_acc_wave_f, 0, c->fwave_cap * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_reps, c->reps.data(), c->reps.size() * sizeof(Rep), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && fb > 0)
        e = hipMemcpyAsync(c->d_fblock_rep, c->fblock_rep.data(), (size_t)fb * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        // neither layout is on the device for sure: leave the fused sweep off
        // (half sweeps need none of it) until a later ising_set_tuning
        for (Rep& r : c->reps) r.f_nblocks = 0;
        c->n_fblocks = 0;
        c->tune &= ~(kTuneFusedMask | kTuneSlabMask);
        return set_err(c, ISING_ERR_HIP, "fused layout: %s", hipGetErrorString(e));
    }
    return ISING_OK;
}

int64_t gcd64(int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a; }

int64_t schedule_period(const ising_schedule& s)
{
    int64_t period = 1;
    for (int v : { s.swap_every, s.measure_every, s.winding_every })
        if (v > 0) period = period / gcd64(period, v) * v;
    return period;
}

// ---- statically allocated series staging (ising_pt_run_static) ------------
// For callers that cannot allocate device memory: a file-scope device array,
// one copy per GPU, each at its own address. That address is taken once per
// device, with the device current, when the first context on it is created,
// and host code uses only that address (the array's name in host code is a
// host-side stand-in, not a device pointer). Every call on a GPU shares its
// copy, so calls that overlap on one GPU race.
constexpr size_t kStaticScratchBytes = (size_t)ISING_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int static_scratch_setup(ising_ctx* c)
{
    if (c->device >= kMaxDevices) return ISING_OK;      // ising_pt_run_static refuses such devices
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return ISING_OK;
    void* p = nullptr;
    HIP_TRY(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(c, ISING_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[c->device] = p;
    return ISING_OK;
}

// Byte layout of one run's series in a staging area: the descriptor, then
// energies, magnetisations, FP64 and integer charges (each 256-byte aligned).
struct SeriesLayout {
    size_t e, m, qr, qi, total;
};

SeriesLayout series_layout(size_t n_meas, size_t n_q, size_t R, size_t Z)
{
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    SeriesLayout L;
    L.e = up(sizeof(SeriesDesc));
    L.m = L.e + up(n_meas * R * sizeof(double));
    L.qr = L.m + up(n_meas * R * 3 * sizeof(double));
    L.qi = L.qr + up(n_q * R * Z * sizeof(double));
    L.total = L.qi + up(n_q * R * Z * sizeof(long long));
    return L;
}

// Series buffers of the context, grown to hold n_meas E/M and n_q charge
// records of Z entries per temperature (they are never shrunk; a reallocation needs no new graph, the
// captured launches read the buffers through the descriptor).
int series_reserve(ising_ctx* c, size_t n_meas, size_t n_q, size_t Z)
{
    const size_t R = (size_t)c->R;
    int rc;
    if (n_meas * R > c->ser_cap_e) {
        if (c->ser_e) (void)hipFree(c->ser_e);
        if (c->ser_m) (void)hipFree(c->ser_m);
        c->ser_e = nullptr; c->ser_m = nullptr; c->ser_cap_e = 0;
        if ((rc = dalloc(c, &c->ser_e, n_meas * R)) || (rc = dalloc(c, &c->ser_m, 3 * n_meas * R))) return rc;
        c->ser_cap_e = n_meas * R;
    }
    if (n_q * R * Z > c->ser_cap_q) {
        if (c->ser_qr) (void)hipFree(c->ser_qr);
        if (c->ser_qi) (void)hipFree(c->ser_qi);
        c->ser_qr = nullptr; c->ser_qi = nullptr; c->ser_cap_q = 0;
        if ((rc = dalloc(c, &c->ser_qr, n_q * R * Z)) || (rc = dalloc(c, &c->ser_qi, n_q * R * Z))) return rc;
        c->ser_cap_q = n_q * R * Z;
    }
    return ISING_OK;
}

void drop_pt_graph(ising_ctx* c)
{
    if (c->pt_exec) (void)hipGraphExecDestroy(c->pt_exec);
    if (c->pt_graph) (void)hipGraphDestroy(c->pt_graph);
    c->pt_exec = nullptr;
    c->pt_graph = nullptr;
    c->pt_key = PtGraphKey{};
}

// The body of the PT entry points. staging == nullptr: the series live in
// the context's buffers and whole periods replay one cached graph;
// otherwise the descriptor and the series sit in `staging` (device memory of
// at least series_layout(...).total bytes) and every period is launched
// eagerly, so nothing is allocated.
int pt_run_impl(ising_ctx* c, const ising_schedule* sched, const ising_series* out, unsigned char* staging)
{
    if (!c || !sched) return ISING_ERR_INVALID_ARG;
    if (c->any_demag)
        return set_err(c, ISING_ERR_UNSUPPORTED, "parallel tempering with the demagnetisation term (the reference's "
                       "PT runs without it)");
    const ising_schedule s = *sched;
    if (s.n_sweeps < 0 || s.swap_every < 0 || s.measure_every < 0 || s.winding_every < 0)
        return set_err(c, ISING_ERR_INVALID_ARG, "negative schedule entry");
    if (s.winding_layer < 0 || s.winding_layer > c->zmax)
        return set_err(c, ISING_ERR_INVALID_ARG, "winding_layer %d outside 0..%d", s.winding_layer, c->zmax);
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    const size_t n_meas = s.measure_every ? (size_t)(s.n_sweeps / s.measure_every) : 0u;
    const size_t n_q = s.winding_every ? (size_t)(s.n_sweeps / s.winding_every) : 0u;
    const size_t R = (size_t)c->R;
    // one layer named: its charges alone, [n_q][R]; else [n_q][R][Zmax]
    const size_t Z = s.winding_layer > 0 ? 1u : (size_t)c->zmax;
    SeriesDesc desc{};
    SeriesDesc* d_desc = nullptr;
    if (staging) {
        const SeriesLayout lay = series_layout(n_meas, n_q, R, Z);
        d_desc = (SeriesDesc*)staging;
        desc.e = (double*)(staging + lay.e);
        desc.m = (double*)(staging + lay.m);
        desc.qr = (double*)(staging + lay.qr);
        desc.qi = (long long*)(staging + lay.qi);
    } else {
        if ((rc = series_reserve(c, n_meas, n_q, Z))) return rc;
        d_desc = c->d_series;
        desc.e = c->ser_e; desc.m = c->ser_m; desc.qr = c->ser_qr; desc.qi = c->ser_qi;
    }
    desc.cap_e = (uint32_t)n_meas;
    desc.cap_q = (uint32_t)n_q;
    desc.zrec = (uint32_t)Z;
    SeriesDev sd;
    sd.desc = d_desc;
    // Requested series of up to kStageMax bytes go to a pinned staging block
    // of the context by copies queued behind the last period, so they cost
    // one synchronisation and a host copy instead of a blocking pageable copy
    // each; larger ones, whose time is bandwidth, and the allocation-free
    // entry point copy directly after the run. The block is (re)allocated
    // before any work is queued, so a failure leaves the context unchanged.
    struct Part { void* dst; const void* src; size_t bytes; };
    Part parts[4] = {};
    int n_parts = 0;
    if (out) {
        if (out->energy && n_meas) parts[n_parts++] = { out->energy, desc.e, n_meas * R * sizeof(double) };
        if (out->magnetisation && n_meas) parts[n_parts++] = { out->magnetisation, desc.m, n_meas * R * 3 * sizeof(double) };
        if (out->winding && n_q) parts[n_parts++] = { out->winding, desc.qi, n_q * R * Z * sizeof(int64_t) };
        if (out->winding_raw && n_q) parts[n_parts++] = { out->winding_raw, desc.qr, n_q * R * Z * sizeof(double) };
    }
    size_t stage_bytes = 0;
    for (int k = 0; k < n_parts; ++k) stage_bytes += (parts[k].bytes + 255u) & ~(size_t)255u;
    constexpr size_t kStageMax = (size_t)64 << 20;
    const bool staged = !staging && n_parts > 0 && stage_bytes <= kStageMax;
    if (staged && stage_bytes > c->h_stage_cap) {
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        c->h_stage_cap = 0;
        if (hipHostMalloc((void**)&c->h_stage, stage_bytes, hipHostMallocDefault) != hipSuccess) {
            c->h_stage = nullptr;
            return set_err(c, ISING_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu bytes) for the series copy-out", stage_bytes);
        }
        c->h_stage_cap = stage_bytes;
    }
    // The descriptor and the record counters of this run (plain kernels: no
    // staging copy, no memset).
    hipLaunchKernelGGL(k_set_series_desc, dim3(1), dim3(64), 0, c->stream, d_desc, desc);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, c->d_ctr + 1, 2, 0u);
    LAUNCH_CHECK(c);

    const int64_t period = schedule_period(s);
    const bool graph = !staging && s.use_graph && period <= 512 && s.n_sweeps >= 2 * period;
    int done = 0;
    if (graph) {
        PtGraphKey key;
        key.swap_every = s.swap_every; key.measure_every = s.measure_every; key.winding_every = s.winding_every;
        key.winding_layer = s.winding_layer; key.period = period; key.tune = c->tune;
        key.learned = c->any_learned ? 1 : 0; key.demag = c->any_demag ? 1 : 0; key.records = 1;
        key.demag_buf = c->d_demag;
        if (!c->pt_exec || !(key == c->pt_key)) {
            drop_pt_graph(c);
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return set_err(c, ISING_ERR_HIP, "hipStreamBeginCapture");
            rc = enqueue_period(c, s, (int)period, sd);
            hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (rc || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : set_err(c, ISING_ERR_HIP, "capture");
            }
            hipGraphExec_t ge = nullptr;
            if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(g);
                return set_err(c, ISING_ERR_HIP, "hipGraphInstantiate");
            }
            c->pt_graph = g;
            c->pt_exec = ge;
            c->pt_key = key;
        }
        const int reps = s.n_sweeps / (int)period;
        for (int k = 0; k < reps; ++k)
            if (hipGraphLaunch(c->pt_exec, c->stream) != hipSuccess) return set_err(c, ISING_ERR_HIP, "hipGraphLaunch");
        done = reps * (int)period;
    }
    // Remaining sweeps eagerly, period by period (records keep their order).
    while (done < s.n_sweeps) {
        const int cnt = (int)std::min<int64_t>(period, s.n_sweeps - done);
        if ((rc = enqueue_period(c, s, cnt, sd))) return rc;
        done += cnt;
    }
    hipError_t ecopy = hipSuccess;
    if (staged) {
        size_t off = 0;
        for (int k = 0; k < n_parts && ecopy == hipSuccess; ++k) {
            ecopy = hipMemcpyAsync(c->h_stage + off, parts[k].src, parts[k].bytes, hipMemcpyDeviceToHost, c->stream);
            off += (parts[k].bytes + 255u) & ~(size_t)255u;
        }
    }
    hipError_t es = hipStreamSynchronize(c->stream);
    if (es != hipSuccess) return set_err(c, ISING_ERR_HIP, "pt_run: %s", hipGetErrorString(es));
    advance_host(c, (uint32_t)s.n_sweeps);
    if (ecopy != hipSuccess) return set_err(c, ISING_ERR_HIP, "series copy-out: %s", hipGetErrorString(ecopy));
    if (staged) {
        size_t off = 0;
        for (int k = 0; k < n_parts; ++k) {
            memcpy(parts[k].dst, c->h_stage + off, parts[k].bytes);
            off += (parts[k].bytes + 255u) & ~(size_t)255u;
        }
        return ISING_OK;
    }
    if (out) {
        hipError_t e1 = hipSuccess;
        if (out->energy && n_meas)
            e1 = hipMemcpy(out->energy, desc.e, n_meas * R * sizeof(double), hipMemcpyDeviceToHost);
        if (e1 == hipSuccess && out->magnetisation && n_meas)
            e1 = hipMemcpy(out->magnetisation, desc.m, n_meas * R * 3 * sizeof(double), hipMemcpyDeviceToHost);
        if (e1 == hipSuccess && out->winding && n_q)
            e1 = hipMemcpy(out->winding, desc.qi, n_q * R * Z * sizeof(int64_t), hipMemcpyDeviceToHost);
        if (e1 == hipSuccess && out->winding_raw && n_q)
            e1 = hipMemcpy(out->winding_raw, desc.qr, n_q * R * Z * sizeof(double), hipMemcpyDeviceToHost);
        if (e1 != hipSuccess) return set_err(c, ISING_ERR_HIP, "series copy-out: %s", hipGetErrorString(e1));
    }
    return ISING_OK;
}

} // namespace

// ===========================================================================
extern "C" {

const char* ising_version(void) { return "ising_hip 0.1 (gfx950)"; }

int ising_device_count(int* n)
{
    if (!n) return ISING_ERR_INVALID_ARG;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    *n = k;
    return ISING_OK;
}

const char* ising_last_error(const ising_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int ising_num_replicas(const ising_ctx* ctx, int* n)
{
    if (!ctx || !n) return ISING_ERR_INVALID_ARG;
    *n = ctx->R;
    return ISING_OK;
}

int ising_replica_sites(const ising_ctx* ctx, int replica, int64_t* n)
{
    if (!ctx || !n || replica < 0 || replica >= ctx->R) return ISING_ERR_INVALID_ARG;
    *n = ctx->reps[replica].n;
    return ISING_OK;
}

void ising_destroy(ising_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = { c->d_spins, c->d_reps, c->d_lads, c->d_lad_override, c->d_block_rep, c->d_qblock_rep, c->d_lblock_rep,
                     c->d_beta_slot, c->d_beta_temp, c->d_slot_of_temp, c->d_lad_of_pos, c->d_lad_start,
                     c->d_lad_n, c->d_partial, c->d_energy, c->d_mag, c->d_qpartial, c->d_qraw, c->d_qint,
                     c->d_ctr, c->d_acc, c->d_acc_wave, c->d_wave_part, c->d_mlp, c->d_mlp_prm, c->d_mlp_m1,
                     c->d_mlp_m2, c->d_mlp_grad, c->d_wave_grad, c->d_grad_part, c->d_recs, c->d_swap_att, c->d_swap_acc, c->d_sites, c->d_props, c->d_de,
                     c->ck_spins, c->ck_sot, c->ck_bslot, c->ck_btemp, c->ck_ctr,
                     c->d_fblock_rep, c->d_acc_wave_f, c->d_acc_base, c->d_wave_part_f,
                     c->d_demag, c->d_dmesh, c->d_dk1, c->d_dk2, c->d_tw, c->d_dreps, c->d_cic,
                     c->ser_e, c->ser_m, c->ser_qr, c->ser_qi, c->d_series };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->pt_exec) (void)hipGraphExecDestroy(c->pt_exec);
    if (c->pt_graph) (void)hipGraphDestroy(c->pt_graph);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ising_create(ising_ctx** out, const ising_ladder_desc* ladders, int n_ladders, uint64_t seed, int device)
{
    if (!out) return ISING_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ladders || n_ladders < 1) return ISING_ERR_INVALID_ARG;
    ising_ctx* c = new ising_ctx();
    c->seed = seed;
    c->device = device;
    int rc = ISING_OK;
    // On failure the partial context is released and its message is kept for
    // ising_last_error(NULL) on this thread.
    auto fail = [&](int code) {
        g_create_err = c->err;
        ising_destroy(c);
        *out = nullptr;
        return code;
    };

    // Validate and lay out replicas.
    int R = 0;
    for (int l = 0; l < n_ladders; ++l) {
        const ising_ladder_desc& d = ladders[l];
        if ((rc = check_lattice(c, d.lattice))) return fail(rc);
        if ((rc = check_ham(c, d.ham))) return fail(rc);
        if (d.ham.demag && (rc = check_demag_lattice(c, d.lattice))) return fail(rc);
        if (d.n_replicas < 1 || !d.betas) return fail(set_err(c, ISING_ERR_INVALID_ARG, "ladder %d: n_replicas/betas", l));
        if (!(d.cone_radius >= 0.0) || !std::isfinite(d.cone_radius))
            return fail(set_err(c, ISING_ERR_INVALID_ARG, "ladder %d: cone radius", l));
        LadHost L;
        L.lat = d.lattice; L.ham = d.ham; L.n = d.n_replicas; L.start = R; L.cone = d.cone_radius;
        L.betas.assign(d.betas, d.betas + d.n_replicas);
        for (double b : L.betas)
            if (!std::isfinite(b) || b < 0) return fail(set_err(c, ISING_ERR_INVALID_ARG, "ladder %d: bad beta", l));
        c->ladders.push_back(L);
        R += d.n_replicas;
    }
    c->R = R;
    int64_t off = 0;
    int blk = 0, qblk = 0, zmax = 0;
    for (int l = 0; l < n_ladders; ++l) {
        const LadHost& L = c->ladders[l];
        for (int k = 0; k < L.n; ++k) {
            Rep r;
            memset(&r, 0, sizeof r);
            r.lx = L.lat.lx; r.ly = L.lat.ly; r.lz = L.lat.lz;
            r.cell = L.lat.cell;
            r.hx = r.cell == ISING_CELL_SC ? r.lx / 2 : r.lx;
            r.per = (L.lat.periodic[0] ? 1 : 0) | (L.lat.periodic[1] ? 2 : 0) | (L.lat.periodic[2] ? 4 : 0);
            r.ladder = l;
            r.n = (int64_t)r.lx * r.ly * r.lz * atoms_of(r.cell);
            r.ncol = r.cell == ISING_CELL_HYPERKAGOME ? kHkAtoms / kHkPerColour : 2;
            r.half = (uint32_t)(r.n / r.ncol);
            c->max_ncol = std::max(c->max_ncol, r.ncol);
            r.off = off;
            // Replica blocks start on 256-byte boundaries, an odd number of
            // 256-byte granules apart, and 64 KB further apart for blocks of 1
            // MB and more: block starts a power-of-two multiple apart send the
            // same site of every replica to the same HBM channels, which cost
            // the fused sweep 6% at 64^3 x 64 (3 MB blocks).
            int64_t stride = (3 * r.n + 63) & ~(int64_t)63;
            if (((stride >> 6) & 1) == 0) stride += 64;
            if (stride >= (int64_t)1 << 18) stride += 16384;
            off += stride;
            r.first_block = blk;
            r.nblocks = (int)((r.half + kBlock - 1) / kBlock);
            blk += r.nblocks;
            r.q_first_block = qblk;
            r.q_blocks_layer = (r.lx * r.ly + kBlock - 1) / kBlock;
            qblk += r.q_blocks_layer * r.lz;
            r.l_first_block = (int32_t)c->lblock_rep.size();
            r.plane = r.ly * r.hx;
            r.div_hx = make_fastdiv((uint32_t)r.hx);
            r.div_ly = make_fastdiv((uint32_t)r.ly);
            r.div_lx = make_fastdiv((uint32_t)r.lx);
            r.div_lz = make_fastdiv((uint32_t)r.lz);
            zmax = std::max(zmax, r.lz);
            c->reps.push_back(r);
            for (int b = 0; b < r.nblocks; ++b) c->block_rep.push_back((int32_t)c->reps.size() - 1);
            for (int b = 0; b < r.q_blocks_layer * r.lz; ++b) c->qblock_rep.push_back((int32_t)c->reps.size() - 1);
            for (int b = 0; b < r.q_blocks_layer; ++b) c->lblock_rep.push_back((int32_t)c->reps.size() - 1);
        }
    }
    for (auto& r : c->reps) r.zbase = 0;
    c->zmax = zmax;
    c->n_blocks = blk;
    c->n_qblocks = qblk;
    c->n_lblocks = (int)c->lblock_rep.size();
    c->total_floats = off;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(set_err(c, ISING_ERR_NO_DEVICE, "no HIP device visible"));
    if (device < 0 || device >= ndev)
        return fail(set_err(c, ISING_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, ndev));
    if (hipSetDevice(device) != hipSuccess) return fail(set_err(c, ISING_ERR_NO_DEVICE, "hipSetDevice(%d)", device));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(c, ISING_ERR_HIP, "hipStreamCreate"));
    if (c->max_ncol > 2) {      // hyper-kagome ladders: the cell table in this device's constant memory
        const HkTable* T = hk_table();
        if (!T) return fail(set_err(c, ISING_ERR_STATE, "hyper-kagome table construction failed"));
        if (hipMemcpyToSymbol(HIP_SYMBOL(c_hk), T, sizeof(HkTable), 0, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(c, ISING_ERR_HIP, "upload of the hyper-kagome table"));
    }

    if ((rc = dalloc(c, &c->d_spins, 2 * (size_t)c->total_floats))) return fail(rc);
    if ((rc = dalloc(c, &c->d_reps, c->reps.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lads, c->ladders.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lad_override, 1))) return fail(rc);
    if ((rc = dalloc(c, &c->d_block_rep, c->block_rep.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_qblock_rep, c->qblock_rep.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lblock_rep, c->lblock_rep.size()))) return fail(rc);
    if ((rc = dalloc(c, &c->d_beta_slot, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_beta_temp, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_slot_of_temp, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lad_of_pos, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lad_start, n_ladders))) return fail(rc);
    if ((rc = dalloc(c, &c->d_lad_n, n_ladders))) return fail(rc);
    if ((rc = dalloc(c, &c->d_partial, 4 * (size_t)blk))) return fail(rc);
    if ((rc = dalloc(c, &c->d_energy, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mag, 3 * (size_t)R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_qpartial, (size_t)qblk))) return fail(rc);
    if ((rc = dalloc(c, &c->d_qraw, (size_t)R * zmax))) return fail(rc);
    if ((rc = dalloc(c, &c->d_qint, (size_t)R * zmax))) return fail(rc);
    if ((rc = dalloc(c, &c->d_ctr, kCounters))) return fail(rc);
    if ((rc = dalloc(c, &c->d_acc, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_acc_wave, (size_t)blk * (kBlock / kWave)))) return fail(rc);
    if ((rc = dalloc(c, &c->d_wave_part, 4 * (size_t)blk * (kBlock / kWave)))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mlp, 1))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mlp_prm, kMlpParams))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mlp_m1, kMlpParams))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mlp_m2, kMlpParams))) return fail(rc);
    if ((rc = dalloc(c, &c->d_mlp_grad, kMlpParams + 1))) return fail(rc);
    if ((rc = dalloc(c, &c->d_grad_part, (size_t)kMlpReduceGroups * (kMlpParams + 1)))) return fail(rc);
    c->learned.assign(n_ladders, 0);
    if ((rc = dalloc(c, &c->d_swap_att, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_swap_acc, R))) return fail(rc);
    if ((rc = dalloc(c, &c->d_series, 1))) return fail(rc);
    if ((rc = fused_setup(c))) return fail(rc);
    if ((rc = static_scratch_setup(c))) return fail(rc);

    std::vector<float> beta_slot(R);
    std::vector<double> beta_temp(R);
    std::vector<int32_t> slot_of_temp(R), lad_of_pos(R), lad_start(n_ladders), lad_n(n_ladders);
    for (int l = 0; l < n_ladders; ++l) {
        const LadHost& L = c->ladders[l];
        lad_start[l] = L.start; lad_n[l] = L.n;
        for (int k = 0; k < L.n; ++k) {
            beta_slot[L.start + k] = (float)L.betas[k];
            beta_temp[L.start + k] = L.betas[k];
            slot_of_temp[L.start + k] = L.start + k;
            lad_of_pos[L.start + k] = l;
        }
    }
    hipStream_t s = c->stream;
#define UP(dst, src, cnt) \
    if (hipMemcpyAsync(dst, src, (cnt) * sizeof(*(src)), hipMemcpyHostToDevice, s) != hipSuccess) \
        return fail(set_err(c, ISING_ERR_HIP, "upload " #dst));
    UP(c->d_reps, c->reps.data(), c->reps.size());
    UP(c->d_block_rep, c->block_rep.data(), c->block_rep.size());
    UP(c->d_qblock_rep, c->qblock_rep.data(), c->qblock_rep.size());
    UP(c->d_lblock_rep, c->lblock_rep.data(), c->lblock_rep.size());
    UP(c->d_beta_slot, beta_slot.data(), R);
    UP(c->d_beta_temp, beta_temp.data(), R);
    UP(c->d_slot_of_temp, slot_of_temp.data(), R);
    UP(c->d_lad_of_pos, lad_of_pos.data(), R);
    UP(c->d_lad_start, lad_start.data(), n_ladders);
    UP(c->d_lad_n, lad_n.data(), n_ladders);
#undef UP
    if (hipMemsetAsync(c->d_spins, 0, 2 * (size_t)c->total_floats * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(c->d_acc_base, 0, R * sizeof(unsigned long long), s) != hipSuccess ||
        (c->fwave_cap && hipMemsetAsync(c->d_acc_wave_f, 0, c->fwave_cap * sizeof(unsigned long long), s) != hipSuccess) ||
        hipMemsetAsync(c->d_ctr, 0, kCounters * sizeof(uint32_t), s) != hipSuccess ||
        hipMemsetAsync(c->d_acc, 0, R * sizeof(unsigned long long), s) != hipSuccess ||
        hipMemsetAsync(c->d_acc_wave, 0, (size_t)blk * (kBlock / kWave) * sizeof(unsigned long long), s) != hipSuccess ||
        hipMemsetAsync(c->d_mlp_prm, 0, kMlpParams * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(c->d_mlp_m1, 0, kMlpParams * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(c->d_mlp_m2, 0, kMlpParams * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(c->d_swap_att, 0, R * sizeof(unsigned long long), s) != hipSuccess ||
        hipMemsetAsync(c->d_swap_acc, 0, R * sizeof(unsigned long long), s) != hipSuccess)
        return fail(set_err(c, ISING_ERR_HIP, "hipMemsetAsync"));
    if ((rc = upload_lads(c))) return fail(rc);
    if (c->any_demag && (rc = demag_setup(c))) return fail(rc);
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, s, (const float*)c->d_mlp_prm, c->d_mlp);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return fail(set_err(c, ISING_ERR_HIP, "network initialisation"));
    *out = c;
    return ISING_OK;
}

int ising_set_hamiltonian(ising_ctx* c, int ladder, const ising_hamiltonian* h)
{
    if (!c || !h) return ISING_ERR_INVALID_ARG;
    if (ladder < 0 || ladder >= (int)c->ladders.size()) return set_err(c, ISING_ERR_INVALID_ARG, "ladder %d", ladder);
    int rc = check_ham(c, *h);
    if (rc) return rc;
    if (h->demag) {
        if ((rc = check_demag_lattice(c, c->ladders[ladder].lat))) return rc;
        if (c->any_learned)
            return set_err(c, ISING_ERR_UNSUPPORTED, "a ladder of this context takes its radius from the network "
                           "(the learned sweep kernel carries no demagnetising field)");
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (h->demag && (rc = demag_setup(c))) return rc;
    c->ladders[ladder].ham = *h;
    return upload_lads(c);
}

int ising_set_cone(ising_ctx* c, int ladder, double radius)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (ladder < 0 || ladder >= (int)c->ladders.size()) return set_err(c, ISING_ERR_INVALID_ARG, "ladder %d", ladder);
    if (!(radius >= 0.0) || !std::isfinite(radius)) return set_err(c, ISING_ERR_INVALID_ARG, "cone radius");
    HIP_TRY(c, hipSetDevice(c->device));
    c->ladders[ladder].cone = radius;
    return upload_lads(c);
}

int ising_set_beta(ising_ctx* c, int replica, double beta)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (!std::isfinite(beta) || beta < 0) return set_err(c, ISING_ERR_INVALID_ARG, "beta");
    HIP_TRY(c, hipSetDevice(c->device));
    // The slot keeps its temperature position: set both the slot's beta and
    // the beta of the temperature position it currently occupies.
    std::vector<int32_t> sot(c->R);
    HIP_TRY(c, hipMemcpyAsync(sot.data(), c->d_slot_of_temp, c->R * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int t = -1;
    for (int k = 0; k < c->R; ++k) if (sot[k] == replica) t = k;
    float bf = (float)beta;
    HIP_TRY(c, hipMemcpyAsync(c->d_beta_slot + replica, &bf, sizeof bf, hipMemcpyHostToDevice, c->stream));
    if (t >= 0) HIP_TRY(c, hipMemcpyAsync(c->d_beta_temp + t, &beta, sizeof beta, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& L : c->ladders)
        if (t >= L.start && t < L.start + L.n) L.betas[t - L.start] = beta;
    return ISING_OK;
}

int ising_get_betas(ising_ctx* c, double* out)
{
    if (!c || !out) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<float> b(c->R);
    HIP_TRY(c, hipMemcpyAsync(b.data(), c->d_beta_slot, c->R * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->R; ++k) out[k] = b[k];
    return ISING_OK;
}

int ising_set_spins(ising_ctx* c, int replica, const double* soa, int64_t n)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    const Rep& r = c->reps[replica];
    if (!soa || n != r.n) return set_err(c, ISING_ERR_INVALID_ARG, "spin buffer of %lld sites, replica has %lld",
                                         (long long)n, (long long)r.n);
    std::vector<float> h((size_t)3 * r.n);
    for (int64_t site = 0; site < n; ++site) {
        const double x = soa[site], y = soa[n + site], z = soa[2 * n + site];
        const double nn = sqrt(x * x + y * y + z * z);
        if (!std::isfinite(nn) || fabs(nn - 1.0) > 1e-3)
            return set_err(c, ISING_ERR_BAD_SPIN, "site %lld: |S| = %g (NaN, inf or not unit)", (long long)site, nn);
        const int64_t idx = dev_index(r, site);
        h[idx] = (float)x; h[r.n + idx] = (float)y; h[2 * r.n + idx] = (float)z;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(cur_spins(c) + r.off, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_get_spins(ising_ctx* c, int replica, double* soa, int64_t n)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    const Rep& r = c->reps[replica];
    if (!soa || n != r.n) return set_err(c, ISING_ERR_INVALID_ARG, "spin buffer size");
    std::vector<float> h((size_t)3 * r.n);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(h.data(), cur_spins(c) + r.off, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int64_t site = 0; site < n; ++site) {
        const int64_t idx = dev_index(r, site);
        soa[site] = h[idx]; soa[n + site] = h[r.n + idx]; soa[2 * n + site] = h[2 * r.n + idx];
    }
    return ISING_OK;
}

static int init_common(ising_ctx* c, int replica, int mode, double frac)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (replica >= c->R) return set_err(c, ISING_ERR_INVALID_ARG, "replica %d", replica);
    HIP_TRY(c, hipSetDevice(c->device));
    int base = 0, nb = c->n_blocks;
    if (replica >= 0) { base = c->reps[replica].first_block; nb = c->reps[replica].nblocks; }
    hipLaunchKernelGGL(k_init, dim3(nb), dim3(kBlock), 0, c->stream, cur_spins(c), c->d_reps, c->d_block_rep, base,
                       mode, frac, (uint32_t)c->seed, (uint32_t)(c->seed >> 32));
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_init_random(ising_ctx* c, int replica) { return init_common(c, replica, 0, 0.0); }

int ising_init_ferro(ising_ctx* c, int replica, double frac)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (!(frac >= 0.0 && frac <= 1.0)) return set_err(c, ISING_ERR_INVALID_ARG, "flip fraction %g", frac);
    return init_common(c, replica, 1, frac);
}

int ising_sweep(ising_ctx* c, int n_sweeps)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (n_sweeps < 0) return set_err(c, ISING_ERR_INVALID_ARG, "n_sweeps %d", n_sweeps);
    if (n_sweeps == 0) return ISING_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    // Metropolis_mcs computes the field before the first sweep and after
    // every demag_update_interval-th sweep (src/energyMin.f90:258-259,362-369).
    const bool demag = c->any_demag;
    if (demag && (rc = launch_demag(c, kDemagLadders))) return rc;
    for (int k = 0; k < n_sweeps; ++k) {
        if ((rc = launch_sweep(c, (uint32_t)k))) return rc;
        if (demag && c->demag_interval > 0 && (k + 1) % c->demag_interval == 0 &&
            (rc = launch_demag(c, kDemagLadders)))
            return rc;
    }
    if ((rc = launch_advance(c, (uint32_t)n_sweeps))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    advance_host(c, (uint32_t)n_sweeps);
    return ISING_OK;
}

int ising_energy(ising_ctx* c, double* e)
{
    if (!c || !e) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    if (c->any_demag && (rc = launch_demag(c, kDemagLadders))) return rc;   // TotalHeisenbergEnergy recomputes it
    rc = launch_measure_all(c);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(e, c->d_energy, c->R * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_magnetisation(ising_ctx* c, double* m)
{
    if (!c || !m) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = launch_measure_all(c);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(m, c->d_mag, 3 * c->R * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_energy_with(ising_ctx* c, int replica, const ising_hamiltonian* h, double* e)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (!h || !e) return set_err(c, ISING_ERR_INVALID_ARG, "null argument");
    if ((rc = check_ham(c, *h))) return rc;
    if (h->demag && (rc = check_demag_lattice(c, c->ladders[c->reps[replica].ladder].lat))) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if (h->demag && (rc = launch_demag(c, replica))) return rc;
    const Rep& r = c->reps[replica];
    Lad L = make_lad(*h, c->ladders[r.ladder].cone, r.cell);
    HIP_TRY(c, hipMemcpyAsync(c->d_lad_override, &L, sizeof L, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_measure, dim3(r.nblocks), dim3(kBlock), 0, c->stream, spin_buf(c), 0u, c->d_reps,
                       c->d_lads, (const Lad*)c->d_lad_override, c->d_block_rep, r.first_block, c->d_partial,
                       (const double*)c->d_demag);
    LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_measure_finalize, dim3(1), dim3(kBlock), 0, c->stream, c->d_reps, c->d_partial, replica,
                       c->d_energy, c->d_mag, (const int32_t*)nullptr, 0, kNoRecord);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(e, c->d_energy + replica, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_winding_all(ising_ctx* c, int64_t* q, double* q_raw)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = launch_winding_all(c);
    if (rc) return rc;
    const size_t cnt = (size_t)c->R * c->zmax;
    if (q) HIP_TRY(c, hipMemcpyAsync(q, c->d_qint, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    if (q_raw) HIP_TRY(c, hipMemcpyAsync(q_raw, c->d_qraw, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_winding_density(ising_ctx* c, int replica, int z, double* density)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (!density) return set_err(c, ISING_ERR_INVALID_ARG, "null density");
    const Rep& r = c->reps[replica];
    if (z < 0 || z >= r.lz) return set_err(c, ISING_ERR_INVALID_ARG, "layer %d out of range", z);
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = (size_t)r.lx * r.ly;
    double* d = nullptr;
    if ((rc = dalloc(c, &d, n))) return rc;
    hipLaunchKernelGGL(k_winding_density, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                       (const float*)cur_spins(c), c->d_reps, replica, z, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(density, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(c, ISING_ERR_HIP, "winding density: %s", hipGetErrorString(e));
    return ISING_OK;
}

// ---- skyrmion count ----------------------------------------------------------

// D_N(u): the band-limited interpolation kernel of N periodic samples, the
// Nyquist term of an even N split between +-N/2, in long double.
static double sk_interp_kernel(int n, long double u)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    long double acc = 1.0L;
    const int kmax = (n % 2 == 0) ? n / 2 - 1 : (n - 1) / 2;
    for (int k = 1; k <= kmax; ++k) acc += 2.0L * cosl(2.0L * pi * (long double)k * u / (long double)n);
    if (n % 2 == 0) acc += cosl(pi * u);
    return (double)(acc / (long double)n);
}

// K[r][m] = D_N(r/s + m): the weight of the sample m cells before the
// fine point r/s past a sample.
static std::vector<double> sk_kernel_table(int n, int s)
{
    std::vector<double> K((size_t)s * n);
    for (int r = 0; r < s; ++r)
        for (int m = 0; m < n; ++m) K[(size_t)r * n + m] = sk_interp_kernel(n, (long double)r / (long double)s + m);
    return K;
}

// Normalised 5x5 Gaussian weights G[a+2][b+2] (Gaussian_filter_2d,
// src/reciprocalSpaceProcesses.f90:814-857, centred).
static void sk_gauss_weights(double sigma, double (&G)[25])
{
    double acc = 0.0;
    for (int a = -2; a <= 2; ++a)
        for (int b = -2; b <= 2; ++b) {
            const double w = std::exp(-(double)(a * a + b * b) / (2.0 * sigma * sigma));
            G[(a + 2) * 5 + (b + 2)] = w;
            acc += w;
        }
    for (double& g : G) g /= acc;
}

// Device buffers of one layer's statistic, released by the destructor.
struct SkBufs {
    double *S = nullptr, *T = nullptr, *F = nullptr, *d = nullptr, *f = nullptr, *Kx = nullptr, *Ky = nullptr,
           *G = nullptr;
    int32_t *root = nullptr, *counts = nullptr;
    uint8_t* lvl = nullptr;
    unsigned long long *acc = nullptr, *mx = nullptr;
    ~SkBufs()
    {
        void* ptrs[] = { S, T, F, d, f, Kx, Ky, G, root, counts, lvl, acc, mx };
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
};

static dim3 grid_of(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// The fine density of layer z of `replica` into b.d (NX*NY doubles).
static int sk_fine_density(ising_ctx* c, int replica, int z, int sx, int sy, SkBufs& b)
{
    const Rep& r = c->reps[replica];
    const int nx = r.lx, ny = r.ly, NX = nx * sx, NY = ny * sy;
    const int64_t n = (int64_t)NX * NY;
    int rc;
    if ((rc = dalloc(c, &b.S, 3 * (size_t)nx * ny)) || (rc = dalloc(c, &b.T, 3 * (size_t)ny * NX)) ||
        (rc = dalloc(c, &b.F, 3 * (size_t)n)) || (rc = dalloc(c, &b.d, (size_t)n)) ||
        (rc = dalloc(c, &b.Kx, (size_t)sx * nx)) || (rc = dalloc(c, &b.Ky, (size_t)sy * ny)))
        return rc;
    const std::vector<double> kx = sk_kernel_table(nx, sx), ky = sk_kernel_table(ny, sy);
    HIP_TRY(c, hipMemcpyAsync(b.Kx, kx.data(), kx.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(b.Ky, ky.data(), ky.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_sk_coarse, grid_of((int64_t)nx * ny), dim3(kBlock), 0, c->stream, (const float*)cur_spins(c),
                       c->d_reps, replica, z, b.S);
    hipLaunchKernelGGL(k_sk_interp_x, grid_of((int64_t)3 * ny * NX), dim3(kBlock), 0, c->stream, (const double*)b.S,
                       nx, ny, sx, (const double*)b.Kx, b.T);
    hipLaunchKernelGGL(k_sk_interp_y, grid_of(3 * n), dim3(kBlock), 0, c->stream, (const double*)b.T, ny, NX, sy,
                       (const double*)b.Ky, b.F);
    hipLaunchKernelGGL(k_sk_density, grid_of(n), dim3(kBlock), 0, c->stream, (const double*)b.F, NX, NY, b.d);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));     // the host tables may go now
    return ISING_OK;
}

static int sk_check(ising_ctx* c, int replica, int z, int sx, int sy)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    const Rep& r = c->reps[replica];
    if (z < 0 || z >= r.lz) return set_err(c, ISING_ERR_INVALID_ARG, "layer %d out of range", z);
    if ((r.per & 3) != 3)
        return set_err(c, ISING_ERR_UNSUPPORTED, "the interpolated charge density needs periodic x and y");
    if (sx < 1 || sy < 1 || (int64_t)r.lx * sx * r.ly * sy > ((int64_t)1 << 26))
        return set_err(c, ISING_ERR_INVALID_ARG, "interpolation scales %d x %d (fine grid up to 2^26 points)", sx, sy);
    return ISING_OK;
}

int ising_winding_density_interpolated(ising_ctx* c, int replica, int z, int scale_x, int scale_y, double* density)
{
    int rc = sk_check(c, replica, z, scale_x, scale_y);
    if (rc) return rc;
    if (!density) return set_err(c, ISING_ERR_INVALID_ARG, "null density");
    HIP_TRY(c, hipSetDevice(c->device));
    SkBufs b;
    if ((rc = sk_fine_density(c, replica, z, scale_x, scale_y, b))) return rc;
    const Rep& r = c->reps[replica];
    const size_t n = (size_t)r.lx * scale_x * r.ly * scale_y;
    HIP_TRY(c, hipMemcpyAsync(density, b.d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_skyrmion_distribution(ising_ctx* c, int replica, int z, const ising_skyrmion_params* prm, int32_t* counts,
                                int32_t* table)
{
    if (!c || !prm || !counts) return ISING_ERR_INVALID_ARG;
    const ising_skyrmion_params P = *prm;
    int rc = sk_check(c, replica, z, P.scale_x, P.scale_y);
    if (rc) return rc;
    if (P.n_particles < 1 || P.num_thresholds < 1 || P.num_thresholds > 254 || P.num_sigmas < 1 ||
        !(P.min_threshold >= 0.0) || !(P.max_threshold >= P.min_threshold) || !(P.max_threshold <= 1.0))
        return set_err(c, ISING_ERR_INVALID_ARG, "skyrmion parameters (particles %d, thresholds %d, widths %d, "
                       "threshold range %g..%g)", P.n_particles, P.num_thresholds, P.num_sigmas, P.min_threshold,
                       P.max_threshold);
    HIP_TRY(c, hipSetDevice(c->device));
    SkBufs b;
    if ((rc = sk_fine_density(c, replica, z, P.scale_x, P.scale_y, b))) return rc;
    const Rep& r = c->reps[replica];
    const int NX = r.lx * P.scale_x, NY = r.ly * P.scale_y;
    const int64_t n = (int64_t)NX * NY;
    const int T = P.num_thresholds, S = P.num_sigmas, NP = P.n_particles;
    const size_t ntab = (size_t)T * S * NP;
    if ((rc = dalloc(c, &b.f, (size_t)n)) || (rc = dalloc(c, &b.root, (size_t)n)) || (rc = dalloc(c, &b.lvl, (size_t)n)) ||
        (rc = dalloc(c, &b.acc, (size_t)n)) || (rc = dalloc(c, &b.mx, 1)) || (rc = dalloc(c, &b.G, 25)) ||
        (rc = dalloc(c, &b.counts, ntab)))
        return rc;
    HIP_TRY(c, hipMemsetAsync(b.counts, 0, ntab * sizeof(int32_t), c->stream));
    std::vector<double> Gs((size_t)S * 25);
    for (int k = 1; k <= S; ++k) {
        const double sigma = 1.0 - ((double)k / (double)S) * (1.0 - 0.00001);
        double G[25];
        sk_gauss_weights(sigma, G);
        std::copy(G, G + 25, Gs.begin() + (size_t)(k - 1) * 25);
    }
    for (int k = 0; k < S; ++k) {
        HIP_TRY(c, hipMemcpyAsync(b.G, Gs.data() + (size_t)k * 25, 25 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_sk_filter, grid_of(n), dim3(kBlock), 0, c->stream, (const double*)b.d, NX, NY,
                           (const double*)b.G, b.f);
        HIP_TRY(c, hipMemsetAsync(b.mx, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_sk_absmax, dim3(1024), dim3(kBlock), 0, c->stream, (const double*)b.f, n, b.mx);
        hipLaunchKernelGGL(k_sk_level, grid_of(n), dim3(kBlock), 0, c->stream, (const double*)b.f, n,
                           (const unsigned long long*)b.mx, T, P.min_threshold, P.max_threshold, b.lvl);
        hipLaunchKernelGGL(k_sk_root, grid_of(n), dim3(kBlock), 0, c->stream, (const double*)b.f, NX, NY,
                           (const uint8_t*)b.lvl, T, b.root);
        HIP_TRY(c, hipMemsetAsync(b.acc, 0, (size_t)n * sizeof(unsigned long long), c->stream));
        for (int t = 1; t <= T; ++t) {
            hipLaunchKernelGGL(k_sk_accum, grid_of(n), dim3(kBlock), 0, c->stream, (const double*)b.f, n,
                               (const int32_t*)b.root, (const uint8_t*)b.lvl, t, b.acc);
            hipLaunchKernelGGL(k_sk_count, grid_of(n), dim3(kBlock), 0, c->stream, n, (const int32_t*)b.root,
                               (const uint8_t*)b.lvl, (const unsigned long long*)b.acc, t, NP,
                               b.counts + ((size_t)(t - 1) * S + k) * NP);
        }
        LAUNCH_CHECK(c);
        HIP_TRY(c, hipStreamSynchronize(c->stream));   // b.G is refilled for the next width
    }
    std::vector<int32_t> tab(ntab);
    HIP_TRY(c, hipMemcpyAsync(tab.data(), b.counts, ntab * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (table) std::copy(tab.begin(), tab.end(), table);
    // The layer's winding number and the search of compute_skyrmion_distribution
    // (src/reciprocalSpaceProcesses.f90:955-981): thresholds outer, widths inner;
    // the first entry within 0.1 of Q, else the first whose nint came closest.
    int64_t qi = 0;
    double q = 0.0;
    if ((rc = ising_winding(c, replica, z, &qi, &q))) return rc;
    const long long nq = llround(q);
    std::vector<int32_t> closest(NP, 0);
    for (int t = 0; t < T; ++t)
        for (int k = 0; k < S; ++k) {
            const int32_t* arr = tab.data() + ((size_t)t * S + k) * NP;
            double winding = 0.0;
            long long current = 0;
            for (int j = 1; j <= NP; ++j) {
                winding += (double)j * arr[j - 1];
                current += (long long)j * closest[j - 1];
            }
            if (std::llabs(llround(winding) - nq) < std::llabs(current - nq)) std::copy(arr, arr + NP, closest.begin());
            if (std::fabs(winding - q) < 0.1) {
                std::copy(arr, arr + NP, counts);
                return ISING_OK;
            }
        }
    std::copy(closest.begin(), closest.end(), counts);
    return ISING_OK;
}

int ising_winding(ising_ctx* c, int replica, int z, int64_t* q, double* q_raw)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (z < 0 || z >= c->reps[replica].lz) return set_err(c, ISING_ERR_INVALID_ARG, "layer %d out of range", z);
    std::vector<int64_t> qi((size_t)c->R * c->zmax);
    std::vector<double> qr((size_t)c->R * c->zmax);
    if ((rc = ising_winding_all(c, qi.data(), qr.data()))) return rc;
    if (q) *q = qi[(size_t)replica * c->zmax + z];
    if (q_raw) *q_raw = qr[(size_t)replica * c->zmax + z];
    return ISING_OK;
}

int ising_delta_energy(ising_ctx* c, int replica, const int32_t* sites, const double* props, int n, double* de)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!sites || !props || !de))) return set_err(c, ISING_ERR_INVALID_ARG, "null argument");
    if (n == 0) return ISING_OK;
    const Rep& r = c->reps[replica];
    std::vector<float> pf((size_t)3 * n);
    for (int k = 0; k < n; ++k) {
        if (sites[k] < 0 || sites[k] >= r.n) return set_err(c, ISING_ERR_INVALID_ARG, "site %d out of range", sites[k]);
        for (int m = 0; m < 3; ++m) pf[3 * k + m] = (float)props[3 * k + m];
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (n > c->de_cap) {
        if (c->d_sites) (void)hipFree(c->d_sites);
        if (c->d_props) (void)hipFree(c->d_props);
        if (c->d_de) (void)hipFree(c->d_de);
        c->d_sites = nullptr; c->d_props = nullptr; c->d_de = nullptr; c->de_cap = 0;
        if ((rc = dalloc(c, &c->d_sites, n)) || (rc = dalloc(c, &c->d_props, 3 * (size_t)n)) || (rc = dalloc(c, &c->d_de, n)))
            return rc;
        c->de_cap = n;
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_sites, sites, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_props, pf.data(), pf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_delta_energy, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                       (const float*)cur_spins(c), c->d_reps, c->d_lads, replica, (const int32_t*)c->d_sites,
                       (const float*)c->d_props, n, c->d_de, (const double*)c->d_demag);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(de, c->d_de, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_pt_run(ising_ctx* c, const ising_schedule* sched, const ising_series* out)
{
    return pt_run_impl(c, sched, out, nullptr);
}

int ising_pt_run_static(ising_ctx* c, const ising_schedule* sched, const ising_series* out)
{
    if (!c || !sched) return ISING_ERR_INVALID_ARG;
    if (c->device >= kMaxDevices || !g_scratch_ptr[c->device])
        return set_err(c, ISING_ERR_STATE, "no static scratch address for device %d", c->device);
    const ising_schedule& s = *sched;
    if (s.n_sweeps < 0 || s.measure_every < 0 || s.winding_every < 0)
        return set_err(c, ISING_ERR_INVALID_ARG, "negative schedule entry");
    const size_t n_meas = s.measure_every ? (size_t)(s.n_sweeps / s.measure_every) : 0u;
    const size_t n_q = s.winding_every ? (size_t)(s.n_sweeps / s.winding_every) : 0u;
    const size_t need = series_layout(n_meas, n_q, (size_t)c->R, s.winding_layer > 0 ? 1u : (size_t)c->zmax).total;
    if (need > kStaticScratchBytes)
        return set_err(c, ISING_ERR_INVALID_ARG, "series of %zu bytes exceed the static scratch of %zu bytes "
                       "(ISING_STATIC_SCRATCH_BYTES)", need, kStaticScratchBytes);
    return pt_run_impl(c, sched, out, (unsigned char*)g_scratch_ptr[c->device]);
}

int ising_save_state(ising_ctx* c)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    if (!c->ck_spins) {
        if ((rc = dalloc(c, &c->ck_spins, (size_t)c->total_floats)) || (rc = dalloc(c, &c->ck_sot, c->R)) ||
            (rc = dalloc(c, &c->ck_bslot, c->R)) || (rc = dalloc(c, &c->ck_btemp, c->R)) || (rc = dalloc(c, &c->ck_ctr, kCounters)))
            return rc;
    }
    HIP_TRY(c, hipMemcpyAsync(c->ck_spins, cur_spins(c), (size_t)c->total_floats * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ck_sot, c->d_slot_of_temp, c->R * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ck_bslot, c->d_beta_slot, c->R * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ck_btemp, c->d_beta_temp, c->R * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->ck_ctr, c->d_ctr, kCounters * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->ck_host_sweep = c->host_sweep;
    c->ck_host_fused = c->host_fused;
    c->ck_valid = true;
    return ISING_OK;
}

int ising_restore_state(ising_ctx* c)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (!c->ck_valid) return set_err(c, ISING_ERR_STATE, "no saved state: call ising_save_state first");
    HIP_TRY(c, hipSetDevice(c->device));
    float* dst = c->d_spins + (int64_t)(c->ck_host_fused & 1u) * c->total_floats;
    HIP_TRY(c, hipMemcpyAsync(dst, c->ck_spins, (size_t)c->total_floats * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_slot_of_temp, c->ck_sot, c->R * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_beta_slot, c->ck_bslot, c->R * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_beta_temp, c->ck_btemp, c->R * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_ctr, c->ck_ctr, kCounters * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->host_sweep = c->ck_host_sweep;
    c->host_fused = c->ck_host_fused;
    return ISING_OK;
}

int ising_temperature_map(ising_ctx* c, int32_t* slot_of_temp)
{
    if (!c || !slot_of_temp) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(slot_of_temp, c->d_slot_of_temp, c->R * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_swap_stats(ising_ctx* c, int64_t* attempts, int64_t* accepts)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    if (attempts) HIP_TRY(c, hipMemcpyAsync(attempts, c->d_swap_att, c->R * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    if (accepts) HIP_TRY(c, hipMemcpyAsync(accepts, c->d_swap_acc, c->R * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_time_sweeps(ising_ctx* c, int n_sweeps, double* ms_per_launch)
{
    if (!c || !ms_per_launch || n_sweeps < 1) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, timing_event(&e0));
    if (timing_event(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_err(c, ISING_ERR_HIP, "hipEventCreate"); }
    int rc = ISING_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    for (int k = 0; k < n_sweeps && !rc; ++k) {
        rc = launch_sweep(c, (uint32_t)k);
    }
    if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    if (!rc) rc = launch_advance(c, (uint32_t)n_sweeps);
    float ms = 0.f;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    advance_host(c, (uint32_t)n_sweeps);
    *ms_per_launch = (double)ms / ((double)c->max_ncol * n_sweeps);
    return ISING_OK;
}

int ising_time_demag(ising_ctx* c, int n_fields, double* ms_per_field)
{
    if (!c || !ms_per_field || n_fields < 1) return ISING_ERR_INVALID_ARG;
    if (!c->any_demag) return set_err(c, ISING_ERR_INVALID_ARG, "no ladder has the demagnetising term");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = demag_setup(c);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, timing_event(&e0));
    if (timing_event(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_err(c, ISING_ERR_HIP, "hipEventCreate"); }
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    for (int k = 0; k < n_fields && !rc; ++k) rc = launch_demag(c, kDemagLadders);
    if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    float ms = 0.f;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms_per_field = (double)ms / n_fields;
    return ISING_OK;
}

int ising_time_pt_sweeps(ising_ctx* c, const ising_schedule* sched, int n_sweeps, double* ms_per_launch)
{
    if (!c || !sched || !ms_per_launch || n_sweeps < 1) return ISING_ERR_INVALID_ARG;
    if (c->any_demag)
        return set_err(c, ISING_ERR_UNSUPPORTED, "parallel tempering with the demagnetisation term");
    const ising_schedule s = *sched;
    if (s.swap_every < 0 || s.measure_every < 0 || s.winding_every < 0 || s.winding_layer < 0 || s.winding_layer > c->zmax)
        return set_err(c, ISING_ERR_INVALID_ARG, "bad schedule");
    HIP_TRY(c, hipSetDevice(c->device));
    int64_t period = 1;
    for (int v : { s.swap_every, s.measure_every, s.winding_every })
        if (v > 0) period = period / gcd64(period, v) * v;
    LaunchTimer timer;
    // at most two plain launches per sweep, two events each (three-colour
    // contexts take more and create the rest as they go)
    if (timer.reserve(4 * (size_t)n_sweeps) != hipSuccess) return set_err(c, ISING_ERR_HIP, "hipEventCreate");
    SeriesDev none;                      // records are computed, not stored
    int rc = ISING_OK, done = 0;
    // The record counters restart as in ising_pt_run (they only index series).
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, c->d_ctr + 1, 2, 0u);
    LAUNCH_CHECK(c);
    while (done < n_sweeps && !rc) {
        const int cnt = (int)std::min<int64_t>(period, n_sweeps - done);
        rc = enqueue_period(c, s, cnt, none, &timer);
        done += cnt;
    }
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    advance_host(c, (uint32_t)n_sweeps);
    double total = 0.0;
    for (size_t k = 0; k + 1 < timer.used; k += 2) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, timer.ev[k], timer.ev[k + 1]));
        total += ms;
    }
    const size_t pairs = timer.used / 2;
    if (pairs == 0) return set_err(c, ISING_ERR_STATE, "the schedule has no plain sweep launch");
    *ms_per_launch = total / ((double)pairs * timer.launches_per_pair);
    return ISING_OK;
}

int ising_synchronize(ising_ctx* c)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());
    return ISING_OK;
}

int ising_bandwidth_probe(ising_ctx* c, int64_t bytes, double* ms)
{
    if (!c || !ms || bytes < 16) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = (size_t)bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = dalloc(c, &a, n))) return rc;
    if ((rc = dalloc(c, &b, n))) { (void)hipFree(a); return rc; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = ISING_OK;
    if (timing_event(&e0) != hipSuccess || timing_event(&e1) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventCreate");
    // one 16-byte element per thread: on MI355X this reaches ~6.1 TB/s over
    // buffers beyond the Infinity Cache, where grid-stride loops stay near 5
    const unsigned grid = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, (size_t)1 << 30);
    if (!rc && hipMemsetAsync(a, 0, n * sizeof(float4), c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "memset");
    if (!rc) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kBlock), 0, c->stream, (const float4*)a, b, n);
    if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    if (!rc) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kBlock), 0, c->stream, (const float4*)b, a, n);
    if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "hipEventRecord");
    float t = 0.f;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "elapsed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc) return rc;
    *ms = t;
    return ISING_OK;
}

int ising_set_tuning(ising_ctx* c, uint32_t flags)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (flags & ~(kTuneXcdRemap | kTuneNtOwn | kTuneFusedMask | kTuneSlabMask))
        return set_err(c, ISING_ERR_INVALID_ARG, "unknown tuning flags %u", flags);
    const int lc = (int)((flags & kTuneFusedMask) >> kTuneFusedShift);
    const int hy = (int)((flags & kTuneSlabMask) >> kTuneSlabShift);
    if (hy > 0 && lc == 0)
        return set_err(c, ISING_ERR_INVALID_ARG, "y slabs (tuning flags %u) are a shape of the fused sweep", flags);
    if (lc > 0 && c->any_demag)
        return set_err(c, ISING_ERR_UNSUPPORTED, "fused sweep: a ladder has the demagnetisation term");
    if (lc > 0 && !c->fused_ok)
        return set_err(c, ISING_ERR_UNSUPPORTED,
                       "fused sweep unavailable: it needs simple-cubic replicas whose z planes of one colour "
                       "fit one workgroup (%d sites) or y slabs of at least two rows of it", kFusedMaxPairs);
    HIP_TRY(c, hipSetDevice(c->device));
    if (lc != fused_lc(c) || hy != fused_hy(c)) {
        int rc = fused_layout(c, lc, hy);
        if (rc) return rc;
    }
    c->tune = flags;
    return ISING_OK;
}

int ising_get_tuning(const ising_ctx* c, uint32_t* flags)
{
    if (!c || !flags) return ISING_ERR_INVALID_ARG;
    *flags = c->tune;
    return ISING_OK;
}

// Estimated time of a fused layout (lc planes, y slabs of hy rows): a sweep
// runs in rounds of one workgroup per compute unit, and a round lasts as long
// as its largest workgroup, whose work is its (lc + 2) planes times its window
// rows. 0 when some replica has no such layout.
static double fused_layout_cost(const ising_ctx* c, int lc, int hy, int n_cu)
{
    int64_t nwg = 0;
    int64_t work = 0;
    for (const Rep& r : c->reps) {
        const SlabShape sh = slab_shape(r, hy);
        if (sh.nslab == 0) return 0.0;
        nwg += (int64_t)((r.lz + lc - 1) / lc) * sh.nslab;
        work = std::max<int64_t>(work, (int64_t)(std::min(lc, r.lz) + 2) * sh.rows * r.hx);
    }
    return (double)((nwg + n_cu - 1) / n_cu) * (double)work;
}

// Flag sets the autotuner times: the half-sweep pair with and without
// die-contiguous blocks and non-temporal own-colour access, then the fused
// sweep with 8, 16, 32 or all z planes per workgroup, with and without
// die-contiguous blocks. When every replica's planes are cut into y slabs,
// each chunk size takes the slab height of the lowest fused_layout_cost
// (balanced heights, from the tallest window down to a quarter of it), and
// a chunk of 24 planes joins the list: the layouts whose workgroup count
// fills the last round win (on 128^3 x 8 replicas, 32 planes x 16 rows).
static std::vector<uint32_t> tuning_candidates(const ising_ctx* c)
{
    std::vector<uint32_t> v = { 0u, kTuneXcdRemap, kTuneNtOwn, kTuneXcdRemap | kTuneNtOwn };
    if (c->fused_ok && !c->any_demag) {
        bool all_slab = true;
        for (const Rep& r : c->reps) all_slab = all_slab && r.plane > kFusedMaxPairs;
        std::vector<int> lcs;
        const std::initializer_list<int> base = { 8, 16, 32, c->zmax }, slabbed = { 8, 16, 24, 32, c->zmax };
        for (int lc : all_slab ? slabbed : base) {
            lc = std::min(std::min(lc, c->zmax), 255);
            if (lc > 0 && std::find(lcs.begin(), lcs.end(), lc) == lcs.end()) lcs.push_back(lc);
        }
        int n_cu = 0;
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || n_cu < 1)
            n_cu = 256;
        for (int lc : lcs) {
            uint32_t f = (uint32_t)lc << kTuneFusedShift;
            if (all_slab) {
                // balanced heights: ceil(ly / n) for n slabs, n from the fewest up to four times as many
                const Rep& r0 = *std::max_element(c->reps.begin(), c->reps.end(),
                                                  [](const Rep& a, const Rep& b) { return a.plane < b.plane; });
                const int hmax = kFusedMaxPairs / r0.hx - kSlabHalo;
                const int n0 = (r0.ly + hmax - 1) / hmax;
                double best = 0.0;
                int best_hy = 0;
                for (int n = n0; n <= 4 * n0 && hmax >= 2; ++n) {
                    const int hy = (r0.ly + n - 1) / n;
                    if (hy < 2) break;
                    const double cost = fused_layout_cost(c, lc, hy, n_cu);
                    if (cost > 0.0 && (best_hy == 0 || cost < best)) { best = cost; best_hy = hy; }
                }
                if (best_hy == 0) continue;
                f |= (uint32_t)std::min(best_hy, 255) << kTuneSlabShift;
            }
            v.push_back(f);
            v.push_back(f | kTuneXcdRemap);
        }
    }
    if (v.size() > ISING_TUNE_MAX_CANDIDATES) v.resize(ISING_TUNE_MAX_CANDIDATES);
    return v;
}

int ising_tuning_candidates(const ising_ctx* c, uint32_t* flags, int* count)
{
    if (!c || !count) return ISING_ERR_INVALID_ARG;
    const std::vector<uint32_t> v = tuning_candidates(c);
    *count = (int)v.size();
    if (flags) std::copy(v.begin(), v.end(), flags);
    return ISING_OK;
}

int ising_autotune(ising_ctx* c, int n_sweeps, uint32_t* best, double* ms)
{
    if (!c || n_sweeps < 1) return ISING_ERR_INVALID_ARG;
    const std::vector<uint32_t> cand = tuning_candidates(c);
    std::vector<double> t(cand.size());
    size_t arg = 0;
    for (size_t k = 0; k < cand.size(); ++k) {
        int rc = ising_set_tuning(c, cand[k]);
        if (!rc) rc = ising_time_sweeps(c, n_sweeps, &t[k]);
        if (rc) return rc;
        if (t[k] < t[arg]) arg = k;
    }
    int rc = ising_set_tuning(c, cand[arg]);
    if (rc) return rc;
    if (best) *best = cand[arg];
    if (ms) for (size_t k = 0; k < cand.size(); ++k) ms[k] = t[k];
    return ISING_OK;
}

// Sweeps-per-swap schedule: times each candidate interval over n_sweeps
// sweeps of the schedule (eager periods, one hipEvent pair around the whole
// run of a candidate; two passes in alternating order, the faster kept) and
// picks the most frequent exchange whose time per sweep is within
// (1 + max_overhead) of the cheapest candidate's. Deterministic even/odd
// exchange mixes temperatures fastest when it runs most often (Syed,
// Bouchard-Cote, Deligiannidis, Doucet 2019: round trips of non-reversible
// PT take O(N) rounds), so the device time of a round is its only cost.
int ising_autotune_schedule(ising_ctx* c, ising_schedule* sched, const int32_t* cand, int n_cand, int n_sweeps,
                            double max_overhead, double* ms_per_sweep)
{
    if (!c || !sched || !cand || n_cand < 1 || n_cand > ISING_TUNE_MAX_CANDIDATES || n_sweeps < 1 ||
        !(max_overhead >= 0.0) || !std::isfinite(max_overhead))
        return ISING_ERR_INVALID_ARG;
    if (c->any_demag)
        return set_err(c, ISING_ERR_UNSUPPORTED, "parallel tempering with the demagnetisation term");
    for (int k = 0; k < n_cand; ++k)
        if (cand[k] < 1) return set_err(c, ISING_ERR_INVALID_ARG, "swap interval candidate %d", cand[k]);
    ising_schedule s = *sched;
    if (s.measure_every < 0 || s.winding_every < 0 || s.winding_layer < 0 || s.winding_layer > c->zmax)
        return set_err(c, ISING_ERR_INVALID_ARG, "bad schedule");
    HIP_TRY(c, hipSetDevice(c->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, timing_event(&e0));
    if (timing_event(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_err(c, ISING_ERR_HIP, "hipEventCreate"); }
    std::vector<double> t(n_cand, 1e300);
    SeriesDev none;
    int rc = ISING_OK;
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        for (int j = 0; j < n_cand && !rc; ++j) {
            const int k = pass == 0 ? j : n_cand - 1 - j;
            s.swap_every = cand[k];
            int64_t period = 1;
            for (int v : { s.swap_every, s.measure_every, s.winding_every })
                if (v > 0) period = period / gcd64(period, v) * v;
            if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = set_err(c, ISING_ERR_HIP, "hipEventRecord"); break; }
            for (int done = 0; done < n_sweeps && !rc;) {
                const int cnt = (int)std::min<int64_t>(period, n_sweeps - done);
                rc = enqueue_period(c, s, cnt, none);
                done += cnt;
            }
            if (rc) break;
            float ms = 0.f;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = set_err(c, ISING_ERR_HIP, "schedule timing");
                break;
            }
            advance_host(c, (uint32_t)n_sweeps);
            t[k] = std::min(t[k], (double)ms / n_sweeps);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    const double fastest = *std::min_element(t.begin(), t.end());
    int best = -1;
    for (int k = 0; k < n_cand; ++k)
        if (t[k] <= (1.0 + max_overhead) * fastest && (best < 0 || cand[k] < cand[best])) best = k;
    sched->swap_every = cand[best];
    if (ms_per_sweep) for (int k = 0; k < n_cand; ++k) ms_per_sweep[k] = t[k];
    return ISING_OK;
}

// ---- learned proposal ------------------------------------------------------

static int mlp_pack(ising_ctx* c)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, c->stream, (const float*)c->d_mlp_prm, c->d_mlp);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

int ising_mlp_set_params(ising_ctx* c, const float* prm, int n)
{
    if (!c || !prm) return ISING_ERR_INVALID_ARG;
    if (n != kMlpParams) return set_err(c, ISING_ERR_INVALID_ARG, "network has %d parameters, got %d", kMlpParams, n);
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(prm[k])) return set_err(c, ISING_ERR_INVALID_ARG, "parameter %d is not finite", k);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->d_mlp_prm, prm, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_mlp_m1, 0, n * sizeof(float), c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_mlp_m2, 0, n * sizeof(float), c->stream));
    c->adam_step = 0;
    int rc = mlp_pack(c);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_mlp_get_params(ising_ctx* c, float* prm, int n)
{
    if (!c || !prm) return ISING_ERR_INVALID_ARG;
    if (n != kMlpParams) return set_err(c, ISING_ERR_INVALID_ARG, "network has %d parameters, got %d", kMlpParams, n);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(prm, c->d_mlp_prm, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_mlp_enable(ising_ctx* c, int ladder, int on)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (ladder >= (int)c->ladders.size()) return set_err(c, ISING_ERR_INVALID_ARG, "ladder %d", ladder);
    // The network's input is the six neighbours of a simple-cubic site.
    for (int k = 0; k < (int)c->ladders.size(); ++k)
    {
        if (on && (ladder < 0 || k == ladder) && c->ladders[k].lat.cell != ISING_CELL_SC)
            return set_err(c, ISING_ERR_UNSUPPORTED, "learned proposal: ladder %d is not simple cubic", k);
        if (on && c->ladders[k].ham.demag)
            return set_err(c, ISING_ERR_UNSUPPORTED, "learned proposal: ladder %d of this context has the "
                           "demagnetisation term", k);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    for (int k = 0; k < (int)c->learned.size(); ++k)
        if (ladder < 0 || k == ladder) c->learned[k] = on ? 1 : 0;
    c->any_learned = false;
    for (int v : c->learned) c->any_learned = c->any_learned || v;
    return upload_lads(c);
}

int ising_mlp_collect(ising_ctx* c, int per_replica)
{
    if (!c || per_replica < 1) return ISING_ERR_INVALID_ARG;
    for (const LadHost& L : c->ladders)
        if (L.lat.cell != ISING_CELL_SC || L.ham.demag)
            return set_err(c, ISING_ERR_UNSUPPORTED,
                           "training records: every ladder must be simple cubic, without the demagnetisation term");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t add = (int64_t)c->R * per_replica;
    if (c->n_recs + add > (int64_t)UINT32_MAX) return set_err(c, ISING_ERR_INVALID_ARG, "too many training records");
    if (c->n_recs + add > c->cap_recs) {
        int64_t cap = std::max<int64_t>(2 * c->cap_recs, c->n_recs + add);
        float* nb = nullptr;
        int rc = dalloc(c, &nb, (size_t)cap * kRecord);
        if (rc) return rc;
        if (c->n_recs)
            HIP_TRY(c, hipMemcpyAsync(nb, c->d_recs, (size_t)c->n_recs * kRecord * sizeof(float),
                                      hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (c->d_recs) (void)hipFree(c->d_recs);
        c->d_recs = nb;
        c->cap_recs = cap;
    }
    hipLaunchKernelGGL(k_mlp_collect, dim3((unsigned)((add + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                       (const float*)cur_spins(c), c->d_reps, c->d_lads, (const float*)c->d_beta_slot, c->R, per_replica,
                       c->collect_calls, (uint32_t)c->seed, (uint32_t)(c->seed >> 32),
                       c->d_recs + (size_t)c->n_recs * kRecord);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->n_recs += add;
    c->collect_calls++;
    return ISING_OK;
}

int ising_mlp_num_records(const ising_ctx* c, int64_t* n)
{
    if (!c || !n) return ISING_ERR_INVALID_ARG;
    *n = c->n_recs;
    return ISING_OK;
}

int ising_mlp_clear_records(ising_ctx* c)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    c->n_recs = 0;
    return ISING_OK;
}

int ising_mlp_get_records(ising_ctx* c, float* out, int64_t n)
{
    if (!c || !out || n < 0 || n > c->n_recs) return ISING_ERR_INVALID_ARG;
    if (n == 0) return ISING_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(out, c->d_recs, (size_t)n * kRecord * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

static int mlp_grad_launch(ising_ctx* c, const float* recs, uint32_t n_recs, uint32_t batch, uint32_t step, int hashed)
{
    const int tiles = (int)((batch + 63) / 64);
    const int blocks = std::max(1, std::min(256, (tiles + 3) / 4));
    const int waves = blocks * 4;
    if (waves > c->wave_grad_waves) {
        if (c->d_wave_grad) (void)hipFree(c->d_wave_grad);
        c->d_wave_grad = nullptr;
        c->wave_grad_waves = 0;
        int rc = dalloc(c, &c->d_wave_grad, (size_t)waves * (kMlpParams + 1));
        if (rc) return rc;
        c->wave_grad_waves = waves;
    }
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, c->stream, recs, n_recs, batch, step, hashed,
                       (const MlpDev*)c->d_mlp, (uint32_t)c->seed, (uint32_t)(c->seed >> 32), c->d_wave_grad);
    LAUNCH_CHECK(c);
    const int groups = std::min(kMlpReduceGroups, waves), per_group = (waves + groups - 1) / groups;
    hipLaunchKernelGGL(k_mlp_reduce_groups, dim3((kMlpParams + 1 + 255) / 256, groups), dim3(256), 0, c->stream,
                       (const float*)c->d_wave_grad, waves, per_group, c->d_grad_part);
    LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, c->stream,
                       (const float*)c->d_grad_part, groups, c->d_mlp_grad);
    LAUNCH_CHECK(c);
    return ISING_OK;
}

int ising_mlp_train(ising_ctx* c, int steps, int batch, double lr, double* loss)
{
    if (!c || steps < 0 || batch < 1 || !(lr > 0.0)) return ISING_ERR_INVALID_ARG;
    if (c->n_recs < 1) return set_err(c, ISING_ERR_STATE, "no training records: call ising_mlp_collect first");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<float> lsum(loss ? steps : 0);
    for (int t = 0; t < steps; ++t) {
        int rc = mlp_grad_launch(c, c->d_recs, (uint32_t)c->n_recs, (uint32_t)batch, (uint32_t)c->adam_step, 1);
        if (rc) return rc;
        c->adam_step++;
        const float bc1 = 1.f - powf(0.9f, (float)c->adam_step), bc2 = 1.f - powf(0.999f, (float)c->adam_step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, c->d_mlp_prm,
                           c->d_mlp_m1, c->d_mlp_m2, (const float*)c->d_mlp_grad, (float)lr, bc1, bc2);
        LAUNCH_CHECK(c);
        if ((rc = mlp_pack(c))) return rc;
        if (loss) HIP_TRY(c, hipMemcpyAsync(&lsum[t], c->d_mlp_grad + kMlpParams, sizeof(float),
                                            hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int t = 0; t < (loss ? steps : 0); ++t) loss[t] = (double)lsum[t] / (double)batch;
    return ISING_OK;
}

int ising_mlp_forward(ising_ctx* c, const float* x, int n, float* f)
{
    if (!c || n < 0 || (n > 0 && (!x || !f))) return ISING_ERR_INVALID_ARG;
    if (n == 0) return ISING_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    float *dx = nullptr, *df = nullptr;
    int rc = dalloc(c, &dx, (size_t)n * kMlpIn);
    if (rc) return rc;
    if ((rc = dalloc(c, &df, (size_t)n))) { (void)hipFree(dx); return rc; }
    if (hipMemcpyAsync(dx, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "upload");
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_forward_rows, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float*)dx, n,
                           (const MlpDev*)c->d_mlp, df);
        if (hipGetLastError() != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "launch");
    }
    if (!rc && hipMemcpyAsync(f, df, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "download");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    (void)hipFree(dx);
    (void)hipFree(df);
    return rc;
}

int ising_mlp_grad(ising_ctx* c, const float* recs, int n, uint32_t step, double* loss, float* grad)
{
    if (!c || n < 1 || !recs || !grad) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    float* dr = nullptr;
    int rc = dalloc(c, &dr, (size_t)n * kRecord);
    if (rc) return rc;
    if (hipMemcpyAsync(dr, recs, (size_t)n * kRecord * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "upload");
    if (!rc) rc = mlp_grad_launch(c, dr, (uint32_t)n, (uint32_t)n, step, 0);
    std::vector<float> g(kMlpParams + 1);
    if (!rc && hipMemcpyAsync(g.data(), c->d_mlp_grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "download");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    (void)hipFree(dr);
    if (rc) return rc;
    for (int k = 0; k < kMlpParams; ++k) grad[k] = g[k];
    if (loss) *loss = (double)g[kMlpParams] / (double)n;
    return ISING_OK;
}

int ising_mfma_probe(ising_ctx* c, const float* A, const float* B, float* D)
{
    if (!c || !A || !B || !D) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    float* d = nullptr;
    int rc = dalloc(c, &d, 32 * 16 + 16 * 32 + 32 * 32);
    if (rc) return rc;
    if (hipMemcpyAsync(d, A, 512 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d + 512, B, 512 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "upload");
    if (!rc) {
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, c->stream, (const float*)d, (const float*)(d + 512), d + 1024);
        if (hipGetLastError() != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "launch");
    }
    if (!rc && hipMemcpyAsync(D, d + 1024, 1024 * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = set_err(c, ISING_ERR_HIP, "download");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, ISING_ERR_HIP, "sync");
    (void)hipFree(d);
    return rc;
}

int ising_accept_counts(ising_ctx* c, int64_t* accepted)
{
    if (!c || !accepted) return ISING_ERR_INVALID_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(k_acc_finalize, dim3(c->R), dim3(kBlock), 0, c->stream, c->d_reps,
                       (const unsigned long long*)c->d_acc_wave, (const unsigned long long*)c->d_acc_wave_f,
                       c->fused_wpb, (const unsigned long long*)c->d_acc_base, c->d_acc);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(accepted, c->d_acc, c->R * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

// ---- demagnetisation ---------------------------------------------------------

static int demag_replica_args(ising_ctx* c, int replica, const void* h, int64_t n)
{
    int rc = check_replica(c, replica);
    if (rc) return rc;
    if (!h) return set_err(c, ISING_ERR_INVALID_ARG, "null field");
    const Rep& r = c->reps[replica];
    if (n != r.n) return set_err(c, ISING_ERR_INVALID_ARG, "n = %lld, replica has %lld sites", (long long)n, (long long)r.n);
    return check_demag_lattice(c, c->ladders[r.ladder].lat);
}

static int copy_out_demag(ising_ctx* c, int replica, double* h)
{
    const Rep& r = c->reps[replica];
    std::vector<double> dev((size_t)3 * r.n);
    HIP_TRY(c, hipMemcpyAsync(dev.data(), c->d_demag + r.off, dev.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < r.n; ++i) {
        const int64_t t = dev_index(r, i);
        for (int k = 0; k < 3; ++k) h[k * r.n + i] = dev[(size_t)k * r.n + t];
    }
    return ISING_OK;
}

int ising_demag_field(ising_ctx* c, int replica, double* h, int64_t n)
{
    int rc = demag_replica_args(c, replica, h, n);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = launch_demag(c, replica))) return rc;
    return copy_out_demag(c, replica, h);
}

int ising_get_demag_field(ising_ctx* c, int replica, double* h, int64_t n)
{
    int rc = demag_replica_args(c, replica, h, n);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = demag_setup(c))) return rc;
    return copy_out_demag(c, replica, h);
}

int ising_set_demag_field(ising_ctx* c, int replica, const double* h, int64_t n)
{
    int rc = demag_replica_args(c, replica, h, n);
    if (rc) return rc;
    const Rep& r = c->reps[replica];
    std::vector<double> dev((size_t)3 * r.n);
    for (int64_t i = 0; i < r.n; ++i) {
        const int64_t t = dev_index(r, i);
        for (int k = 0; k < 3; ++k) {
            const double v = h[k * r.n + i];
            if (!std::isfinite(v)) return set_err(c, ISING_ERR_INVALID_ARG, "non-finite field at site %lld", (long long)i);
            dev[(size_t)k * r.n + t] = v;
        }
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = demag_setup(c))) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->d_demag + r.off, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ISING_OK;
}

int ising_set_demag_interval(ising_ctx* c, int sweeps)
{
    if (!c) return ISING_ERR_INVALID_ARG;
    if (sweeps < 0) return set_err(c, ISING_ERR_INVALID_ARG, "interval %d", sweeps);
    c->demag_interval = sweeps;
    return ISING_OK;
}

} // extern "C"
