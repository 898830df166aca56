// mlp.hpp -- the learned cone-width network on CDNA4 matrix cores.
//
// Network (per site): x[21] = six neighbour spins (+x,-x,+y,-y,+z,-z) and the
// reduced field beta*h; z1 = W1 x + b1, h1 = g(z1); z2 = W2 h1 + b2,
// h2 = g(z2); f = w3.h2 + b3; cone radius r = exp(clamp(f)); g is SiLU in
// base 2 (see silu below). The input
// never contains the site's own spin, so the proposal radius is the same for
// S -> S' and S' -> S and the Metropolis step stays exact.
//
// One wavefront evaluates 64 sites (one per lane) with v_mfma_f32_32x32x16_bf16
// in the transposed form Z^T = W X^T: the 32 hidden units are the M rows, the
// sites the N columns (two blocks of 32 sites: lanes 0-31 and 32-63), so a
// lane keeps its own site throughout:
//   * operand A (weights, M x K): lane l holds row l%32, k = 8*(l/32) + 0..7;
//   * operand B (activations, K x N): lane l holds column l%32, same k;
//   * result (M x N): lane l holds column l%32, rows 8*(r/4) + 4*(l/32) + r%4
//     in accumulator register r (0..15).
// Layer 1 exchanges half of each site's features between lanes l and l^32
// (the other half of the B column); layer 2 takes layer 1's result registers
// directly as its B operand because the K order of W2 is permuted to match
// (unit U(kc, g, j) = 8*(2kc + j/4) + 4g + j%4 sits at k = 16kc + 8g + j);
// layer 3 is a 16-term dot product per lane plus one exchange. No LDS.
// Weights are bfloat16, accumulation FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"

namespace ising {

constexpr int kMlpIn = 21;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;  // 1793
constexpr float kLogRMin = -4.605170185988091f;   // ln 0.01
constexpr float kLogRMax = 2.0794415416798357f;   // ln 8
constexpr uint32_t kStreamTrain = 3u;
constexpr uint32_t kStreamCollect = 4u;
constexpr uint32_t kTrainKey = 0x6D6C7000u;

// Offsets of the parameter groups in the FP32 master vector.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Device copy of the network, refreshed from the FP32 master parameters.
struct MlpDev {
    bf16x8 w1f[2][64];    // A fragments of W1:      W1[l%32][16kc + 8(l/32) + j]
    bf16x8 w2f[2][64];    // A fragments of W2:      W2[l%32][U(kc, l/32, j)]
    bf16x8 w2tf[2][64];   // A fragments of W2^T:    W2[U(kc, l/32, j)][l%32]
    float b1[kMlpH], b2[kMlpH], w3[kMlpH];   // w3 holds its bfloat16 value
    float b3;
    float pad[3];
};

__host__ __device__ constexpr int acc_unit(int r, int g) { return 8 * (r >> 2) + 4 * g + (r & 3); }
__host__ __device__ constexpr int perm_unit(int kc, int g, int j) { return 8 * (2 * kc + (j >> 2)) + 4 * g + (j & 3); }

__device__ __forceinline__ f32x16 mfma32(const bf16x8& a, const bf16x8& b, const f32x16& c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4u as_u32(const bf16x8& v) { return __builtin_bit_cast(v4u, v); }
__device__ __forceinline__ bf16x8 as_bf(const v4u& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ v4u sel4(bool c, const v4u& a, const v4u& b)
{
    return v4u{ c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w };
}

__device__ __forceinline__ v4u xor32(const v4u& v)
{
    return v4u{ (uint32_t)__shfl_xor((int)v.x, 32, 64), (uint32_t)__shfl_xor((int)v.y, 32, 64),
                (uint32_t)__shfl_xor((int)v.z, 32, 64), (uint32_t)__shfl_xor((int)v.w, 32, 64) };
}

// The activation is SiLU in base 2, g(z) = z / (1 + 2^-z) = log2(e) silu(z ln 2):
// a SiLU network with its weights rescaled by ln 2, so it represents the same
// functions, and each activation needs no extra multiply before the exp2.
__device__ __forceinline__ float silu(float z)
{
    return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z));
}
__device__ __forceinline__ float dsilu(float z)
{
    const float s = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z));
    return s * (1.f + z * 0.6931471805599453f * (1.f - s));
}

// Accumulator registers 8kc..8kc+7 as a bfloat16 B fragment.
__device__ __forceinline__ bf16x8 chunk_bf16(const f32x16& a, int kc)
{
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)a[8 * kc + j];
    return v;
}

// Per-lane copies of a 32-vector in accumulator order: out[r] = v[acc_unit(r, g)].
__device__ __forceinline__ f32x16 lane_vec(const float* __restrict__ v, int g)
{
    f32x16 o;
    const float4* p = (const float4*)v;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 t = p[2 * q + g];
        o[4 * q] = t.x; o[4 * q + 1] = t.y; o[4 * q + 2] = t.z; o[4 * q + 3] = t.w;
    }
    return o;
}

// Saved activations of one wavefront for the backward pass.
struct MlpTape {
    bf16x8 xb[2][2];     // layer-1 B fragments [block][kc]
    bf16x8 h1b[2][2];    // layer-2 B fragments [block][kc]
    f32x16 z1[2], z2[2];
};

// Forward pass of the 64 sites of a wavefront; x[] are this lane's 21
// features. Returns f of this lane's site. Every lane of the wavefront must
// call it (it exchanges data across lanes).
template <bool kTape>
__device__ __forceinline__ float mlp_forward(const MlpDev& P, const float (&x)[kMlpIn], MlpTape* tape)
{
    const int l = threadIdx.x & 63, g = l >> 5;
    // Features in four bfloat16 groups of 8 (21..31 are zero).
    bf16x8 F[4];
#pragma unroll
    for (int k = 0; k < 32; ++k) F[k >> 3][k & 7] = (__bf16)(k < kMlpIn ? x[k] : 0.f);
    // The group (2kc + g) stays; the group (2kc + 1 - g) goes to the partner lane.
    v4u own[2], recv[2];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        own[kc] = sel4(g == 0, as_u32(F[2 * kc]), as_u32(F[2 * kc + 1]));
        recv[kc] = xor32(sel4(g == 0, as_u32(F[2 * kc + 1]), as_u32(F[2 * kc])));
    }
    const bf16x8 a1[2] = { P.w1f[0][l], P.w1f[1][l] };
    const bf16x8 a2[2] = { P.w2f[0][l], P.w2f[1][l] };
    const f32x16 b1 = lane_vec(P.b1, g), b2 = lane_vec(P.b2, g), w3 = lane_vec(P.w3, g);
    float f[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f32x16 acc = b1;           // the bias enters as the accumulator's initial value
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 bx = as_bf(sel4(m == g, own[kc], recv[kc]));
            if (kTape) tape->xb[m][kc] = bx;
            acc = mfma32(a1[kc], bx, acc);
        }
        f32x16 h;
#pragma unroll
        for (int r = 0; r < 16; ++r) h[r] = silu(acc[r]);
        if (kTape) tape->z1[m] = acc;
        f32x16 acc2 = b2;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 hb = chunk_bf16(h, kc);
            if (kTape) tape->h1b[m][kc] = hb;
            acc2 = mfma32(a2[kc], hb, acc2);
        }
        float part = 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) part += w3[r] * silu(acc2[r]);
        if (kTape) tape->z2[m] = acc2;
        f[m] = part + __shfl_xor(part, 32, 64) + P.b3;
    }
    return g == 0 ? f[0] : f[1];
}

__device__ __forceinline__ float clamp_log_radius(float f) { return fminf(fmaxf(f, kLogRMin), kLogRMax); }

// ---------------------------------------------------------------------------
// Packing of the master parameters into MlpDev (one workgroup of 64 lanes).
__global__ void k_mlp_pack(const float* __restrict__ prm, MlpDev* __restrict__ P)
{
    const int l = threadIdx.x, g = l >> 5, i = l & 31;
    if (l >= 64) return;
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        bf16x8 a, b, c;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 16 * kc + 8 * g + j;
            a[j] = (__bf16)(k < kMlpIn ? prm[kOffW1 + i * kMlpIn + k] : 0.f);
            const int u = perm_unit(kc, g, j);
            b[j] = (__bf16)prm[kOffW2 + i * kMlpH + u];
            c[j] = (__bf16)prm[kOffW2 + u * kMlpH + i];
        }
        P->w1f[kc][l] = a;
        P->w2f[kc][l] = b;
        P->w2tf[kc][l] = c;
    }
    if (l < kMlpH) {
        P->b1[l] = prm[kOffB1 + l];
        P->b2[l] = prm[kOffB2 + l];
        P->w3[l] = (float)(__bf16)prm[kOffW3 + l];
    }
    if (l == 0) P->b3 = prm[kOffB3];
}

// Unit-ball vector of the training objective (record rec, step).
__device__ __forceinline__ void unit_ball(uint32_t seed_lo, uint32_t seed_hi, uint32_t rec, uint32_t step, float u[3])
{
    const u32x4 p = philox4x32_10(u32x4{ rec, step, kStreamTrain, seed_hi }, seed_lo, kTrainKey);
    const float ct = s24f(p.x);
    const float up = u24f(p.y);
    const float rad = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(u24f(p.z)) * (1.f / 3.f));
    const float st = sqrtf(fmaxf(0.f, (1.f - ct) * (1.f + ct)));
    u[0] = rad * st * __builtin_amdgcn_cosf(up);
    u[1] = rad * st * __builtin_amdgcn_sinf(up);
    u[2] = rad * ct;
}

} // namespace ising

namespace ising {

// ---------------------------------------------------------------------------
// Training. A record is 32 floats: x[0..20], S[21..23], h[24..26], beta[27],
// K[28]. Loss of a record with unit-ball vector u and r = exp(clamp(f)):
// S' = (S + r u)/|S + r u|, dE = -(S'-S).h - K (S'z-Sz)(S'z+Sz),
// a = min(1, exp(-beta dE)), loss = -|S'-S|^2 a. Returns d loss / d f.
constexpr int kRecord = 32;

__device__ __forceinline__ float record_loss(const float* __restrict__ rc, const float u[3], float f, float* dldf)
{
    const bool inside = f > kLogRMin && f < kLogRMax;
    const float r = __builtin_amdgcn_exp2f(clamp_log_radius(f) * 1.4426950408889634f);
    const float S[3] = { rc[21], rc[22], rc[23] };
    const float h[3] = { rc[24], rc[25], rc[26] };
    const float beta = rc[27], K = rc[28];
    float t[3], sn[3], del[3], ds[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = S[k] + r * u[k];
    const float nn = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const float inv = 1.f / nn;
#pragma unroll
    for (int k = 0; k < 3; ++k) { sn[k] = t[k] * inv; del[k] = sn[k] - S[k]; }
    const float dE = -(del[0] * h[0] + del[1] * h[1] + del[2] * h[2]) - K * del[2] * (sn[2] + S[2]);
    const float a = dE <= 0.f ? 1.f : __builtin_amdgcn_exp2f(-beta * dE * 1.4426950408889634f);
    const float d2 = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
    const float su = sn[0] * u[0] + sn[1] * u[1] + sn[2] * u[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) ds[k] = (u[k] - sn[k] * su) * inv;
    const float dd2 = 2.f * (del[0] * ds[0] + del[1] * ds[1] + del[2] * ds[2]);
    const float ddE = -(h[0] * ds[0] + h[1] * ds[1] + h[2] * ds[2]) - 2.f * K * sn[2] * ds[2];
    const float da = dE > 0.f ? -beta * a * ddE : 0.f;
    *dldf = inside ? -(dd2 * a + d2 * da) * r : 0.f;
    return -d2 * a;
}

__device__ __forceinline__ void wave_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the 32 lanes of each half-wavefront.
__device__ __forceinline__ float half_sum(float v)
{
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Gradient of the mean loss over `batch` records for one optimiser step.
// Record of batch position b: b itself (hashed = 0) or a multiplicative hash
// of (step, b) over the n_recs records. Each wavefront writes its partial
// gradient (and loss sum, last entry) to wave_grad[wave][kMlpParams + 1].
// Weight gradients are bfloat16 MFMA products of tiles transposed in LDS.
__global__ __launch_bounds__(256) void k_mlp_grad(
    const float* __restrict__ recs, uint32_t n_recs, uint32_t batch, uint32_t step, int hashed,
    const MlpDev* __restrict__ Pp, uint32_t seed_lo, uint32_t seed_hi, float* __restrict__ wave_grad)
{
    __shared__ __attribute__((aligned(16))) __bf16 T[4][4][kMlpH][64];   // [wave][dz2,h1,dz1,x][unit][site]
    const MlpDev& P = *Pp;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 5, i = l & 31;
    const int gwave = blockIdx.x * 4 + w, n_waves = gridDim.x * 4;
    const f32x16 w3 = lane_vec(P.w3, g);
    f32x16 gW1 = {}, gW2 = {}, gb1 = {}, gb2 = {}, gw3 = {};
    float gb3 = 0.f, lsum = 0.f;
    const float inv_batch = 1.f / (float)batch;
    for (uint32_t tile = gwave; tile * 64u < batch; tile += n_waves) {
        const uint32_t b = tile * 64u + (uint32_t)l;
        const bool valid = b < batch;
        uint32_t ri = 0;
        if (valid) ri = hashed ? (uint32_t)(((uint64_t)step * batch + b) * 2654435761ull % n_recs) : b;
        const float* rc = recs + (size_t)ri * kRecord;
        float x[kMlpIn];
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) x[k] = valid ? rc[k] : 0.f;
        MlpTape tp;
        const float f = mlp_forward<true>(P, x, &tp);
        float dldf = 0.f, loss = 0.f;
        if (valid) {
            float u[3];
            unit_ball(seed_lo, seed_hi, ri, step, u);
            loss = record_loss(rc, u, f, &dldf);
        }
        const float gown = dldf * inv_batch;
        lsum += loss;
        gb3 += gown;
        const float gpart = __shfl_xor(gown, 32, 64);
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const float gm = (m == g) ? gown : gpart;
            f32x16 dz2;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float z = tp.z2[m][r];
                gw3[r] += gm * silu(z);
                dz2[r] = gm * w3[r] * dsilu(z);
                gb2[r] += dz2[r];
            }
            f32x16 dh1 = {};
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) dh1 = mfma32(P.w2tf[kc][l], chunk_bf16(dz2, kc), dh1);
            f32x16 dz1;
#pragma unroll
            for (int r = 0; r < 16; ++r) { dz1[r] = dh1[r] * dsilu(tp.z1[m][r]); gb1[r] += dz1[r]; }
            const int site = 32 * m + i;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                T[w][0][acc_unit(r, g)][site] = (__bf16)dz2[r];
                T[w][2][acc_unit(r, g)][site] = (__bf16)dz1[r];
            }
#pragma unroll
            for (int kc = 0; kc < 2; ++kc)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    T[w][1][perm_unit(kc, g, j)][site] = tp.h1b[m][kc][j];
                    T[w][3][16 * kc + 8 * g + j][site] = tp.xb[m][kc][j];
                }
        }
        wave_fence();
        // dW2[u_out][u_in] += sum_s dz2[s][u_out] h1[s][u_in]; dW1 likewise.
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            const int s0 = 16 * kc + 8 * g;
            const bf16x8 a2 = *(const bf16x8*)&T[w][0][i][s0];
            const bf16x8 b2 = *(const bf16x8*)&T[w][1][i][s0];
            gW2 = mfma32(a2, b2, gW2);
            const bf16x8 a1 = *(const bf16x8*)&T[w][2][i][s0];
            const bf16x8 b1v = *(const bf16x8*)&T[w][3][i][s0];
            gW1 = mfma32(a1, b1v, gW1);
        }
        wave_fence();
    }
    float* out = wave_grad + (size_t)gwave * (kMlpParams + 1);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int u = acc_unit(r, g);
        out[kOffW2 + u * kMlpH + i] = gW2[r];
        if (i < kMlpIn) out[kOffW1 + u * kMlpIn + i] = gW1[r];
        const float s1 = half_sum(gb1[r]), s2 = half_sum(gb2[r]), s3 = half_sum(gw3[r]);
        if (i == 0) { out[kOffB1 + u] = s1; out[kOffB2 + u] = s2; out[kOffW3 + u] = s3; }
    }
    float t3 = gb3, tl = lsum;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { t3 += __shfl_xor(t3, o, 64); tl += __shfl_xor(tl, o, 64); }
    if (l == 0) { out[kOffB3] = t3; out[kMlpParams] = tl; }
}

// Deterministic sum of the per-wavefront gradients (and loss, entry
// kMlpParams) in two fixed-order stages: group g = blockIdx.y of
// `per_group` consecutive wavefronts into part[g][p], then the groups in
// order into grad[p].
constexpr int kMlpReduceGroups = 32;

__global__ void k_mlp_reduce_groups(const float* __restrict__ wave_grad, int n_waves, int per_group,
                                    float* __restrict__ part)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    const int w0 = blockIdx.y * per_group, w1 = min(w0 + per_group, n_waves);
    float s = 0.f;
#pragma unroll 8
    for (int w = w0; w < w1; ++w) s += wave_grad[(size_t)w * (kMlpParams + 1) + p];
    part[(size_t)blockIdx.y * (kMlpParams + 1) + p] = s;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int n_groups, float* __restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.f;
    for (int g = 0; g < n_groups; ++g) s += part[(size_t)g * (kMlpParams + 1) + p];
    grad[p] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float bc1, float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gr = grad[p];
    const float a = 0.9f * m1[p] + 0.1f * gr;
    const float v = 0.999f * m2[p] + 0.001f * gr * gr;
    m1[p] = a;
    m2[p] = v;
    prm[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// Test hook: f for rows of 21 features (one wavefront per 64 rows).
__global__ __launch_bounds__(256) void k_mlp_forward_rows(const float* __restrict__ xs, int n, const MlpDev* __restrict__ P,
                                                          float* __restrict__ out)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = row < n;
    float x[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) x[k] = ok ? xs[(size_t)row * kMlpIn + k] : 0.f;
    const float f = mlp_forward<false>(*P, x, nullptr);
    if (ok) out[row] = f;
}

// Test hook: D (32x32) = A (32x16) B (16x32) through one MFMA, operands
// loaded with the lane layout the kernels above assume.
__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D)
{
    const int l = threadIdx.x, g = l >> 5, i = l & 31;
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = (__bf16)A[i * 16 + 8 * g + j];
        b[j] = (__bf16)B[(8 * g + j) * 32 + i];
    }
    const f32x16 d = mfma32(a, b, f32x16{});
#pragma unroll
    for (int r = 0; r < 16; ++r) D[acc_unit(r, g) * 32 + i] = d[r];
}

} // namespace ising
