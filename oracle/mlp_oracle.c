/*
 * mlp_oracle.c -- FP64 CPU oracle of the learned cone-width network and of
 * its training objective. TEST INFRASTRUCTURE ONLY (tests/ and bench.py's
 * checker); the product's kernels live in ising_amd/csrc/mlp.hpp.
 *
 * The network has no reference counterpart (the reference samples with a
 * fixed radius, src/energyMin.f90:49-68, :326-329). Semantics fixed here:
 *
 *   input x[21]: the six neighbour spins (+x, -x, +y, -y, +z, -z; 3 each)
 *                followed by beta*h, the reduced local field (3);
 *   z1 = W1 x + b1,  h1 = g(z1)           W1: 32 x 21
 *   z2 = W2 h1 + b2, h2 = g(z2)           W2: 32 x 32
 *   f  = w3 . h2 + b3                     log of the cone radius
 *   r  = exp(clamp(f, log_rmin, log_rmax))
 * with g(z) = z / (1 + 2^-z), SiLU in base 2 (= log2(e) silu(z ln 2), the
 * same function family as SiLU with weights rescaled by ln 2).
 *
 * Parameter vector (1793): W1 row-major [32][21], b1[32], W2 row-major
 * [32][32], b2[32], w3[32], b3.
 *
 * With quantize = 1 the oracle rounds exactly the operands that the device
 * kernels round to bfloat16 (round to nearest even): x, W1, h1, W2, w3 in the
 * forward pass; dz2, dz1, h1, x in the backward products. Everything else is
 * FP64. quantize = 0 is the plain FP64 network (for finite differences).
 *
 * Training objective per record (site spin S, field h, beta, K, a unit-ball
 * vector u): S' = (S + r u)/|S + r u|, dE = -(S'-S).h - K (S'z-Sz)(S'z+Sz),
 * a = min(1, exp(-beta dE)), loss = -|S'-S|^2 a (negative expected squared
 * jump distance); the batch loss is the mean.
 *
 * Pinning (tests/test_oracle_mlp.py): the gradient by central finite
 * differences of the loss for every parameter group; the forward pass by
 * special cases with a closed form (zero weights give f = b3; a single active
 * path gives f = w3 silu(W2 silu(W1 x + b1) + b2) + b3 for one unit); the
 * objective by its limits (r -> 0 gives loss 0; beta = 0 gives a = 1 and the
 * loss is minus the squared chord length).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ising_oracle.h"

#define NIN 21
#define NH 32
#define NPARAM (NH * NIN + NH + NH * NH + NH + NH + 1)

static double bf16q(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double q(double v, int quantize) { return quantize ? bf16q(v) : v; }
static double sigm(double z) { return 1.0 / (1.0 + exp2(-z)); }

typedef struct {
    const double *W1, *b1, *W2, *b2, *w3;
    double b3;
} net_t;

static net_t unpack(const double* p)
{
    net_t n;
    n.W1 = p;
    n.b1 = n.W1 + NH * NIN;
    n.W2 = n.b1 + NH;
    n.b2 = n.W2 + NH * NH;
    n.w3 = n.b2 + NH;
    n.b3 = n.w3[NH];
    return n;
}

typedef struct { double xq[NIN], z1[NH], h1[NH], h1q[NH], z2[NH], h2[NH], f; } fwd_t;

static void forward(const net_t* n, const double* x, int quantize, fwd_t* o)
{
    for (int k = 0; k < NIN; ++k) o->xq[k] = q(x[k], quantize);
    for (int u = 0; u < NH; ++u) {
        double s = 0.0;
        for (int k = 0; k < NIN; ++k) s += q(n->W1[u * NIN + k], quantize) * o->xq[k];
        o->z1[u] = s + n->b1[u];
        o->h1[u] = o->z1[u] * sigm(o->z1[u]);
        o->h1q[u] = q(o->h1[u], quantize);
    }
    double f = n->b3;
    for (int u = 0; u < NH; ++u) {
        double s = 0.0;
        for (int k = 0; k < NH; ++k) s += q(n->W2[u * NH + k], quantize) * o->h1q[k];
        o->z2[u] = s + n->b2[u];
        o->h2[u] = o->z2[u] * sigm(o->z2[u]);
        f += q(n->w3[u], quantize) * o->h2[u];
    }
    o->f = f;
}

double orc_mlp_forward(const double* params, const double* x, int quantize)
{
    net_t n = unpack(params);
    fwd_t o;
    forward(&n, x, quantize, &o);
    return o.f;
}

static double dsilu(double z)
{
    const double s = sigm(z);
    return s * (1.0 + z * M_LN2 * (1.0 - s));
}

/* Loss of one record and d loss / d f. rec: x[0..20], S[21..23], h[24..26],
 * beta[27], K[28]. */
static double record_loss(const double* rec, const double* uvec, double f, double lo, double hi, double* dldf)
{
    const double* S = rec + 21;
    const double* h = rec + 24;
    const double beta = rec[27], K = rec[28];
    const int inside = (f > lo && f < hi);
    const double lr = f < lo ? lo : (f > hi ? hi : f);
    const double r = exp(lr);
    double t[3], sn[3], del[3], ds[3];
    for (int k = 0; k < 3; ++k) t[k] = S[k] + r * uvec[k];
    const double nn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; ++k) { sn[k] = t[k] / nn; del[k] = sn[k] - S[k]; }
    const double dE = -(del[0] * h[0] + del[1] * h[1] + del[2] * h[2]) - K * del[2] * (sn[2] + S[2]);
    const double a = dE <= 0.0 ? 1.0 : exp(-beta * dE);
    const double d2 = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
    const double su = sn[0] * uvec[0] + sn[1] * uvec[1] + sn[2] * uvec[2];
    for (int k = 0; k < 3; ++k) ds[k] = (uvec[k] - sn[k] * su) / nn;
    const double dd2 = 2.0 * (del[0] * ds[0] + del[1] * ds[1] + del[2] * ds[2]);
    const double ddE = -(h[0] * ds[0] + h[1] * ds[1] + h[2] * ds[2]) - 2.0 * K * sn[2] * ds[2];
    const double da = dE > 0.0 ? -beta * a * ddE : 0.0;
    const double dldr = -(dd2 * a + d2 * da);
    *dldf = inside ? dldr * r : 0.0;
    return -d2 * a;
}

/* Mean loss over n records and its gradient (mean) w.r.t. all parameters. */
double orc_mlp_loss_grad(const double* params, const double* recs, const double* uvecs, int n, int quantize,
                         double log_rmin, double log_rmax, double* grad)
{
    net_t nt = unpack(params);
    memset(grad, 0, sizeof(double) * NPARAM);
    double* gW1 = grad;
    double* gb1 = gW1 + NH * NIN;
    double* gW2 = gb1 + NH;
    double* gb2 = gW2 + NH * NH;
    double* gw3 = gb2 + NH;
    double* gb3 = gw3 + NH;
    double total = 0.0;
    for (int i = 0; i < n; ++i) {
        const double* rec = recs + 32 * (size_t)i;
        fwd_t o;
        forward(&nt, rec, quantize, &o);
        double g;
        total += record_loss(rec, uvecs + 3 * (size_t)i, o.f, log_rmin, log_rmax, &g);
        g /= (double)n;
        double dz2[NH], dz2q[NH], dz1[NH];
        *gb3 += g;
        for (int u = 0; u < NH; ++u) {
            gw3[u] += g * o.h2[u];
            dz2[u] = g * q(nt.w3[u], quantize) * dsilu(o.z2[u]);
            dz2q[u] = q(dz2[u], quantize);
            gb2[u] += dz2[u];
        }
        for (int u = 0; u < NH; ++u)
            for (int k = 0; k < NH; ++k) gW2[u * NH + k] += dz2q[u] * o.h1q[k];
        for (int k = 0; k < NH; ++k) {
            double s = 0.0;
            for (int u = 0; u < NH; ++u) s += q(nt.W2[u * NH + k], quantize) * dz2q[u];
            dz1[k] = s * dsilu(o.z1[k]);
            gb1[k] += dz1[k];
        }
        for (int u = 0; u < NH; ++u) {
            const double dq = q(dz1[u], quantize);
            for (int k = 0; k < NIN; ++k) gW1[u * NIN + k] += dq * o.xq[k];
        }
    }
    return total / (double)n;
}

/* Unit-ball vector of the training objective for record `rec` at training
 * step `step`: Philox key (seed_lo, 0x6D6C7000), counter (rec, step, 3,
 * seed_hi); cos(theta) = 2U - 1, phi = 2 pi U', radius U''^(1/3). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

void orc_mlp_unit_ball(uint64_t seed, uint32_t rec, uint32_t step, double u[3])
{
    uint32_t ctr[4] = { rec, step, 3u, (uint32_t)(seed >> 32) };
    uint32_t key[2] = { (uint32_t)seed, 0x6D6C7000u }, r[4];
    orc_philox4x32_10(ctr, key, r);
    const double ct = (double)(r[0] >> 8) * 0x1p-23 - 1.0;
    const double ph = 2.0 * M_PI * (double)(r[1] >> 8) * 0x1p-24;
    const double rad = cbrt((double)(r[2] >> 8) * 0x1p-24);
    const double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    u[0] = rad * st * cos(ph);
    u[1] = rad * st * sin(ph);
    u[2] = rad * ct;
}

int orc_mlp_nparam(void) { return NPARAM; }

/* Sweeps with the learned radius (build form; the reference has none): each
 * colour of the checkerboard in turn, every site of that colour one proposal
 * with radius r = exp(clamp(f(x), log_rmin, log_rmax)), x = the six
 * neighbour spins (+x, -x, +y, -y, +z, -z; zero across an open face) and
 * beta*h (h the local field of orc_local_field: exchange, DMI and Zeeman,
 * no anisotropy), then the Metropolis test of orc_sweeps with the same
 * Philox counters. The radius never reads the site's own spin, so the
 * proposal is symmetric. Simple-cubic lattices only (-1 otherwise).
 *
 * Pinned (tests/test_oracle_mlp.py) by: a network with zero weights and
 * b3 = ln r repeats orc_sweeps with radius exp(b3) bit for bit; with an
 * arbitrary network a paramagnet keeps the Langevin law <Sz> =
 * coth(beta B) - 1/(beta B) (any symmetric proposal does; one that read the
 * site's own spin would not). */
int64_t orc_sweeps_learned(const orc_lattice* L, const orc_ham* h, double* s, double beta, const double* params,
                           int quantize, double log_rmin, double log_rmax, uint64_t seed, uint32_t slot,
                           uint32_t sweep0, int n_sweeps)
{
    if (L->cell != ORC_CELL_SC) return -1;
    const int64_t N = orc_nsites(L);
    net_t nt = unpack(params);
    int64_t acc = 0;
    for (int k = 0; k < n_sweeps; ++k) {
        for (int c = 0; c < 2; ++c) {
            #pragma omp parallel for reduction(+:acc) schedule(static)
            for (int64_t i = 0; i < N; ++i) {
                if (orc_site_colour(L, i) != c) continue;
                double x[NIN];
                int f = 0;
                for (int axis = 0; axis < 3; ++axis)
                    for (int d = 1; d >= -1; d -= 2) {
                        const int64_t j = orc_neighbour(L, i, axis, d);
                        for (int m = 0; m < 3; ++m) x[f + m] = j >= 0 ? s[m * N + j] : 0.0;
                        f += 3;
                    }
                double hf[3];
                orc_local_field(L, h, s, i, hf);
                for (int m = 0; m < 3; ++m) x[18 + m] = beta * hf[m];
                fwd_t o;
                forward(&nt, x, quantize, &o);
                const double lr = o.f < log_rmin ? log_rmin : (o.f > log_rmax ? log_rmax : o.f);
                uint32_t rnd[4];
                orc_metropolis_random(seed, slot, (uint32_t)i, sweep0 + (uint32_t)k, rnd);
                double so[3] = { s[i], s[N + i], s[2 * N + i] }, sn[3], u;
                if (!orc_propose(so, exp(lr), rnd, sn, &u)) continue;
                const double de = orc_delta_energy(L, h, s, i, sn);
                if (u < exp(-beta * de)) {
                    s[i] = sn[0]; s[N + i] = sn[1]; s[2 * N + i] = sn[2];
                    acc++;
                }
            }
        }
    }
    return acc;
}
