// oc_common.hpp -- structures shared by the gfx950 kernels and the host side
// of the C ABI.  All arithmetic that must be bit-identical to the FP64
// restatement of cel/engine/oercel.c (integration, cell keys, chemistry
// distances, bonded forces) is written as plain expressions; the library is
// compiled with -ffp-contract=off so none of them is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/oercel_mi.h"

namespace oc {

constexpr int kGidBits = 29;                      // gt = gid | type << 29 (N < 2^29 per replica)
constexpr uint32_t kGidMask = (1u << kGidBits) - 1u;
constexpr int kBeadBlock = 256;                   // beads per block of the per-bead kernels
constexpr int kRedChunk = 2048;                   // beads per partial of the stats reduction

// sticky device error bits
enum : uint32_t { ERR_EVOVF = 1u, ERR_CHAIN = 2u, ERR_SORT = 4u, ERR_TILE = 8u, ERR_CATCAP = 16u, ERR_DISP = 32u,
                  ERR_INJ = 64u };

// Replica constants (one system of the ensemble).
struct RepInfo {
    int32_t N, off;           // beads, first global slot
    int32_t ncx, ncy, ncell, coff;  // cells (column-major), first global cell
    int32_t mlen, nH;
    int32_t chem, n_chem;
    int32_t has_stir, bblk0, rblk0, col0;  // stirring on; first bead / reduction block; columns of earlier replicas
    double L, halfL, icsx, icsy, csx, csy;
    double dt, dt2, cvt, hdt;       // dt, dt^2, (1/2 + lambda) dt, dt/2
    double gamma, kT, kb, r0, ka;
    double stir_A, stir_kn, kT_amp, kT_period;
    double t_anchor;
    int64_t step_anchor, warmup, warm_end;
    uint64_t seed;
    double p_conv, rc2, p_form, rf2, p_break;
    double rwc, rwf;  // half-heights of the chemistry search windows (r_conv, r_form + 1e-6)
    // the reference chemistry without catalysts: decay and food supply
    int32_t rchem, feed_mode, vent_shape, rpad;
    double p_ld, p_nd, p_feed, vent_x, vent_y, vent_r2, vent_hw;
    double a[64];
    // lipid template of a branched architecture (arch != 0): bonds (bi, bj,
    // rest br0) and angle terms (ai, aj = middle, ak; kind 0: K (1 - cos),
    // kind 1: K (cos - c0)^2), in template order
    int32_t arch, tnb, tna, tpad;
    int32_t bi[2 * OC_MAXMLEN], bj[2 * OC_MAXMLEN];
    int32_t ai[2 * OC_MAXMLEN], aj[2 * OC_MAXMLEN], ak[2 * OC_MAXMLEN], akind[2 * OC_MAXMLEN];
    double br0[2 * OC_MAXMLEN], aK[2 * OC_MAXMLEN], ac0[2 * OC_MAXMLEN];
    // catalyst chemistry (the reference GPU port's form) and the feed cap
    int32_t react, syn_tries, syn_relax, nsynN;  // react: some p_syn / p_rep > 0
    double p_syn[2], p_rep[2], mu, p_kd, r2cat;
    // windows (cells) of the r_cat search, the overlap energy and the
    // relaxation neighbours; two catalysts closer than (cfx, cfy) cells may
    // touch each other's beads
    int32_t rwx, rwy, owx, owy, lwx, lwy, cfx, cfy;
    int32_t fcap_on, fcap_n;  // feed cap on; floor(feed_cap N + 1e-9)
    int32_t rdx, rdy;         // cells (from the catalyst's) that hold every bead a reaction reads
    int32_t stile0, ntc;      // first sort tile of the replica, sort tiles per column
    double rg2;               // squared radius around the catalyst of every bead a reaction reads
    double ro2;               // squared radius around the catalyst of every overlap-energy partner
    double tpx[OC_MAXMLEN], tpy[OC_MAXMLEN];  // template coordinates (centroid 0), branched lipids
};

// Per-step values of one replica, written on the device by k_step_params.
struct RepStep {
    int64_t step;             // step whose forces are (being) computed
    double t;
    uint64_t sseed;           // pair-noise key of this step
    uint64_t kconv, kform, kbreak;  // chemistry keys
    uint64_t kld, knd, kfeed;       // reference chemistry keys (salts 11, 13, 14)
    uint64_t kkd, ksyn, krep, kmu, kpri, kcap;  // salts 12, 21, 22, 23, 24, 25
    uint64_t disp;            // the step's integrate moved some bead by the chemistry margin 0.5 or more
    uint64_t feed_thr;        // feed cap: largest kept key
    int32_t feed_add, rpad;   // feed cap: conversions kept (0: none)
    float a_eff[64];          // repulsion at this step's warm-up ramp
    float gamma, sdt;         // gamma, sigma(t)/sqrt(dt)
    int32_t Rx, Ry;           // sort window radii (cells) of the last integrate
    int32_t chem_now;         // precursor chemistry check at this step
    int32_t rchem_now;        // reference chemistry (decay, feed) check at this step
    double stir_mean;
    uint64_t maxbond_bits;    // running max bond length (bits of a positive double)
};

static_assert(offsetof(RepStep, kcap) - offsetof(RepStep, sseed) == 12 * sizeof(uint64_t),
              "the step keys sseed .. kcap are consecutive");

struct Event {
    int64_t step;
    int32_t kind;
    uint32_t gid_a, gid_b;
    int32_t rep;
    double x, y;
};

// Pair tile: replica, column, sub-cell rows [k0, k1).
struct Tile {
    int32_t rep, cx, k0, k1;
};

// --------------------------------------------------------------------------- hashes
// splitmix64 finaliser and the keys of cel/engine/oercel.c:123-141, 819-826
__host__ __device__ inline uint64_t mix13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t step_seed(uint64_t seed, int64_t step) {
    return mix13(mix13(seed ^ 0x5851f42d4c957f2dULL) ^ mix13((uint64_t)step + 0x632be59bd9b4e019ULL));
}
__host__ __device__ inline uint64_t hash_key(uint64_t seed, int64_t step, uint64_t salt) {
    return mix13(seed ^ mix13((uint64_t)step * 0x9e3779b97f4a7c15ULL + salt));
}
__host__ __device__ inline double hash_u01(uint64_t key, uint32_t g) {
    return (double)(mix13(key ^ ((uint64_t)g * 0xd1b54a32d192ed03ULL)) >> 11) * 0x1.0p-53;
}
__host__ __device__ inline uint64_t pair_hash(uint64_t key, uint32_t gi, uint32_t gj) {
    const uint32_t lo = gi < gj ? gi : gj, hi = gi < gj ? gj : gi;
    const uint64_t k = ((uint64_t)lo << 32) | hi;
    return mix13((k * 0x9e3779b97f4a7c15ULL) ^ key);
}
__host__ __device__ inline double pair_u01(uint64_t key, uint32_t gi, uint32_t gj) {
    return (double)(pair_hash(key, gi, gj) >> 11) * 0x1.0p-53;
}

// chemistry salts (distinct from the reference's 11-14)
enum : uint64_t { SALT_CONV = 31, SALT_FORM = 32, SALT_BREAK = 33 };

__host__ __device__ inline double minimg(double d, double L, double h) {
    if (d > h) d -= L;
    else if (d < -h) d += L;
    return d;
}

}  // namespace oc
