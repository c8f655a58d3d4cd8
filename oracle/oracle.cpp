/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY.  CPU (FP64) restatement of the 2D DPD
 * step of the reference engine cel/engine/oercel.c, used as the checker of the
 * HIP library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product never does.
 *
 * Every function cites the reference lines it restates.  Parts with no
 * reference counterpart (the precursor chemistry, the energy, the full-shell
 * traversal order) are this project's own definitions and say so.
 *
 * Pinning (what checks this file, see tests/):
 *   hashes, soup init, cell keys + stable sort : bit-exact against golden
 *       vectors written by the reference itself (tests/golden), and the
 *       feed known-answer count 1138 (cel/engine/README.md:519-520);
 *   pair forces   : golden reference forces (<= 1e-10 rel), O(N^2) brute force,
 *       sum F = 0, noise moments;
 *   bonded forces : golden reference forces, central finite differences of an
 *       independently written energy;
 *   integrator + full step : golden reference trajectories (20 steps);
 *   reductions    : closed forms on synthetic inputs;
 *   precursor chemistry (or_chem) : PARITY UNPINNED against the reference (it has
 *       no such chemistry); pinned only by brute-force neighbour search and
 *       binomial rate checks in the tests.
 *
 * Built with -ffp-contract=off so that the arithmetic is exactly the written
 * expression order (the HIP library is built the same way for its FP64 parts).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../include/oercel_mi.h"

#define OR_API extern "C" __attribute__((visibility("default")))

static const double SQRT12 = 3.4641016151377545870548926830117;

/* ------------------------------------------------------------------ hashes */

/* splitmix64 finaliser, cel/engine/oercel.c:123-128 */
static inline uint64_t mix13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
/* per-step pair-noise key, cel/engine/oercel.c:130-133 */
static inline uint64_t step_seed(uint64_t seed, int64_t step) {
    return mix13(mix13(seed ^ 0x5851f42d4c957f2dULL) ^ mix13((uint64_t)step + 0x632be59bd9b4e019ULL));
}
static inline uint64_t pair_hash(uint64_t key, uint32_t gi, uint32_t gj) {
    const uint64_t k = gi < gj ? (((uint64_t)gi << 32) | gj) : (((uint64_t)gj << 32) | gi);
    return mix13((k * 0x9e3779b97f4a7c15ULL) ^ key);
}
/* symmetric pair noise on [-sqrt3, sqrt3), cel/engine/oercel.c:135-141 */
static inline double pair_theta(uint64_t sseed, uint32_t gi, uint32_t gj) {
    return (double)(int64_t)pair_hash(sseed, gi, gj) * (SQRT12 / 18446744073709551616.0);
}
/* the same deviate from the top 32 bits of the hash: the form the device
 * computes in FP32 (cel/engine_gpu/oercel_gpu.cu:246-251 uses the same cut) */
static inline double pair_theta32(uint64_t sseed, uint32_t gi, uint32_t gj) {
    return (double)(int32_t)(uint32_t)(pair_hash(sseed, gi, gj) >> 32) * (SQRT12 / 4294967296.0);
}
/* counter-based per-bead uniform, cel/engine/oercel.c:819-826 */
static inline uint64_t hash_key(uint64_t seed, int64_t step, uint64_t salt) {
    return mix13(seed ^ mix13((uint64_t)step * 0x9e3779b97f4a7c15ULL + salt));
}
static inline double hash_u01(uint64_t key, uint32_t g) {
    return (double)(mix13(key ^ ((uint64_t)g * 0xd1b54a32d192ed03ULL)) >> 11) * 0x1.0p-53;
}
/* the full 64-bit per-bead hash (the GPU port's hash_u64,
 * cel/engine_gpu/oercel_gpu.cu:259): priorities and feed-cap keys */
static inline uint64_t hash_u64(uint64_t key, uint32_t g) { return mix13(key ^ ((uint64_t)g * 0xd1b54a32d192ed03ULL)); }
/* symmetric per-pair uniform for bond formation (own definition; same hash
 * shape as pair_theta, keyed by a chemistry salt) */
static inline double pair_u01(uint64_t key, uint32_t gi, uint32_t gj) {
    return (double)(pair_hash(key, gi, gj) >> 11) * 0x1.0p-53;
}

/* chemistry salts (own; the reference uses 11-14, its GPU port 21-25 and 40+) */
enum { SALT_CONV = 31, SALT_FORM = 32, SALT_BREAK = 33 };

OR_API uint64_t or_mix13(uint64_t z) { return mix13(z); }
OR_API uint64_t or_step_seed(uint64_t seed, int64_t step) { return step_seed(seed, step); }
OR_API double or_pair_theta(uint64_t s, uint32_t gi, uint32_t gj) { return pair_theta(s, gi, gj); }
OR_API double or_pair_theta32(uint64_t s, uint32_t gi, uint32_t gj) { return pair_theta32(s, gi, gj); }
OR_API uint64_t or_hash_key(uint64_t seed, int64_t step, uint64_t salt) { return hash_key(seed, step, salt); }
OR_API double or_hash_u01(uint64_t key, uint32_t g) { return hash_u01(key, g); }
OR_API double or_pair_u01(uint64_t key, uint32_t gi, uint32_t gj) { return pair_u01(key, gi, gj); }

/* Known-answer count of the uniform food supply on pure water
 * (cel/engine/oercel.c:1713-1753 with feed_cap = 0 and no decay): a bead
 * converts W -> N at the first check whose hash_u01(hash_key(seed, step, 14),
 * gid) < p_feed, so the total is the number of gids with any success. */
OR_API int64_t or_feed_count(uint64_t seed, int32_t N, double p_feed, int64_t n_chem, int64_t steps) {
    std::vector<uint8_t> done(N, 0);
    int64_t tot = 0;
    for (int64_t st = n_chem; st <= steps; st += n_chem) {
        const uint64_t kf = hash_key(seed, st, 14);
        for (int32_t g = 0; g < N; g++)
            if (!done[g] && hash_u01(kf, (uint32_t)g) < p_feed) { done[g] = 1; tot++; }
    }
    return tot;
}

/* ------------------------------------------------------------------ serial RNG (init only) */

/* xoshiro256**, splitmix seeding, Box-Muller: cel/engine/oercel.c:143-174 */
struct Rng { uint64_t s[4]; };
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t rng_next(Rng *r) {
    uint64_t *s = r->s;
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
static void rng_seed(Rng *r, uint64_t seed) {
    uint64_t z = seed;
    for (int k = 0; k < 4; k++) {
        z += 0x9e3779b97f4a7c15ULL;
        r->s[k] = mix13(z);
    }
}
static double rng_unif(Rng *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static double rng_normal(Rng *r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ------------------------------------------------------------------ geometry */

struct Geo {
    int N, ncx, ncy, ncell;
    double L, icsx, icsy, csx, csy;
};
/* cells: ncx = floor(L) columns, ncy = floor(L)*nsub sub-cells, column-major,
 * cel/engine/oercel.c:739-745 */
static Geo make_geo(const oc_params *p, int N) {
    Geo g;
    g.N = N;
    g.L = p->L;
    g.ncx = (int)floor(p->L);
    g.ncy = (int)floor(p->L) * p->nsub;
    g.ncell = g.ncx * g.ncy;
    g.csx = p->L / g.ncx;
    g.csy = p->L / g.ncy;
    g.icsx = g.ncx / p->L;
    g.icsy = g.ncy / p->L;
    return g;
}
/* cel/engine/oercel.c:761-786 */
static inline void wrap1(double *q, double L) {
    double v = *q;
    if (v < 0.0) {
        v += L;
        if (v >= L) v -= L;
    } else if (v >= L) {
        v -= L;
        if (v >= L || v < 0.0) v = fmod(v, L);
    }
    *q = v;
}
static inline double minimg(double d, double L) {
    const double h = 0.5 * L;
    if (d > h) d -= L;
    else if (d < -h) d += L;
    return d;
}
/* the cell key of sort_cells, cel/engine/oercel.c:860-866 */
static inline int cell_key(const Geo &g, double x, double y) {
    int cx = (int)(x * g.icsx), cy = (int)(y * g.icsy);
    cx = cx >= g.ncx ? g.ncx - 1 : cx;
    cy = cy >= g.ncy ? g.ncy - 1 : cy;
    return cx * g.ncy + cy;
}

OR_API void or_cell_keys(const oc_params *p, int32_t N, const double *x, const double *y, int32_t *key) {
    const Geo g = make_geo(p, N);
    for (int i = 0; i < N; i++) key[i] = cell_key(g, x[i], y[i]);
}

/* Stable counting sort by key: src[new slot] = old slot, cstart[ncell+1].
 * The unique stable order that sort_cells produces, cel/engine/oercel.c:828-933. */
OR_API void or_sort(int32_t N, int32_t ncell, const int32_t *key, int32_t *src, int32_t *cstart) {
    std::vector<int32_t> cnt(ncell + 1, 0);
    for (int i = 0; i < N; i++) cnt[key[i]]++;
    int run = 0;
    for (int c = 0; c < ncell; c++) {
        cstart[c] = run;
        run += cnt[c];
    }
    cstart[ncell] = run;
    std::vector<int32_t> fill(cstart, cstart + ncell);
    for (int i = 0; i < N; i++) src[fill[key[i]]++] = i;
}

/* ------------------------------------------------------------------ parameters of one step */

/* warm-up ramp of the repulsion, cel/engine/oercel.c:1122-1129 */
static double a_eff(const oc_params *p, int i, int j, int64_t step) {
    const double a = p->a[i][j];
    if (step >= p->warm_end || p->warmup <= 0 || a <= 25.0) return a;
    const double f = 1.0 - (double)(p->warm_end - step) / (double)p->warmup;
    return 25.0 + (a - 25.0) * (f < 0 ? 0 : f);
}
/* cel/engine/oercel.c:622-624 */
static inline double kT_at(const oc_params *p, double t) {
    return p->kT_amp == 0.0 ? p->kT : p->kT + p->kT_amp * sin(2.0 * M_PI * t / p->kT_period);
}
static inline double time_of(const oc_params *p, int64_t step) {
    return p->t_anchor + (double)(step - p->step_anchor) * p->dt;
}

/* ------------------------------------------------------------------ pair forces */

struct PairPar {
    double a[64];
    double gamma, sdt;
    uint64_t sseed;
};
/* cel/engine/oercel.c:1198-1205 */
static PairPar pair_par(const oc_params *p, int64_t step, double t) {
    PairPar pp;
    memset(&pp, 0, sizeof pp);
    for (int i = 0; i < OC_NTYPES; i++)
        for (int j = 0; j < OC_NTYPES; j++) pp.a[8 * i + j] = a_eff(p, i, j, step);
    pp.gamma = p->gamma;
    pp.sdt = sqrt(2.0 * p->gamma * kT_at(p, t)) / sqrt(p->dt);
    pp.sseed = step_seed(p->seed, step);
    return pp;
}

/* one pair term on bead i from j (dx = x_i - x_j image), the arithmetic of
 * pair_seg, cel/engine/oercel.c:968-979; returns the force on i */
/* Caller-supplied draws, the oracle side of oc_forces_theta / oc_chem_uniforms:
 * per-pair theta and formation uniforms keyed by (min gid << 32 | max gid),
 * per-gid conversion and break uniforms.  While set they replace the hashes
 * in pair_term and or_chem. */
static struct {
    std::vector<uint64_t> tk, fk;
    std::vector<double> tv, fv, uconv, ubreak;
    bool theta = false, chem = false;
    int64_t missing = 0;
} g_inj;

static double inj_lookup(const std::vector<uint64_t> &k, const std::vector<double> &v, uint32_t gi, uint32_t gj,
                         double absent) {
    const uint64_t key = gi < gj ? ((uint64_t)gi << 32) | gj : ((uint64_t)gj << 32) | gi;
    const auto it = std::lower_bound(k.begin(), k.end(), key);
    if (it == k.end() || *it != key) {
#pragma omp atomic
        g_inj.missing++;
        return absent;
    }
    return v[it - k.begin()];
}

static std::vector<uint64_t> inj_keys(const uint32_t *lo, const uint32_t *hi, int64_t n) {
    std::vector<uint64_t> k((size_t)n);
    for (int64_t q = 0; q < n; q++) k[q] = ((uint64_t)lo[q] << 32) | hi[q];
    return k;
}

/* theta table (theta != NULL), chemistry uniforms (uconv != NULL); NULL clears */
OR_API void or_inject(const uint32_t *tlo, const uint32_t *thi, const double *theta, int64_t nt, int32_t N,
                      const double *uconv, const double *ubreak, const uint32_t *flo, const uint32_t *fhi,
                      const double *uform, int64_t nf) {
    g_inj.theta = theta != nullptr;
    g_inj.tk = theta ? inj_keys(tlo, thi, nt) : std::vector<uint64_t>();
    g_inj.tv = theta ? std::vector<double>(theta, theta + nt) : std::vector<double>();
    g_inj.chem = uconv != nullptr;
    g_inj.uconv = uconv ? std::vector<double>(uconv, uconv + N) : std::vector<double>();
    g_inj.ubreak = uconv ? std::vector<double>(ubreak, ubreak + N) : std::vector<double>();
    g_inj.fk = uconv ? inj_keys(flo, fhi, nf) : std::vector<uint64_t>();
    g_inj.fv = uconv ? std::vector<double>(uform, uform + nf) : std::vector<double>();
    g_inj.missing = 0;
}

OR_API int64_t or_inject_missing(void) { return g_inj.missing; }

static inline void pair_term(const PairPar &pp, int theta32, double dx, double dy, double dvx, double dvy, int ti, int tj,
                             uint32_t gi, uint32_t gj, double *fx, double *fy, double *e, double *sabs) {
    const double r2 = dx * dx + dy * dy;
    if (!(r2 < 1.0 && r2 > 1e-24)) return;
    const double r = sqrt(r2);
    const double ir = 1.0 / r;
    const double w = 1.0 - r;
    const double rv = (dx * dvx + dy * dvy) * ir;
    const double th = g_inj.theta ? inj_lookup(g_inj.tk, g_inj.tv, gi, gj, 0.0)
                                  : (theta32 ? pair_theta32(pp.sseed, gi, gj) : pair_theta(pp.sseed, gi, gj));
    const double a = pp.a[8 * ti + tj];
    const double fm = w * (a - pp.gamma * w * rv + pp.sdt * th) * ir;
    *fx += fm * dx;
    *fy += fm * dy;
    if (e) *e += 0.25 * a * w * w; /* half of a/2 (1-r)^2: every pair is visited from both sides */
    if (sabs) *sabs += fabs(fm) * r;
}

/* Stirring force minus its mean over NBSUM fixed blocks,
 * cel/engine/oercel.c:1131-1196 (sin(2 pi u) through libm here). */
static void stir_init(const oc_params *p, int N, const double *y, double *fx, double *fy) {
    for (int i = 0; i < N; i++) { fx[i] = 0.0; fy[i] = 0.0; }
    if (p->stir_A == 0.0) return;
    const double A = p->stir_A, kn = (double)p->stir_n / p->L;
    double tot = 0.0;
    for (int b = 0; b < 64; b++) {
        const int lo = (int)(((long)b * N) / 64), hi = (int)(((long)(b + 1) * N) / 64);
        double sum = 0.0;
        for (int i = lo; i < hi; i++) {
            const double f = A * sin(2.0 * M_PI * (kn * y[i]));
            fx[i] = f;
            sum += f;
        }
        tot += sum;
    }
    const double mean = tot / N;
    for (int i = 0; i < N; i++) fx[i] -= mean;
}

/*
 * Full-shell pair forces on slot-ordered, cell-sorted arrays (own traversal:
 * the reference visits each pair once from a half shell, pair_column,
 * cel/engine/oercel.c:1054-1120; the device and this oracle visit both
 * orientations).  For bead i in column cx the candidates are, per column
 * cx-1, cx, cx+1, the sub-cells floor((y_i-1)/csy) .. floor((y_i+1)/csy)
 * (the window expressions of cel/engine/oercel.c:1073-1074), each piece with
 * the periodic image shift of its column / rows.  Forces are added to fx, fy
 * (which the caller initialises, e.g. with stir_init).
 */
OR_API void or_pair_forces(const oc_params *p, int32_t N, const double *x, const double *y, const double *vtx,
                           const double *vty, const uint8_t *type, const uint32_t *gid, const int32_t *cstart,
                           int64_t step, int32_t theta32, int32_t init_f, double *fx, double *fy, double *epot,
                           double *sabs) {
    const Geo g = make_geo(p, N);
    const PairPar pp = pair_par(p, step, time_of(p, step));
    if (init_f) stir_init(p, N, y, fx, fy);
    const int ncx = g.ncx, ncy = g.ncy;
    const double L = g.L;
#pragma omp parallel for schedule(dynamic, 4)
    for (int cx = 0; cx < ncx; cx++) {
        for (int i = cstart[cx * ncy]; i < cstart[(cx + 1) * ncy]; i++) {
            const double xi = x[i], yi = y[i];
            const int khi = (int)((yi + 1.0) * g.icsy);
            const int klo = (int)((yi - 1.0) * g.icsy + ncy) - ncy;
            double ax = 0.0, ay = 0.0, e = 0.0, sa = 0.0;
            for (int dc = -1; dc <= 1; dc++) {
                int col = cx + dc;
                double sx = 0.0;
                if (col < 0) { col += ncx; sx = -L; }
                else if (col >= ncx) { col -= ncx; sx = L; }
                const int base = col * ncy;
                /* pieces: rows below 0 (image -L), rows 0..ncy-1, rows >= ncy (image +L) */
                int k0[3], k1[3];
                double sy[3];
                int np = 0;
                if (klo < 0) { k0[np] = klo + ncy; k1[np] = ncy - 1; sy[np] = -L; np++; }
                k0[np] = klo < 0 ? 0 : klo; k1[np] = khi >= ncy ? ncy - 1 : khi; sy[np] = 0.0; np++;
                if (khi >= ncy) { k0[np] = 0; k1[np] = khi - ncy; sy[np] = L; np++; }
                for (int q = 0; q < np; q++) {
                    for (int j = cstart[base + k0[q]]; j < cstart[base + k1[q] + 1]; j++) {
                        const double dx = (xi - x[j]) - sx, dy = (yi - y[j]) - sy[q];
                        pair_term(pp, theta32, dx, dy, vtx[i] - vtx[j], vty[i] - vty[j], type[i], type[j], gid[i], gid[j],
                                  &ax, &ay, epot ? &e : nullptr, sabs ? &sa : nullptr);
                    }
                }
            }
            fx[i] += ax;
            fy[i] += ay;
            if (epot) epot[i] = e;
            if (sabs) sabs[i] = sa;
        }
    }
}

/* O(N^2) brute force with minimum image (cel/engine/tests/forcetest.c:91-103
 * computes the same sum this way); any order of arrays. */
OR_API void or_pair_brute(const oc_params *p, int32_t N, const double *x, const double *y, const double *vtx,
                          const double *vty, const uint8_t *type, const uint32_t *gid, int64_t step, int32_t theta32,
                          double *fx, double *fy) {
    const PairPar pp = pair_par(p, step, time_of(p, step));
    const double L = p->L;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
        double ax = 0.0, ay = 0.0;
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            const double dx = minimg(x[i] - x[j], L), dy = minimg(y[i] - y[j], L);
            pair_term(pp, theta32, dx, dy, vtx[i] - vtx[j], vty[i] - vty[j], type[i], type[j], gid[i], gid[j], &ax,
                      &ay, nullptr, nullptr);
        }
        fx[i] = ax;
        fy[i] = ay;
    }
}

/* ------------------------------------------------------------------ lipid templates */

/* mol_build, cel/engine/oercel.c:479-576: bead count, heads, bonds (bi, bj,
 * rest) and angle terms (i, j = middle, k; kind 0: K (1 - cos), kind 1:
 * K (cos - c0)^2), in the reference's order, and each bead's soup-init offset
 * from bead 0 (template centred on its centroid first). */
struct Tmpl {
    int arch = 0, mlen = 0, nH = 0;
    std::vector<int> bi, bj, ai, aj, ak, akind;
    std::vector<double> br0, aK, ac0, ix, iy;
    std::vector<double> px, py; /* centroid-0 coordinates (branched architectures) */
    void bond(int i, int j, double r0) { bi.push_back(i); bj.push_back(j); br0.push_back(r0); }
    void angle(int i, int j, int k, int kind, double K, double c0) {
        if (K == 0.0) return;
        ai.push_back(i); aj.push_back(j); ak.push_back(k); akind.push_back(kind); aK.push_back(K); ac0.push_back(c0);
    }
    void chain(const std::vector<int> &c, double r0, double ka) {
        for (size_t b = 0; b + 1 < c.size(); b++) bond(c[b], c[b + 1], r0);
        for (size_t b = 0; b + 2 < c.size(); b++) angle(c[b], c[b + 1], c[b + 2], 0, ka, 0.0);
    }
};

static Tmpl make_tmpl(const oc_params *p) {
    Tmpl M;
    const double r0 = p->r0, h = p->r0_head;
    const int nT = p->lip_nT;
    M.arch = p->lip_arch;
    std::vector<double> px, py;
    if (p->lip_arch == OC_ARCH_LINEAR) {
        M.nH = p->lip_nH;
        M.mlen = p->lip_nH + nT;
        for (int b = 0; b + 1 < M.mlen; b++) M.bond(b, b + 1, r0);
        for (int b = 0; b + 2 < M.mlen; b++) M.angle(b, b + 1, b + 2, 0, p->k_angle, 0.0);
        for (int b = 0; b < M.mlen; b++) { M.ix.push_back(b * r0); M.iy.push_back(0.0); }
        return M;
    }
    std::vector<int> ca, cb;
    if (p->lip_arch == OC_ARCH_PI) {
        M.nH = 2;
        M.mlen = 2 + 2 * nT;
        M.bond(0, 1, h);
        ca.push_back(0); cb.push_back(1);
        for (int k = 1; k <= nT; k++) { ca.push_back(1 + k); cb.push_back(1 + nT + k); }
        M.chain(ca, r0, p->k_angle);
        M.chain(cb, r0, p->k_angle);
        for (int m = 1; m <= p->lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], h);
        M.angle(1, 0, ca[1], 1, p->k_head, p->cos_head);
        M.angle(0, 1, cb[1], 1, p->k_head, p->cos_head);
        px.assign(M.mlen, 0.0); py.assign(M.mlen, 0.0);
        py[0] = -0.5 * h; py[1] = 0.5 * h;
        for (int k = 1; k <= nT; k++) { px[ca[k]] = k * r0; py[ca[k]] = -0.5 * h; px[cb[k]] = k * r0; py[cb[k]] = 0.5 * h; }
    } else if (p->lip_arch == OC_ARCH_Y) {
        M.nH = 1;
        M.mlen = 1 + 2 * nT;
        ca.push_back(0); cb.push_back(0);
        for (int k = 1; k <= nT; k++) { ca.push_back(k); cb.push_back(nT + k); }
        M.chain(ca, r0, p->k_angle);
        M.chain(cb, r0, p->k_angle);
        for (int m = 1; m <= p->lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], 1.2 * r0);
        M.angle(ca[1], 0, cb[1], 1, p->k_head, p->cos_head);
        px.assign(M.mlen, 0.0); py.assign(M.mlen, 0.0);
        for (int k = 1; k <= nT; k++) {
            px[ca[k]] = 0.8 * r0 + (k - 1) * r0; py[ca[k]] = -0.6 * r0;
            px[cb[k]] = 0.8 * r0 + (k - 1) * r0; py[cb[k]] = 0.6 * r0;
        }
    } else {
        M.nH = 3;
        M.mlen = 3 + 2 * nT;
        M.bond(0, 1, h);
        ca.push_back(0); cb.push_back(1);
        for (int k = 1; k <= nT; k++) { ca.push_back(2 + k); cb.push_back(2 + nT + k); }
        ca.push_back(2); cb.push_back(2);
        M.chain(ca, r0, p->k_angle);
        M.chain(cb, r0, p->k_angle);
        for (int m = 1; m <= p->lip_tie && m <= nT; m++) M.bond(ca[m], cb[m], h * (1.0 - (double)m / (nT + 1)));
        M.angle(1, 0, ca[1], 1, p->k_head, p->cos_head);
        M.angle(0, 1, cb[1], 1, p->k_head, p->cos_head);
        const double X = sqrt((nT + 1) * r0 * (nT + 1) * r0 - 0.25 * h * h);
        px.assign(M.mlen, 0.0); py.assign(M.mlen, 0.0);
        py[0] = -0.5 * h; py[1] = 0.5 * h; px[2] = X;
        for (int k = 1; k <= nT; k++) {
            const double f = (double)k / (nT + 1);
            px[ca[k]] = f * X; py[ca[k]] = -0.5 * h * (1.0 - f);
            px[cb[k]] = f * X; py[cb[k]] = 0.5 * h * (1.0 - f);
        }
    }
    double mx = 0, my = 0;
    for (int b = 0; b < M.mlen; b++) { mx += px[b]; my += py[b]; }
    mx /= M.mlen; my /= M.mlen;
    for (int b = 0; b < M.mlen; b++) { px[b] -= mx; py[b] -= my; }
    for (int b = 0; b < M.mlen; b++) { M.ix.push_back(px[b] - px[0]); M.iy.push_back(py[b] - py[0]); }
    M.px = px;
    M.py = py;
    return M;
}

/* ------------------------------------------------------------------ bonded forces */

/* one harmonic bond i -> j (d = x_j - x_i, minimum image):
 * chain_forces_linear's bond loop, cel/engine/oercel.c:1280-1289 */
static inline void bond_term(double L, double kb, double r0, int i, int j, const double *x, const double *y, double *fx,
                             double *fy, double *mb, double *eb, double *bx, double *by, double *bl) {
    const double dx = minimg(x[j] - x[i], L), dy = minimg(y[j] - y[i], L);
    const double r = sqrt(dx * dx + dy * dy);
    if (bx) { *bx = dx; *by = dy; *bl = r; }
    if (r > *mb) *mb = r;
    const double f = -kb * (r - r0) / r;
    fx[j] += f * dx;
    fy[j] += f * dy;
    fx[i] -= f * dx;
    fy[i] -= f * dy;
    if (eb) *eb += 0.5 * kb * (r - r0) * (r - r0);
}

/*
 * Bonded forces on slot-ordered arrays, added to fx, fy.
 * Lipids (mpos == 0 heads, chain through lnk): bonds then angle terms,
 * chain_forces_linear, cel/engine/oercel.c:1267-1308, walked as in
 * bonded_forces, cel/engine/oercel.c:1317-1343.
 * Chemistry bonds (mpos < 0, lnk = partner): one harmonic bond with k_bond and
 * r0, applied once from the bead with the smaller gid (own definition).
 * Returns the max bond length through *maxbond (max with its input).
 */
OR_API void or_bonded(const oc_params *p, int32_t N, const double *x, const double *y, const int8_t *mpos,
                      const int32_t *lnk, const uint32_t *gid, double *fx, double *fy, double *maxbond, double *ebond) {
    const Tmpl M = make_tmpl(p);
    const int ml = M.mlen;
    const double L = p->L, kb = p->k_bond, r0 = p->r0, ka = p->k_angle;
    double mb = *maxbond, eb = 0.0;
    for (int h = 0; h < N; h++) {
        if (mpos[h] == 0 && M.arch != OC_ARCH_LINEAR) {
            /* template bonds, then angle terms: mol_forces, cel/engine/oercel.c:1218-1257 */
            int sl[OC_MAXMLEN];
            int n = 0;
            for (int d = h; d >= 0 && n < ml; d = lnk[d]) sl[n++] = d;
            if (n != ml) continue;
            for (size_t q = 0; q < M.bi.size(); q++)
                bond_term(L, kb, M.br0[q], sl[M.bi[q]], sl[M.bj[q]], x, y, fx, fy, &mb, ebond ? &eb : nullptr, nullptr,
                          nullptr, nullptr);
            for (size_t q = 0; q < M.ai.size(); q++) {
                const int i = sl[M.ai[q]], j = sl[M.aj[q]], k = sl[M.ak[q]];
                const double ax = minimg(x[j] - x[i], L), ay = minimg(y[j] - y[i], L);
                const double cxv = minimg(x[k] - x[j], L), cyv = minimg(y[k] - y[j], L);
                const double la = sqrt(ax * ax + ay * ay), lc = sqrt(cxv * cxv + cyv * cyv);
                const double ilac = 1.0 / (la * lc);
                const double cs = (ax * cxv + ay * cyv) * ilac;
                const double g = M.akind[q] == 0 ? -M.aK[q] : 2.0 * M.aK[q] * (cs - M.ac0[q]);
                const double dUax = g * (cxv * ilac - cs * ax / (la * la));
                const double dUay = g * (cyv * ilac - cs * ay / (la * la));
                const double dUcx = g * (ax * ilac - cs * cxv / (lc * lc));
                const double dUcy = g * (ay * ilac - cs * cyv / (lc * lc));
                fx[i] += dUax;
                fy[i] += dUay;
                fx[j] += -dUax + dUcx;
                fy[j] += -dUay + dUcy;
                fx[k] -= dUcx;
                fy[k] -= dUcy;
                eb += M.akind[q] == 0 ? M.aK[q] * (1.0 - cs) : M.aK[q] * (cs - M.ac0[q]) * (cs - M.ac0[q]);
            }
        } else if (mpos[h] == 0) {
            int sl[OC_MAXMLEN];
            double bx[OC_MAXMLEN], by[OC_MAXMLEN], bl[OC_MAXMLEN];
            int n = 0;
            for (int d = h; d >= 0 && n < ml; d = lnk[d]) sl[n++] = d;
            for (int b = 0; b < n - 1; b++)
                bond_term(L, kb, r0, sl[b], sl[b + 1], x, y, fx, fy, &mb, ebond ? &eb : nullptr, &bx[b], &by[b], &bl[b]);
            if (ka != 0.0) {
                /* cel/engine/oercel.c:1290-1306 */
                for (int b = 0; b < n - 2; b++) {
                    const int i = sl[b], j = sl[b + 1], k = sl[b + 2];
                    const double ax = bx[b], ay = by[b], cxv = bx[b + 1], cyv = by[b + 1];
                    const double la = bl[b], lc = bl[b + 1];
                    const double ilac = 1.0 / (la * lc);
                    const double cs = (ax * cxv + ay * cyv) * ilac;
                    const double dUax = -ka * (cxv * ilac - cs * ax / (la * la));
                    const double dUay = -ka * (cyv * ilac - cs * ay / (la * la));
                    const double dUcx = -ka * (ax * ilac - cs * cxv / (lc * lc));
                    const double dUcy = -ka * (ay * ilac - cs * cyv / (lc * lc));
                    fx[i] += dUax;
                    fy[i] += dUay;
                    fx[j] += -dUax + dUcx;
                    fy[j] += -dUay + dUcy;
                    fx[k] -= dUcx;
                    fy[k] -= dUcy;
                    eb += ka * (1.0 - cs);
                }
            }
        } else if (mpos[h] < 0 && lnk[h] >= 0 && gid[h] < gid[lnk[h]]) {
            bond_term(L, kb, r0, h, lnk[h], x, y, fx, fy, &mb, ebond ? &eb : nullptr, nullptr, nullptr, nullptr);
        }
    }
    *maxbond = mb;
    if (ebond) *ebond = eb;
}

/* ------------------------------------------------------------------ integrator */

/* Groot-Warren update with stored half-step velocity, cel/engine/oercel.c:2303-2334 */
OR_API void or_integrate(const oc_params *p, int32_t N, double *x, double *y, double *vx, double *vy, const double *fx,
                         const double *fy, double *vtx, double *vty) {
    const double dt = p->dt, dt2 = dt * dt, cvt = (0.5 + p->lambda) * dt, L = p->L;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        const double ddx = dt * vx[i] + dt2 * fx[i], ddy = dt * vy[i] + dt2 * fy[i];
        double xi = x[i] + ddx, yi = y[i] + ddy;
        xi = xi < 0.0 ? xi + L : xi;
        yi = yi < 0.0 ? yi + L : yi;
        xi = xi >= L ? xi - L : xi;
        yi = yi >= L ? yi - L : yi;
        x[i] = xi;
        y[i] = yi;
        vtx[i] = vx[i] + cvt * fx[i];
        vty[i] = vy[i] + cvt * fy[i];
        vx[i] += dt * fx[i];
        vy[i] += dt * fy[i];
    }
}

/* ------------------------------------------------------------------ soup init */

/* init_soup + set_velocities, cel/engine/oercel.c:1757-1814, after
 * rng_seed(seed) (cel/engine/oercel.c:2195); gid order = creation order;
 * nxt as gid; velocities are the full v.  Returns 0 or -1 (composition > N). */
OR_API int or_soup(const oc_params *p, int32_t N, double *x, double *y, double *vx, double *vy, uint8_t *type,
                   int8_t *mpos, int32_t *nxt) {
    Rng rng;
    rng_seed(&rng, p->seed);
    const Tmpl M = make_tmpl(p);
    const int ml = M.mlen;
    const double L = p->L;
    const long nlip = lround(p->f_lip * N / ml);
    const long nKa = lround(p->f_Ka * N), nKb = lround(p->f_Kb * N), nN = lround(p->c_N * N);
    if (nlip * ml + nKa + nKb + nN > N) return -1;
    int gi = 0;
    for (long m = 0; m < nlip; m++) {
        const double hx = L * rng_unif(&rng), hy = L * rng_unif(&rng);
        const double phi = 2.0 * M_PI * rng_unif(&rng);
        const double cph = cos(phi), sph = sin(phi);
        for (int b = 0; b < ml; b++) {
            if (M.arch == OC_ARCH_LINEAR) {
                x[gi] = hx + b * p->r0 * cos(phi);
                y[gi] = hy + b * p->r0 * sin(phi);
            } else {
                x[gi] = hx + (M.ix[b] * cph - M.iy[b] * sph);
                y[gi] = hy + (M.ix[b] * sph + M.iy[b] * cph);
            }
            wrap1(&x[gi], L);
            wrap1(&y[gi], L);
            type[gi] = b < M.nH ? OC_H : OC_T;
            mpos[gi] = (int8_t)b;
            nxt[gi] = b + 1 < ml ? gi + 1 : -1;
            gi++;
        }
    }
    long cnt = 0;
    for (; gi < N; gi++, cnt++) {
        x[gi] = L * rng_unif(&rng);
        y[gi] = L * rng_unif(&rng);
        mpos[gi] = -1;
        nxt[gi] = -1;
        if (cnt < nKa) type[gi] = OC_KA;
        else if (cnt < nKa + nKb) type[gi] = OC_KB;
        else if (cnt < nKa + nKb + nN) type[gi] = OC_N;
        else type[gi] = OC_W;
    }
    double px = 0, py = 0;
    for (int i = 0; i < N; i++) {
        vx[i] = rng_normal(&rng) * sqrt(p->kT);
        vy[i] = rng_normal(&rng) * sqrt(p->kT);
        px += vx[i];
        py += vy[i];
    }
    px /= N;
    py /= N;
    double k2 = 0;
    for (int i = 0; i < N; i++) {
        vx[i] -= px;
        vy[i] -= py;
        k2 += vx[i] * vx[i] + vy[i] * vy[i];
    }
    const double sc = sqrt(p->kT * (2.0 * N - 2.0) / k2);
    for (int i = 0; i < N; i++) {
        vx[i] *= sc;
        vy[i] *= sc;
    }
    return 0;
}

/* ------------------------------------------------------------------ precursor chemistry (own definition) */

/*
 * One chemistry check at `step` (only if chem && step % n_chem == 0), on a
 * cell-sorted state.  All decisions read the state at the start of the check:
 *   break : a chemistry bond (mpos < 0, lnk >= 0) breaks if
 *           hash_u01(hash_key(seed, step, 33), min gid) < p_break;
 *   conv  : a P bead (type 3) becomes a free T if
 *           hash_u01(hash_key(seed, step, 31), gid) < p_conv and some lipid T
 *           (type T, mpos >= 0) lies within r < r_conv;
 *   form  : a free T (type T, mpos < 0, lnk < 0) proposes to the free T within
 *           r < r_form with the smallest pair_u01(hash_key(seed, step, 32))
 *           below p_form (ties: smaller gid); mutual proposals bond.
 * Distances are FP64 minimum-image, r^2 < rc*rc.  Events are returned sorted by
 * (kind, gid_a, gid_b).  Returns the number of events (<= cap copied).
 * PARITY UNPINNED against the reference, which has no such chemistry.
 */
OR_API int64_t or_chem(const oc_params *p, int32_t N, int64_t step, const double *x, const double *y, uint8_t *type,
                       const int8_t *mpos, int32_t *lnk, const uint32_t *gid, const int32_t *cstart, oc_event *ev,
                       int64_t cap) {
    if (!p->chem || p->n_chem < 1 || step % p->n_chem != 0) return 0;
    const Geo g = make_geo(p, N);
    const double L = g.L;
    const uint64_t kc = hash_key(p->seed, step, SALT_CONV), kf = hash_key(p->seed, step, SALT_FORM),
                   kb = hash_key(p->seed, step, SALT_BREAK);
    const double rc2 = p->r_conv * p->r_conv, rf2 = p->r_form * p->r_form;
    std::vector<uint8_t> conv(N, 0), brk(N, 0);
    std::vector<int32_t> best(N, -1);
    /* neighbour visitor over the cell window of bead i (cutoff <= 1) */
    auto visit = [&](int i, auto &&fn) {
        const int cx = cell_key(g, x[i], y[i]) / g.ncy;
        const int khi = (int)((y[i] + 1.0) * g.icsy);
        const int klo = (int)((y[i] - 1.0) * g.icsy + g.ncy) - g.ncy;
        for (int dc = -1; dc <= 1; dc++) {
            const int col = (cx + dc + g.ncx) % g.ncx;
            for (int k = klo; k <= khi; k++) {
                const int kk = (k % g.ncy + g.ncy) % g.ncy;
                const int c = col * g.ncy + kk;
                for (int j = cstart[c]; j < cstart[c + 1]; j++) {
                    if (j == i) continue;
                    const double dx = minimg(x[j] - x[i], L), dy = minimg(y[j] - y[i], L);
                    if (!fn(j, dx * dx + dy * dy)) return;
                }
            }
        }
    };
    for (int i = 0; i < N; i++) {
        const double ub = g_inj.chem ? g_inj.ubreak[gid[i]] : hash_u01(kb, gid[i]);
        const double uc = g_inj.chem ? g_inj.uconv[gid[i]] : hash_u01(kc, gid[i]);
        if (mpos[i] < 0 && lnk[i] >= 0 && gid[i] < gid[lnk[i]] && ub < p->p_break) brk[i] = 1;
        if (type[i] == OC_N && uc < p->p_conv) {
            bool found = false;
            visit(i, [&](int j, double r2) {
                if (type[j] == OC_T && mpos[j] >= 0 && r2 < rc2) { found = true; return false; }
                return true;
            });
            conv[i] = found;
        }
        if (type[i] == OC_T && mpos[i] < 0 && lnk[i] < 0) {
            double bu = 2.0;
            uint32_t bg = 0xffffffffu;
            int bj = -1;
            visit(i, [&](int j, double r2) {
                if (type[j] == OC_T && mpos[j] < 0 && lnk[j] < 0 && r2 < rf2) {
                    const double u = g_inj.chem ? inj_lookup(g_inj.fk, g_inj.fv, gid[i], gid[j], 2.0)
                                                : pair_u01(kf, gid[i], gid[j]);
                    if (u < p->p_form && (u < bu || (u == bu && gid[j] < bg))) { bu = u; bg = gid[j]; bj = j; }
                }
                return true;
            });
            best[i] = bj;
        }
    }
    std::vector<oc_event> out;
    for (int i = 0; i < N; i++) {
        if (brk[i]) {
            const int j = lnk[i];
            out.push_back({step, OC_EV_BREAK, std::min(gid[i], gid[j]), std::max(gid[i], gid[j]), 0});
        }
        if (conv[i]) out.push_back({step, OC_EV_CONV, gid[i], gid[i], 0});
        if (best[i] >= 0 && best[best[i]] == i && gid[i] < gid[best[i]])
            out.push_back({step, OC_EV_FORM, gid[i], gid[best[i]], 0});
    }
    /* apply */
    for (int i = 0; i < N; i++) {
        if (brk[i]) { const int j = lnk[i]; lnk[i] = -1; lnk[j] = -1; }
    }
    for (int i = 0; i < N; i++) {
        if (conv[i]) type[i] = OC_T;
        if (best[i] >= 0 && best[best[i]] == i && gid[i] < gid[best[i]]) { lnk[i] = best[i]; lnk[best[i]] = i; }
    }
    std::sort(out.begin(), out.end(), [](const oc_event &a, const oc_event &b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.gid_a != b.gid_a) return a.gid_a < b.gid_a;
        return a.gid_b < b.gid_b;
    });
    const int64_t n = (int64_t)out.size();
    for (int64_t k = 0; k < n && k < cap; k++) ev[k] = out[k];
    return n;
}

/* ------------------------------------------------------------------ reference chemistry: decay and food supply */

/*
 * The decay pass and the food supply of the reference's chemistry check
 * (cel/engine/oercel.c:1656-1753), after the catalyst reactions: a lipid whose
 * first bead draws hash_u01(kl, gid) < p_ld becomes water along its chain, a
 * catalyst with hash_u01(kk, gid) < p_kd and an N with hash_u01(kn, gid) < p_nd
 * become W (decisions collected first, then applied); then every W, or every W
 * inside the vent region, with hash_u01(kf, gid) < p_feed becomes N.  Keys
 * hash_key(seed, step, 11 / 12 / 13 / 14).  With feed_cap > 0 and more
 * conversions than room = floor(feed_cap N + 1e-9) - (number of N), the kept
 * ones are the room smallest keys (hash_u64(hash_key(seed, step, 25), gid) with
 * the low 32 bits replaced by the gid): the reference GPU port's cap
 * (cel/engine_gpu/oercel_gpu.cu:1944-1990); the CPU engine picks the subset
 * from its serial random stream.  Events (dlip: first bead; dK; dN; feed: one
 * per bead) with the bead's position, sorted by (step, kind, gid_a, gid_b).
 */
static bool in_vent(const oc_params *p, double x, double y) {
    const double L = p->L;
    if (p->vent_shape == OC_VENT_DISK) {
        const double dx = minimg(x - p->vent_x, L), dy = minimg(y - p->vent_y, L);
        return dx * dx + dy * dy < p->vent_r * p->vent_r;
    }
    if (p->vent_shape == OC_VENT_VSTRIP) return fabs(minimg(x - p->vent_x, L)) < 0.5 * p->vent_w;
    return fabs(minimg(y - p->vent_y, L)) < 0.5 * p->vent_w;
}

OR_API int64_t or_rchem(const oc_params *p, int32_t N, int64_t step, const double *x, const double *y, uint8_t *type,
                        int8_t *mpos, int32_t *lnk, const uint32_t *gid, oc_event *ev, int64_t cap) {
    if (!p->rchem || p->n_chem < 1 || step % p->n_chem != 0) return 0;
    const Tmpl M = make_tmpl(p);
    const uint64_t kl = hash_key(p->seed, step, 11), kk = hash_key(p->seed, step, 12),
                   kn = hash_key(p->seed, step, 13), kf = hash_key(p->seed, step, 14),
                   kc = hash_key(p->seed, step, 25);
    std::vector<oc_event> out;
    std::vector<int> dl;
    for (int d = 0; d < N; d++) {
        const bool kat = type[d] == OC_KA || type[d] == OC_KB;
        if ((mpos[d] == 0 && p->p_ld > 0 && hash_u01(kl, gid[d]) < p->p_ld) ||
            (kat && p->p_kd > 0 && hash_u01(kk, gid[d]) < p->p_kd) ||
            (type[d] == OC_N && p->p_nd > 0 && hash_u01(kn, gid[d]) < p->p_nd))
            dl.push_back(d);
    }
    for (int d : dl) {
        const int t = type[d];
        if (mpos[d] == 0) {
            oc_event e{step, OC_EV_DLIP, gid[d], gid[d], 0, x[d], y[d]};
            int c = d, n = 0;
            while (c >= 0 && n < M.mlen) {
                const int nx = lnk[c];
                type[c] = OC_W;
                mpos[c] = -1;
                lnk[c] = -1;
                c = nx;
                n++;
            }
            out.push_back(e);
        } else if (t == OC_KA || t == OC_KB) {
            type[d] = OC_W;
            out.push_back(oc_event{step, OC_EV_DK, gid[d], (uint32_t)(t == OC_KB), 0, x[d], y[d]});
        } else if (t == OC_N) {
            type[d] = OC_W;
            out.push_back(oc_event{step, OC_EV_DN, gid[d], gid[d], 0, x[d], y[d]});
        }
    }
    if (p->feed_mode != OC_FEED_NONE && p->p_feed > 0) {
        std::vector<int> fc;
        for (int d = 0; d < N; d++)
            if (type[d] == OC_W && (p->feed_mode == OC_FEED_UNIFORM || in_vent(p, x[d], y[d])) &&
                hash_u01(kf, gid[d]) < p->p_feed)
                fc.push_back(d);
        int64_t add = (int64_t)fc.size();
        uint64_t thr = ~0ULL;
        if (p->feed_cap > 0.0 && add > 0) {
            int64_t nN = 0;
            for (int d = 0; d < N; d++) nN += type[d] == OC_N;
            int64_t room = (int64_t)floor(p->feed_cap * (double)N + 1e-9) - nN;
            if (room < 0) room = 0;
            if (add > room) {
                add = room;
                if (room > 0) {
                    std::vector<uint64_t> keys;
                    for (int d : fc) keys.push_back((hash_u64(kc, gid[d]) & 0xFFFFFFFF00000000ULL) | gid[d]);
                    std::nth_element(keys.begin(), keys.begin() + (room - 1), keys.end());
                    thr = keys[room - 1];
                }
            }
        }
        if (add > 0)
            for (int d : fc)
                if (((hash_u64(kc, gid[d]) & 0xFFFFFFFF00000000ULL) | gid[d]) <= thr) {
                    type[d] = OC_N;
                    out.push_back(oc_event{step, OC_EV_FEED, gid[d], gid[d], 0, x[d], y[d]});
                }
    }
    std::sort(out.begin(), out.end(), [](const oc_event &a, const oc_event &b) {
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.gid_a != b.gid_a) return a.gid_a < b.gid_a;
        return a.gid_b < b.gid_b;
    });
    const int64_t n = (int64_t)out.size();
    for (int64_t k = 0; k < n && k < cap; k++) ev[k] = out[k];
    return n;
}

/* ------------------------------------------------------------------ system (composition of the above) */

struct OrSys {
    oc_params p;
    Geo g;
    int N;
    std::vector<double> x, y, vx, vy, fx, fy, vtx, vty;
    std::vector<uint8_t> type;
    std::vector<uint32_t> gid;
    std::vector<int8_t> mpos;
    std::vector<int32_t> lnk, cstart, key, src;
    int64_t step;
    double maxbond;
    std::vector<oc_event> ev;
};

OR_API OrSys *or_create(const oc_params *p, int32_t N) {
    OrSys *s = new OrSys();
    s->p = *p;
    s->g = make_geo(p, N);
    s->N = N;
    s->x.assign(N, 0); s->y.assign(N, 0); s->vx.assign(N, 0); s->vy.assign(N, 0);
    s->fx.assign(N, 0); s->fy.assign(N, 0); s->vtx.assign(N, 0); s->vty.assign(N, 0);
    s->type.assign(N, 0); s->gid.assign(N, 0); s->mpos.assign(N, -1); s->lnk.assign(N, -1);
    s->cstart.assign(s->g.ncell + 1, 0);
    s->key.assign(N, 0); s->src.assign(N, 0);
    s->step = 0;
    s->maxbond = 0;
    return s;
}
OR_API void or_destroy(OrSys *s) { delete s; }

/* sort_cells on the system: keys, stable order, gather every slot array and
 * remap the links through dest (cel/engine/oercel.c:897-932) */
static void sys_sort(OrSys *s) {
    const int N = s->N;
    or_cell_keys(&s->p, N, s->x.data(), s->y.data(), s->key.data());
    or_sort(N, s->g.ncell, s->key.data(), s->src.data(), s->cstart.data());
    std::vector<int32_t> dest(N);
    for (int d = 0; d < N; d++) dest[s->src[d]] = d;
    auto gather = [&](auto &v) {
        auto w = v;
        for (int d = 0; d < N; d++) w[d] = v[s->src[d]];
        v.swap(w);
    };
    gather(s->x); gather(s->y); gather(s->vx); gather(s->vy); gather(s->vtx); gather(s->vty);
    gather(s->fx); gather(s->fy); gather(s->type); gather(s->gid); gather(s->mpos);
    std::vector<int32_t> l2(N);
    for (int d = 0; d < N; d++) {
        const int32_t o = s->lnk[s->src[d]];
        l2[d] = o < 0 ? -1 : dest[o];
    }
    s->lnk.swap(l2);
}

/* ------------------------------------------------------------------ catalyst chemistry */

/*
 * The catalyst reactions of a chemistry check in the reference GPU port's
 * form (chem_react and the reaction part of chemistry,
 * cel/engine_gpu/oercel_gpu.cu:1645-1709, 1800-1880).  It keeps the CPU
 * engine's per-catalyst body (cel/engine/oercel.c:1575-1653) and replaces its
 * serial random stream by per-catalyst hashes:
 *   - a catalyst (Ka / Kb, gid g) is active if hash_u01(k21, g) < p_syn or
 *     hash_u01(k22, g) < p_rep of its kind;
 *   - the active ones react one after the other in increasing
 *     (hash_u64(k24, g), g), each on the state the earlier ones left;
 *   - synthesis orientations from hash_u01(hash_key(seed, step, 40 + try), g),
 *     mutation from hash_u01(k23, g);
 *   - geometry (search windows, overlap energy, relaxation and their visiting
 *     order) is the CPU engine's, on the cell index of the last sort (every
 *     bead moved less than the chemistry margin 0.5 since).
 */
static const double kChemMargin = 0.5;

struct Near {
    int slot;
    double d2;
};

/* the K nearest beads of `type` within r_cat of slot k, by (d2, gid); returns
 * how many lie within r_cat, at most 256 (find_type_near,
 * cel/engine/oercel.c:1353-1397; its list keeps the first 256 it visits, which
 * differs only with more than 256 such beads within r_cat) */
static int near_of_type(const OrSys *s, int k, int type, Near *out, int K) {
    const Geo &g = s->g;
    const double xk = s->x[k], yk = s->y[k], L = g.L, r2c = s->p.r_cat * s->p.r_cat;
    const int c = cell_key(g, xk, yk), cx = c / g.ncy, cy = c % g.ncy;
    const int rx = (int)ceil((s->p.r_cat + kChemMargin) * g.icsx);
    const int ry = (int)ceil((s->p.r_cat + kChemMargin) * g.icsy);
    int n = 0, nk = 0;
    for (int ox = -rx; ox <= rx; ox++) {
        const int xx = (cx + ox + g.ncx) % g.ncx;
        for (int oy = -ry; oy <= ry; oy++) {
            const int cc = xx * g.ncy + (cy + oy + g.ncy) % g.ncy;
            for (int j = s->cstart[cc]; j < s->cstart[cc + 1]; j++) {
                if (s->type[j] != type || j == k) continue;
                const double dx = minimg(s->x[j] - xk, L), dy = minimg(s->y[j] - yk, L);
                const double d2 = dx * dx + dy * dy;
                if (!(d2 < r2c)) continue;
                if (n < 256) n++;
                if (K <= 0) continue;
                const uint32_t gj = s->gid[j];
                if (nk == K && !(d2 < out[K - 1].d2 || (d2 == out[K - 1].d2 && gj < s->gid[out[K - 1].slot]))) continue;
                int b = nk < K ? nk++ : K - 1;
                while (b > 0 && (out[b - 1].d2 > d2 || (out[b - 1].d2 == d2 && s->gid[out[b - 1].slot] > gj))) {
                    out[b] = out[b - 1];
                    b--;
                }
                out[b].slot = j;
                out[b].d2 = d2;
            }
        }
    }
    return n;
}

/* bead b of a new lipid centred at (cx0, cy0) along (ux, uy), wrapped */
static inline void chain_pos(const OrSys *s, const Tmpl &M, int b, double cx0, double cy0, double ux, double uy,
                             double *px, double *py) {
    if (M.arch == OC_ARCH_LINEAR) {
        const double off = (b - 0.5 * (M.mlen - 1)) * s->p.r0;
        *px = cx0 + off * ux;
        *py = cy0 + off * uy;
    } else {
        *px = cx0 + (M.px[b] * ux - M.py[b] * uy);
        *py = cy0 + (M.px[b] * uy + M.py[b] * ux);
    }
    wrap1(px, s->g.L);
    wrap1(py, s->g.L);
}

static inline bool in_list(int j, const int *sl, int n) {
    bool hit = false;
    for (int q = 0; q < n; q++) hit |= j == sl[q];
    return hit;
}

/* conservative overlap energy of a would-be lipid with every other bead within
 * 1, chain_overlap_energy, cel/engine/oercel.c:1423-1468 (the raw a_ij) */
static double overlap_energy(const OrSys *s, const Tmpl &M, double cx0, double cy0, double ux, double uy,
                             const int *sl) {
    const Geo &g = s->g;
    const double L = g.L;
    const int rx = (int)ceil((1.0 + kChemMargin) * g.icsx), ry = (int)ceil((1.0 + kChemMargin) * g.icsy);
    double U = 0.0;
    for (int b = 0; b < M.mlen; b++) {
        double px, py;
        chain_pos(s, M, b, cx0, cy0, ux, uy, &px, &py);
        const int tb = b < M.nH ? OC_H : OC_T;
        const int c = cell_key(g, px, py), cx = c / g.ncy, cy = c % g.ncy;
        for (int ox = -rx; ox <= rx; ox++) {
            const int xx = (cx + ox + g.ncx) % g.ncx;
            for (int oy = -ry; oy <= ry; oy++) {
                const int cc = xx * g.ncy + (cy + oy + g.ncy) % g.ncy;
                for (int j = s->cstart[cc]; j < s->cstart[cc + 1]; j++) {
                    if (in_list(j, sl, M.mlen)) continue;
                    const double dx = minimg(s->x[j] - px, L), dy = minimg(s->y[j] - py, L);
                    const double r2 = dx * dx + dy * dy;
                    if (r2 >= 1.0) continue;
                    const double w = 1.0 - sqrt(r2);
                    U += 0.5 * s->p.a[tb][s->type[j]] * w * w;
                }
            }
        }
    }
    return U;
}

/* steepest descent of a new lipid in the frozen field of its neighbours
 * (pair overlap + its own bonds and angles), relax_new_chain,
 * cel/engine/oercel.c:1474-1556: neighbours within 1 + 0.05 syn_relax of any
 * bead collected once (per-bead windows, first visit order, at most 512),
 * then syn_relax moves of 0.002 F capped at 0.05 per bead */
static void relax_chain(OrSys *s, const Tmpl &M, const int *sl) {
    const Geo &g = s->g;
    const int ml = M.mlen;
    const double L = g.L, cap = 0.05, alpha = 0.002;
    const int maxnb = 512;
    std::vector<int> nbj;
    const double reach = 1.0 + cap * s->p.syn_relax, reach2 = reach * reach;
    const int rx = (int)ceil((reach + kChemMargin) * g.icsx), ry = (int)ceil((reach + kChemMargin) * g.icsy);
    for (int b = 0; b < ml && (int)nbj.size() < maxnb; b++) {
        const double px = s->x[sl[b]], py = s->y[sl[b]];
        const int c = cell_key(g, px, py), cx = c / g.ncy, cy = c % g.ncy;
        for (int ox = -rx; ox <= rx; ox++) {
            const int xx = ((cx + ox) % g.ncx + g.ncx) % g.ncx;
            for (int oy = -ry; oy <= ry; oy++) {
                const int cc = xx * g.ncy + ((cy + oy) % g.ncy + g.ncy) % g.ncy;
                for (int j = s->cstart[cc]; j < s->cstart[cc + 1]; j++) {
                    if (in_list(j, sl, ml) || in_list(j, nbj.data(), (int)nbj.size())) continue;
                    const double dx = minimg(s->x[j] - px, L), dy = minimg(s->y[j] - py, L);
                    if (dx * dx + dy * dy < reach2 && (int)nbj.size() < maxnb) nbj.push_back(j);
                }
            }
        }
    }
    for (int it = 0; it < s->p.syn_relax; it++) {
        double fx[OC_MAXMLEN] = {0}, fy[OC_MAXMLEN] = {0};
        for (int b = 0; b < ml; b++) {
            const int i = sl[b];
            for (int j : nbj) {
                const double dx = minimg(s->x[i] - s->x[j], L), dy = minimg(s->y[i] - s->y[j], L);
                const double r2 = dx * dx + dy * dy;
                if (r2 >= 1.0 || r2 < 1e-24) continue;
                const double r = sqrt(r2);
                const double f = s->p.a[s->type[i]][s->type[j]] * (1.0 - r) / r;
                fx[b] += f * dx;
                fy[b] += f * dy;
            }
        }
        const double kb = s->p.k_bond;
        if (M.arch == OC_ARCH_LINEAR) {
            double bx[OC_MAXMLEN], by[OC_MAXMLEN], bl[OC_MAXMLEN];
            for (int b = 0; b < ml - 1; b++) {
                const int i = sl[b], j = sl[b + 1];
                bx[b] = minimg(s->x[j] - s->x[i], L);
                by[b] = minimg(s->y[j] - s->y[i], L);
                bl[b] = sqrt(bx[b] * bx[b] + by[b] * by[b]);
                const double f = -kb * (bl[b] - s->p.r0) / bl[b];
                fx[b + 1] += f * bx[b];
                fy[b + 1] += f * by[b];
                fx[b] -= f * bx[b];
                fy[b] -= f * by[b];
            }
            const double ka = s->p.k_angle;
            for (int b = 0; b < ml - 2 && ka != 0.0; b++) {
                const double ax = bx[b], ay = by[b], cxv = bx[b + 1], cyv = by[b + 1], la = bl[b], lc = bl[b + 1];
                const double ilac = 1.0 / (la * lc), cs = (ax * cxv + ay * cyv) * ilac;
                const double dUax = -ka * (cxv * ilac - cs * ax / (la * la)), dUay = -ka * (cyv * ilac - cs * ay / (la * la));
                const double dUcx = -ka * (ax * ilac - cs * cxv / (lc * lc)), dUcy = -ka * (ay * ilac - cs * cyv / (lc * lc));
                fx[b] += dUax;
                fy[b] += dUay;
                fx[b + 1] += -dUax + dUcx;
                fy[b + 1] += -dUay + dUcy;
                fx[b + 2] -= dUcx;
                fy[b + 2] -= dUcy;
            }
        } else {
            /* mol_forces by template position, cel/engine/oercel.c:1218-1257 */
            for (size_t q = 0; q < M.bi.size(); q++) {
                const int bi = M.bi[q], bj = M.bj[q], i = sl[bi], j = sl[bj];
                const double dx = minimg(s->x[j] - s->x[i], L), dy = minimg(s->y[j] - s->y[i], L);
                const double r = sqrt(dx * dx + dy * dy);
                const double f = -kb * (r - M.br0[q]) / r;
                fx[bj] += f * dx;
                fy[bj] += f * dy;
                fx[bi] -= f * dx;
                fy[bi] -= f * dy;
            }
            for (size_t q = 0; q < M.ai.size(); q++) {
                const int pi_ = M.ai[q], pj = M.aj[q], pk = M.ak[q];
                const int i = sl[pi_], j = sl[pj], k = sl[pk];
                const double ax = minimg(s->x[j] - s->x[i], L), ay = minimg(s->y[j] - s->y[i], L);
                const double cxv = minimg(s->x[k] - s->x[j], L), cyv = minimg(s->y[k] - s->y[j], L);
                const double la = sqrt(ax * ax + ay * ay), lc = sqrt(cxv * cxv + cyv * cyv);
                const double ilac = 1.0 / (la * lc);
                const double cs = (ax * cxv + ay * cyv) * ilac;
                const double gg = M.akind[q] == 0 ? -M.aK[q] : 2.0 * M.aK[q] * (cs - M.ac0[q]);
                const double dUax = gg * (cxv * ilac - cs * ax / (la * la));
                const double dUay = gg * (cyv * ilac - cs * ay / (la * la));
                const double dUcx = gg * (ax * ilac - cs * cxv / (lc * lc));
                const double dUcy = gg * (ay * ilac - cs * cyv / (lc * lc));
                fx[pi_] += dUax;
                fy[pi_] += dUay;
                fx[pj] += -dUax + dUcx;
                fy[pj] += -dUay + dUcy;
                fx[pk] -= dUcx;
                fy[pk] -= dUcy;
            }
        }
        for (int b = 0; b < ml; b++) {
            double mx = alpha * fx[b], my = alpha * fy[b];
            const double m = sqrt(mx * mx + my * my);
            if (m > cap) {
                mx *= cap / m;
                my *= cap / m;
            }
            s->x[sl[b]] += mx;
            s->y[sl[b]] += my;
            wrap1(&s->x[sl[b]], L);
            wrap1(&s->y[sl[b]], L);
        }
    }
}

/* one active catalyst's reactions: synthesis, then replication
 * (chem_react, cel/engine_gpu/oercel_gpu.cu:1645-1709) */
static void react_one(OrSys *s, const Tmpl &M, int64_t step, uint64_t ksyn, uint64_t krep, uint64_t kmu, int k,
                      std::vector<oc_event> &out) {
    const oc_params &p = s->p;
    const int tk = s->type[k];
    if (tk != OC_KA && tk != OC_KB) return;
    const int var = tk == OC_KB;
    const uint32_t g = s->gid[k];
    const double us = hash_u01(ksyn, g), ur = hash_u01(krep, g);
    const double psyn = var ? p.p_syn_b : p.p_syn_a, prep = var ? p.p_rep_b : p.p_rep_a;
    const int ml = M.mlen, nsynN = p.syn_nN > 0 ? p.syn_nN : ml;
    const double L = s->g.L;
    Near nb[OC_MAXMLEN + 2], nbw[OC_MAXMLEN];
    const int nn = near_of_type(s, k, OC_N, nb, nsynN + 2);
    if (nn < 2) return;
    int used = 0;
    const double xk = s->x[k], yk = s->y[k];
    int nw = 0;
    if (nn >= nsynN && nsynN < ml) nw = near_of_type(s, k, OC_W, nbw, ml - nsynN);
    if (nn >= nsynN && nsynN + nw >= ml && us < psyn) {
        double cxm = 0, cym = 0, mvx = 0, mvy = 0, mtx = 0, mty = 0;
        int sl[OC_MAXMLEN];
        for (int b = 0; b < ml; b++) {
            const int j = b < nsynN ? nb[b].slot : nbw[b - nsynN].slot;
            sl[b] = j;
            cxm += minimg(s->x[j] - xk, L);
            cym += minimg(s->y[j] - yk, L);
            mvx += s->vx[j];
            mvy += s->vy[j];
            mtx += s->vtx[j];
            mty += s->vty[j];
        }
        cxm = xk + cxm / ml;
        cym = yk + cym / ml;
        mvx /= ml; mvy /= ml; mtx /= ml; mty /= ml;
        double ux = 1.0, uy = 0.0, ubest = INFINITY;
        for (int tr = 0; tr < p.syn_tries; tr++) {
            const double phi = 2.0 * M_PI * hash_u01(hash_key(p.seed, step, 40 + (uint64_t)tr), g);
            const double cx_ = cos(phi), cy_ = sin(phi);
            const double U = p.syn_tries > 1 ? overlap_energy(s, M, cxm, cym, cx_, cy_, sl) : 0.0;
            if (U < ubest) {
                ubest = U;
                ux = cx_;
                uy = cy_;
            }
        }
        for (int b = 0; b < ml; b++) {
            const int j = sl[b];
            chain_pos(s, M, b, cxm, cym, ux, uy, &s->x[j], &s->y[j]);
            s->vx[j] = mvx;
            s->vy[j] = mvy;
            s->vtx[j] = mtx;
            s->vty[j] = mty;
            s->type[j] = b < M.nH ? OC_H : OC_T;
        }
        for (int b = 0; b < ml; b++) {
            s->mpos[sl[b]] = (int8_t)b;
            s->lnk[sl[b]] = b + 1 < ml ? sl[b + 1] : -1;
        }
        if (p.syn_relax > 0) relax_chain(s, M, sl);
        out.push_back(oc_event{step, OC_EV_SYN, g, (uint32_t)var, 0, xk, yk});
        used = nsynN;
    }
    if (nn - used >= 2 && ur < prep) {
        int nv = var, mut = 0;
        if (hash_u01(kmu, g) < p.mu) {
            nv = 1 - var;
            mut = 1;
        }
        const int j1 = nb[used].slot, j2 = nb[used + 1].slot;
        s->type[j1] = nv ? OC_KB : OC_KA;
        s->type[j2] = OC_W;
        out.push_back(oc_event{step, mut ? OC_EV_MUT : OC_EV_REP, s->gid[j1], (uint32_t)nv, 0, xk, yk});
    }
}

/* the reactions of one check: active catalysts in priority order */
static void sys_react(OrSys *s, int64_t step, std::vector<oc_event> &out) {
    const oc_params &p = s->p;
    if (!(p.p_syn_a > 0 || p.p_syn_b > 0 || p.p_rep_a > 0 || p.p_rep_b > 0)) return;
    const Tmpl M = make_tmpl(&p);
    const uint64_t ksyn = hash_key(p.seed, step, 21), krep = hash_key(p.seed, step, 22),
                   kmu = hash_key(p.seed, step, 23), kpri = hash_key(p.seed, step, 24);
    struct Act {
        uint64_t key;
        uint32_t g;
        int slot;
    };
    std::vector<Act> act;
    for (int d = 0; d < s->N; d++) {
        const int t = s->type[d];
        if (t != OC_KA && t != OC_KB) continue;
        const int var = t == OC_KB;
        const uint32_t g = s->gid[d];
        if (hash_u01(ksyn, g) < (var ? p.p_syn_b : p.p_syn_a) || hash_u01(krep, g) < (var ? p.p_rep_b : p.p_rep_a))
            act.push_back(Act{hash_u64(kpri, g), g, d});
    }
    std::sort(act.begin(), act.end(), [](const Act &a, const Act &b) { return a.key != b.key ? a.key < b.key : a.g < b.g; });
    for (const Act &a : act) react_one(s, M, step, ksyn, krep, kmu, a.slot, out);
}

static void sys_forces(OrSys *s, int64_t step, int theta32) {
    or_pair_forces(&s->p, s->N, s->x.data(), s->y.data(), s->vtx.data(), s->vty.data(), s->type.data(), s->gid.data(),
                   s->cstart.data(), step, theta32, 1, s->fx.data(), s->fy.data(), nullptr, nullptr);
    or_bonded(&s->p, s->N, s->x.data(), s->y.data(), s->mpos.data(), s->lnk.data(), s->gid.data(), s->fx.data(),
              s->fy.data(), &s->maxbond, nullptr);
}

static void load_gid_order(OrSys *s, const oc_state *st) {
    const int N = s->N;
    for (int g = 0; g < N; g++) {
        s->x[g] = st->x[g]; s->y[g] = st->y[g];
        s->vx[g] = st->vx[g]; s->vy[g] = st->vy[g];
        s->fx[g] = st->fx ? st->fx[g] : 0.0; s->fy[g] = st->fy ? st->fy[g] : 0.0;
        s->type[g] = st->type[g]; s->gid[g] = (uint32_t)g; s->mpos[g] = st->mpos[g];
        const int32_t nx = st->mpos[g] >= 0 ? st->nxt[g] : (st->bond ? st->bond[g] : -1);
        s->lnk[g] = nx;
    }
}

/* fresh start, cel/engine/oercel.c:2196-2212: v~ = v, sort, forces at step 0, vh = v - dt/2 f */
OR_API int or_start(OrSys *s, const oc_state *st, int32_t theta32) {
    load_gid_order(s, st);
    s->p.warm_end = s->p.step_anchor + s->p.warmup; /* g_warm_end = warmup, cel/engine/oercel.c:2197 */
    s->vtx = s->vx;
    s->vty = s->vy;
    s->step = s->p.step_anchor;
    sys_sort(s);
    sys_forces(s, s->step, theta32);
    for (int i = 0; i < s->N; i++) {
        s->vx[i] -= 0.5 * s->p.dt * s->fx[i];
        s->vy[i] -= 0.5 * s->p.dt * s->fy[i];
    }
    return 0;
}

/* restart-style load (vh and f as given) + sort; the predicted velocity is
 * vh rounded to FP32, the value the device keeps after oc_set_state */
OR_API int or_set_state(OrSys *s, const oc_state *st, int64_t step) {
    load_gid_order(s, st);
    for (int i = 0; i < s->N; i++) { s->vtx[i] = (double)(float)s->vx[i]; s->vty[i] = (double)(float)s->vy[i]; }
    s->step = step;
    sys_sort(s);
    return 0;
}

OR_API void or_get_state(OrSys *s, oc_state *st) {
    const int N = s->N;
    for (int d = 0; d < N; d++) {
        const uint32_t g = s->gid[d];
        if (st->x) st->x[g] = s->x[d];
        if (st->y) st->y[g] = s->y[d];
        if (st->vx) st->vx[g] = s->vx[d];
        if (st->vy) st->vy[g] = s->vy[d];
        if (st->fx) st->fx[g] = s->fx[d];
        if (st->fy) st->fy[g] = s->fy[d];
        if (st->type) st->type[g] = s->type[d];
        if (st->mpos) st->mpos[g] = s->mpos[d];
        const int32_t l = s->lnk[d];
        const int32_t lg = l < 0 ? -1 : (int32_t)s->gid[l];
        if (st->nxt) st->nxt[g] = s->mpos[d] >= 0 ? lg : -1;
        if (st->bond) st->bond[g] = s->mpos[d] < 0 ? lg : -1;
    }
}

/* n steps.  Without the reference chemistry: integrate, sort, precursor
 * chemistry, pair + bonded forces (the precursor chemistry moves no bead, so
 * running it on the fresh cell list makes the decisions it would make on the
 * previous one).  With it, the reference's order (cel/engine/oercel.c:2300-2362):
 * integrate, the reference chemistry on the previous cell index (after an
 * extra sort if a bead moved 0.5 or more), sort, then the precursor chemistry
 * and the forces. */
static void sys_rchem(OrSys *s, int64_t step) {
    sys_react(s, step, s->ev);
    std::vector<oc_event> tmp(s->N + 16);
    const int64_t ne = or_rchem(&s->p, s->N, step, s->x.data(), s->y.data(), s->type.data(), s->mpos.data(),
                                s->lnk.data(), s->gid.data(), tmp.data(), (int64_t)tmp.size());
    s->ev.insert(s->ev.end(), tmp.begin(), tmp.begin() + std::min<int64_t>(ne, (int64_t)tmp.size()));
}

OR_API void or_step(OrSys *s, int64_t n, int32_t theta32) {
    for (int64_t k = 0; k < n; k++) {
        const int64_t ns = s->step + 1;
        const bool rcheck = s->p.rchem && s->p.n_chem >= 1 && ns % s->p.n_chem == 0;
        double dmax = 0.0;
        if (rcheck) {
            const double dt = s->p.dt, dt2 = dt * dt;
            for (int i = 0; i < s->N; i++) {
                const double ddx = fabs(dt * s->vx[i] + dt2 * s->fx[i]), ddy = fabs(dt * s->vy[i] + dt2 * s->fy[i]);
                dmax = std::max(dmax, std::max(ddx, ddy));
            }
        }
        or_integrate(&s->p, s->N, s->x.data(), s->y.data(), s->vx.data(), s->vy.data(), s->fx.data(), s->fy.data(),
                     s->vtx.data(), s->vty.data());
        if (rcheck) {
            if (dmax >= kChemMargin) sys_sort(s);
            sys_rchem(s, ns);
        }
        sys_sort(s);
        if (s->p.chem) {
            std::vector<oc_event> tmp(s->N * 2 + 16);
            const int64_t ne = or_chem(&s->p, s->N, ns, s->x.data(), s->y.data(), s->type.data(), s->mpos.data(),
                                       s->lnk.data(), s->gid.data(), s->cstart.data(), tmp.data(), (int64_t)tmp.size());
            s->ev.insert(s->ev.end(), tmp.begin(), tmp.begin() + ne);
        }
        sys_forces(s, ns, theta32);
        s->step = ns;
    }
}

/* one chemistry check on the loaded (sorted) state (the reference chemistry:
 * reactions, decay, feed; then the precursor chemistry), events appended */
OR_API int64_t or_sys_chem(OrSys *s, int64_t step) {
    if (s->p.rchem && s->p.n_chem >= 1 && step % s->p.n_chem == 0) sys_rchem(s, step);
    std::vector<oc_event> tmp(s->N * 2 + 16);
    const int64_t ne = or_chem(&s->p, s->N, step, s->x.data(), s->y.data(), s->type.data(), s->mpos.data(),
                               s->lnk.data(), s->gid.data(), s->cstart.data(), tmp.data(), (int64_t)tmp.size());
    s->ev.insert(s->ev.end(), tmp.begin(), tmp.begin() + ne);
    return ne;
}

/* pair + bonded forces of the loaded state at `step` (the predicted velocity
 * is whatever the last set_state/step left: vh after set_state) */
OR_API void or_sys_forces(OrSys *s, int64_t step, int32_t theta32, int32_t with_bonded) {
    or_pair_forces(&s->p, s->N, s->x.data(), s->y.data(), s->vtx.data(), s->vty.data(), s->type.data(), s->gid.data(),
                   s->cstart.data(), step, theta32, 1, s->fx.data(), s->fy.data(), nullptr, nullptr);
    if (with_bonded)
        or_bonded(&s->p, s->N, s->x.data(), s->y.data(), s->mpos.data(), s->lnk.data(), s->gid.data(), s->fx.data(),
                  s->fy.data(), &s->maxbond, nullptr);
}

/* pending events, sorted by (step, kind, gid_a, gid_b) as the device reports them */
OR_API int64_t or_get_events(OrSys *s, oc_event *ev, int64_t cap) {
    std::stable_sort(s->ev.begin(), s->ev.end(), [](const oc_event &a, const oc_event &b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.gid_a != b.gid_a) return a.gid_a < b.gid_a;
        return a.gid_b < b.gid_b;
    });
    const int64_t n = (int64_t)s->ev.size();
    for (int64_t k = 0; k < n && k < cap; k++) ev[k] = s->ev[k];
    s->ev.clear();
    return n;
}

OR_API void or_get_cells(OrSys *s, int32_t *cstart, uint32_t *gid_of_slot) {
    if (cstart) memcpy(cstart, s->cstart.data(), sizeof(int32_t) * (s->g.ncell + 1));
    if (gid_of_slot) memcpy(gid_of_slot, s->gid.data(), sizeof(uint32_t) * s->N);
}

/* Conservative virial sum_{i<j} r_ij . F^C_ij = sum_{i<j} a_ij (1 - r) r over
 * pairs within r < 1 (FP64, minimum image; every pair visited from both
 * sides and halved), the sum of the reference's water check
 * (cel/engine_gpu/tests/compare.py:82-110, there for a_WW only). */
static double pair_virial(const oc_params *p, int N, const double *x, const double *y, const uint8_t *type,
                          const int32_t *cstart, int64_t step) {
    const Geo g = make_geo(p, N);
    const PairPar pp = pair_par(p, step, time_of(p, step));
    const int ncx = g.ncx, ncy = g.ncy;
    const double L = g.L;
    double vir = 0.0;
    for (int cx = 0; cx < ncx; cx++)
        for (int i = cstart[cx * ncy]; i < cstart[(cx + 1) * ncy]; i++) {
            const int khi = (int)((y[i] + 1.0) * g.icsy);
            const int klo = (int)((y[i] - 1.0) * g.icsy + ncy) - ncy;
            for (int dc = -1; dc <= 1; dc++) {
                const int col = (cx + dc + ncx) % ncx;
                for (int k = klo; k <= khi; k++) {
                    const int row = (k + ncy) % ncy;
                    for (int j = cstart[col * ncy + row]; j < cstart[col * ncy + row + 1]; j++) {
                        if (j == i) continue;
                        double dx = x[i] - x[j], dy = y[i] - y[j];
                        dx -= L * std::round(dx / L);
                        dy -= L * std::round(dy / L);
                        const double r = sqrt(dx * dx + dy * dy);
                        if (r < 1.0) vir += 0.5 * pp.a[8 * type[i] + type[j]] * (1.0 - r) * r;
                    }
                }
            }
        }
    return vir;
}

/* kinetic_T and count_types, cel/engine/oercel.c:2053-2071, plus energies,
 * the conservative virial and the 2D virial pressure */
OR_API void or_stats(OrSys *s, oc_stats *o) {
    const int N = s->N;
    const double h = 0.5 * s->p.dt;
    double k2 = 0, sx = 0, sy = 0;
    memset(o, 0, sizeof *o);
    for (int i = 0; i < N; i++) {
        const double vx = s->vx[i] + h * s->fx[i], vy = s->vy[i] + h * s->fy[i];
        k2 += vx * vx + vy * vy;
        sx += vx;
        sy += vy;
        o->count[s->type[i]]++;
        o->n_lipids += s->mpos[i] == 0;
        o->n_bonds += (s->mpos[i] < 0 && s->lnk[i] >= 0 && s->gid[i] < s->gid[s->lnk[i]]);
    }
    o->step = s->step;
    o->time = time_of(&s->p, s->step);
    o->T_kin = k2 / (2.0 * N - 2.0);
    o->px = sx;
    o->py = sy;
    o->E_kin = 0.5 * k2;
    std::vector<double> ep(N), fx(N, 0.0), fy(N, 0.0);
    or_pair_forces(&s->p, N, s->x.data(), s->y.data(), s->vtx.data(), s->vty.data(), s->type.data(), s->gid.data(),
                   s->cstart.data(), s->step, 1, 0, fx.data(), fy.data(), ep.data(), nullptr);
    double e = 0;
    for (int i = 0; i < N; i++) e += ep[i];
    o->E_pair = e;
    double mb = 0, eb = 0;
    or_bonded(&s->p, N, s->x.data(), s->y.data(), s->mpos.data(), s->lnk.data(), s->gid.data(), fx.data(), fy.data(), &mb,
              &eb);
    o->E_bond = eb;
    o->max_bond = s->maxbond;
    s->maxbond = 0;
    o->virial = pair_virial(&s->p, N, s->x.data(), s->y.data(), s->type.data(), s->cstart.data(), s->step);
    o->pressure = (double)N / (s->p.L * s->p.L) * o->T_kin + o->virial / (2.0 * s->p.L * s->p.L);
}

OR_API int64_t or_get_step(OrSys *s) { return s->step; }

/* ------------------------------------------------------------------ adjoint (own definition, FP64 reference gradient) */

/*
 * Gradient of a smooth radial-distribution loss over a window of steps with
 * respect to theta = (a_WW, a_WH, a_WT, a_HH, a_HT, a_TT, gamma), the noise
 * stream held fixed (no reference counterpart; BASELINE adjoint window).
 *
 * Window: start state (x0, vh0, v~0) loaded in the system; f0 = F(x0, v~0;
 * theta) is recomputed, then nsteps steps of the Groot-Warren update
 * (or_integrate + sort + forces; chemistry must be off).
 * RDF of type pair (ta, tb) on nbins bins of [0, rmax): a triweight kernel
 * K of half-width `width` per pair,
 *   g_b(x) = C_b sum_{i<j} K(r_ij - c_b),  C_b = dr * norm / (pi (r_{b+1}^2 - r_b^2)),
 *   norm = 2 / (n_a rho_a) for ta == tb, 1 / (n_a rho_b) otherwise,
 * averaged over steps 1..nsteps; loss = sum_b (gbar_b - target_b)^2.
 * Reverse sweep (f_n = F(x_n, v~_n)):
 *   X_n += Jx^T F_n ; T_n = Jv^T F_n ; theta_bar += Jtheta^T F_n
 *   X_{n-1} = X_n + dL/dx_{n-1};  V_{n-1} = dt X_n + T_n + V_n;
 *   F_{n-1} = dt^2 X_n + (1/2 + lambda) dt T_n + dt V_n
 * with adjoints carried through the sort permutations.  Pair Jacobians use
 * f_ij = m e, m = w (a - gamma w (e.u) + s theta_ij), s = sqrt(2 gamma kT / dt);
 * bonds f = -k (r - r0)/r d.  Angle terms, stirring, the kT cycle and the
 * warm-up ramp must be off.
 */
struct RdfSpec {
    int32_t ta, tb, nbins, pad;
    double rmax, width;
};

static inline int theta_index(int a, int b) {
    if (a > b) { const int t = a; a = b; b = t; }
    if (b > 2) return -1;
    static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return idx[a][b];
}

static void rdf_norm(const RdfSpec &sp, int na, int nb, double L, std::vector<double> &C) {
    const double dr = sp.rmax / sp.nbins, area = L * L;
    const double norm = sp.ta == sp.tb ? 2.0 / (na * (na / area)) : 1.0 / (na * (nb / area));
    C.resize(sp.nbins);
    for (int b = 0; b < sp.nbins; b++) {
        const double r0 = b * dr, r1 = (b + 1) * dr;
        C[b] = dr * norm / (M_PI * (r1 * r1 - r0 * r0));
    }
}

/* all-pairs smooth RDF (and, with gw != null, X += dL/dx for bin weights gw).
 * Bin kernel: triweight K(u) = 35/(32 h) (1 - (u/h)^2)^3 on |u| < h
 * (h = spec width; integrates to 1, twice differentiable at its edge). */
static void rdf_pass(const OrSys *s, const RdfSpec &sp, const std::vector<double> &C, double *g, const double *gw,
                     double *X) {
    const int N = s->N;
    const double L = s->p.L, dr = sp.rmax / sp.nbins, h = sp.width, cut = sp.rmax + h;
    const double pref = 35.0 / (32.0 * h);
    for (int i = 0; i < N; i++) {
        const int ti = s->type[i];
        if (ti != sp.ta && ti != sp.tb) continue;
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            const int tj = s->type[j];
            const bool pr = (ti == sp.ta && tj == sp.tb) || (ti == sp.tb && tj == sp.ta);
            if (!pr) continue;
            const double dx = minimg(s->x[i] - s->x[j], L), dy = minimg(s->y[i] - s->y[j], L);
            const double r = sqrt(dx * dx + dy * dy);
            if (r >= cut || r == 0.0) continue;
            const int b0 = std::max(0, (int)floor((r - h) / dr)), b1 = std::min(sp.nbins - 1, (int)floor((r + h) / dr));
            double dsum = 0.0;
            for (int b = b0; b <= b1; b++) {
                const double t = (r - (b + 0.5) * dr) / h;
                if (fabs(t) >= 1.0) continue;
                const double q = 1.0 - t * t;
                /* each unordered pair once for g (ta == tb: i < j) */
                if (g && (sp.ta != sp.tb ? (ti == sp.ta) : (i < j))) g[b] += C[b] * pref * q * q * q;
                if (gw) dsum += gw[b] * C[b] * pref * 3.0 * q * q * (-2.0 * t / h);
            }
            if (gw) {
                /* d/dx_i of the pair's terms (counted once per unordered pair) */
                X[2 * i] += dsum * dx / r;
                X[2 * i + 1] += dsum * dy / r;
            }
        }
    }
}

/* J^T products of the pair + bond forces at the current state for force adjoint F:
 * X += Jx^T F, T = Jv^T F, th += Jtheta^T F */
static void force_adjoint(const OrSys *s, int64_t step, int theta32, const double *F, double *X, double *T, double *th) {
    const int N = s->N;
    const oc_params *p = &s->p;
    const Geo &g = s->g;
    const PairPar pp = pair_par(p, step, time_of(p, step));
    const int ncx = g.ncx, ncy = g.ncy;
    const double L = g.L, sdg = pp.sdt / (2.0 * p->gamma);
    for (int i = 0; i < 2 * N; i++) T[i] = 0.0;
    for (int cx = 0; cx < ncx; cx++) {
        for (int i = s->cstart[cx * ncy]; i < s->cstart[(cx + 1) * ncy]; i++) {
            const double xi = s->x[i], yi = s->y[i];
            const int khi = (int)((yi + 1.0) * g.icsy);
            const int klo = (int)((yi - 1.0) * g.icsy + ncy) - ncy;
            double xa = 0, ya = 0, ta = 0, tb = 0, tha[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int dc = -1; dc <= 1; dc++) {
                int col = cx + dc;
                double sx = 0.0;
                if (col < 0) { col += ncx; sx = -L; }
                else if (col >= ncx) { col -= ncx; sx = L; }
                for (int k = klo; k <= khi; k++) {
                    const int kk = k < 0 ? k + ncy : (k >= ncy ? k - ncy : k);
                    const double sy = k < 0 ? -L : (k >= ncy ? L : 0.0);
                    const int c = col * ncy + kk;
                    for (int j = s->cstart[c]; j < s->cstart[c + 1]; j++) {
                        const double dx = (xi - s->x[j]) - sx, dy = (yi - s->y[j]) - sy;
                        const double r2 = dx * dx + dy * dy;
                        if (!(r2 < 1.0 && r2 > 1e-24)) continue;
                        const double r = sqrt(r2), w = 1.0 - r;
                        const double ex = dx / r, ey = dy / r;
                        const double ux = s->vtx[i] - s->vtx[j], uy = s->vty[i] - s->vty[j];
                        const double eu = ex * ux + ey * uy;
                        const double th = theta32 ? pair_theta32(pp.sseed, s->gid[i], s->gid[j])
                                                  : pair_theta(pp.sseed, s->gid[i], s->gid[j]);
                        const int ti = s->type[i], tj = s->type[j];
                        const double a = pp.a[8 * ti + tj];
                        const double m = w * (a - pp.gamma * w * eu + pp.sdt * th);
                        const double dmr = -(a + pp.sdt * th) + 2.0 * pp.gamma * w * eu;
                        const double dmu = -pp.gamma * w * w;
                        const double lx = F[2 * i] - F[2 * j], ly = F[2 * i + 1] - F[2 * j + 1];
                        const double el = ex * lx + ey * ly;
                        const double pux = ux - eu * ex, puy = uy - eu * ey; /* P u */
                        const double plx = lx - el * ex, ply = ly - el * ey; /* P lambda */
                        xa += (dmr * ex + dmu * pux / r) * el + m * plx / r;
                        ya += (dmr * ey + dmu * puy / r) * el + m * ply / r;
                        ta += dmu * el * ex;
                        tb += dmu * el * ey;
                        const int q = theta_index(ti, tj);
                        if (q >= 0) tha[q] += 0.5 * w * el;
                        tha[6] += 0.5 * el * (-w * w * eu + sdg * w * th);
                    }
                }
            }
            X[2 * i] += xa;
            X[2 * i + 1] += ya;
            T[2 * i] = ta;
            T[2 * i + 1] = tb;
            for (int q = 0; q < 7; q++) th[q] += tha[q];
        }
    }
    /* harmonic bonds (lipid chains and chemistry bonds), d = x_j - x_i */
    auto bond_adj = [&](int i, int j) {
        const double dx = minimg(s->x[j] - s->x[i], L), dy = minimg(s->y[j] - s->y[i], L);
        const double r = sqrt(dx * dx + dy * dy);
        const double gg = -p->k_bond * (r - p->r0) / r, gp = -p->k_bond * p->r0 / (r * r);
        const double ex = dx / r, ey = dy / r;
        const double lx = F[2 * j] - F[2 * i], ly = F[2 * j + 1] - F[2 * i + 1];
        const double el = ex * lx + ey * ly;
        const double bx = gg * lx + gp * r * el * ex, by = gg * ly + gp * r * el * ey;
        X[2 * j] += bx;
        X[2 * j + 1] += by;
        X[2 * i] -= bx;
        X[2 * i + 1] -= by;
    };
    const int ml = p->lip_nH + p->lip_nT;
    for (int h = 0; h < N; h++) {
        if (s->mpos[h] == 0) {
            int prev = h, n = 1;
            for (int d = s->lnk[h]; d >= 0 && n < ml; d = s->lnk[d], n++) { bond_adj(prev, d); prev = d; }
        } else if (s->mpos[h] < 0 && s->lnk[h] >= 0 && s->gid[h] < s->gid[s->lnk[h]]) {
            bond_adj(h, s->lnk[h]);
        }
    }
}

struct TapeStep {
    std::vector<double> x, y, vtx, vty;
    std::vector<uint8_t> type;
    std::vector<uint32_t> gid;
    std::vector<int8_t> mpos;
    std::vector<int32_t> lnk, cstart, src;
};

static void tape_load(OrSys *s, const TapeStep &t) {
    s->x = t.x; s->y = t.y; s->vtx = t.vtx; s->vty = t.vty; s->type = t.type; s->gid = t.gid;
    s->mpos = t.mpos; s->lnk = t.lnk; s->cstart = t.cstart;
}

/* Window gradient on the loaded state (which the call advances by nsteps).
 * out: loss, grad[7], gmean[nbins].  Returns 0, or -1 if the physics is not
 * covered (angles, stirring, kT cycle, ramp or chemistry on). */
OR_API int or_window_grad(OrSys *s, int32_t nsteps, const RdfSpec *spec, const double *target, int32_t theta32,
                          double *loss, double *grad, double *gmean) {
    const oc_params *p = &s->p;
    if (p->k_angle != 0.0 || p->stir_A != 0.0 || p->kT_amp != 0.0 || p->chem || s->step < p->warm_end ||
        p->lip_arch != OC_ARCH_LINEAR)
        return -1;
    const RdfSpec sp = *spec;
    const int N = s->N, nb = sp.nbins;
    int na = 0, nbb = 0;
    for (int i = 0; i < N; i++) { na += s->type[i] == sp.ta; nbb += s->type[i] == sp.tb; }
    std::vector<double> C;
    rdf_norm(sp, na, nbb, p->L, C);
    /* forward with tape: state 0 (after recomputing f0), then steps 1..n */
    std::vector<TapeStep> tape(nsteps + 1);
    auto record = [&](int n) {
        TapeStep &t = tape[n];
        t.x = s->x; t.y = s->y; t.vtx = s->vtx; t.vty = s->vty; t.type = s->type; t.gid = s->gid;
        t.mpos = s->mpos; t.lnk = s->lnk; t.cstart = s->cstart; t.src = s->src;
    };
    sys_forces(s, s->step, theta32);
    record(0);
    std::vector<double> gacc(nb, 0.0);
    const int64_t step0 = s->step;
    for (int n = 1; n <= nsteps; n++) {
        or_step(s, 1, theta32);
        record(n);
        rdf_pass(s, sp, C, gacc.data(), nullptr, nullptr);
    }
    double Lsum = 0.0;
    std::vector<double> gw(nb);
    for (int b = 0; b < nb; b++) {
        gmean[b] = gacc[b] / nsteps;
        const double d = gmean[b] - target[b];
        Lsum += d * d;
        gw[b] = 2.0 * d / nsteps;
    }
    *loss = Lsum;
    /* reverse sweep; the live system is restored to the window end afterwards */
    OrSys saved = *s;
    std::vector<double> X(2 * N, 0.0), V(2 * N, 0.0), F(2 * N, 0.0), T(2 * N, 0.0);
    std::vector<double> th(7, 0.0);
    const double dt = p->dt, c = (0.5 + p->lambda) * dt;
    tape_load(s, tape[nsteps]);
    rdf_pass(s, sp, C, nullptr, gw.data(), X.data());
    for (int n = nsteps; n >= 1; n--) {
        tape_load(s, tape[n]);
        force_adjoint(s, step0 + n, theta32, F.data(), X.data(), T.data(), th.data());
        std::vector<double> Xp(2 * N), Vp(2 * N), Fp(2 * N);
        for (int k = 0; k < 2 * N; k++) {
            Xp[k] = X[k];
            Vp[k] = dt * X[k] + T[k] + V[k];
            Fp[k] = dt * dt * X[k] + c * T[k] + dt * V[k];
        }
        const std::vector<int32_t> &src = tape[n].src;
        for (int d = 0; d < N; d++) {
            const int o = src[d];
            X[2 * o] = Xp[2 * d]; X[2 * o + 1] = Xp[2 * d + 1];
            V[2 * o] = Vp[2 * d]; V[2 * o + 1] = Vp[2 * d + 1];
            F[2 * o] = Fp[2 * d]; F[2 * o + 1] = Fp[2 * d + 1];
        }
        if (n - 1 >= 1) {
            tape_load(s, tape[n - 1]);
            rdf_pass(s, sp, C, nullptr, gw.data(), X.data());
        }
    }
    tape_load(s, tape[0]);
    force_adjoint(s, step0, theta32, F.data(), X.data(), T.data(), th.data());
    for (int q = 0; q < 7; q++) grad[q] = th[q];
    *s = saved;
    return 0;
}

/* smooth RDF of the loaded state (for targets) */
OR_API void or_rdf(OrSys *s, const RdfSpec *spec, double *g) {
    const RdfSpec sp = *spec;
    int na = 0, nbb = 0;
    for (int i = 0; i < s->N; i++) { na += s->type[i] == sp.ta; nbb += s->type[i] == sp.tb; }
    std::vector<double> C;
    rdf_norm(sp, na, nbb, s->p.L, C);
    for (int b = 0; b < sp.nbins; b++) g[b] = 0.0;
    rdf_pass(s, sp, C, g, nullptr, nullptr);
}

/* change the pair parameters of the loaded system (a_ij symmetric 8x8, gamma) */
OR_API void or_set_pair_params(OrSys *s, const double *a64, double gamma) {
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) s->p.a[i][j] = a64[8 * i + j];
    s->p.gamma = gamma;
}

