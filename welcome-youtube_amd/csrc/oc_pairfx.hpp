// oc_pairfx.hpp -- the pair-force kernel on fixed-point periodic coordinates
// with a per-lane interaction list (pair variant 3).
//
// Same tiles, staging geometry and results as k_pair (full-shell DPD pair
// forces of cel/engine/oercel.c:1054-1120 with the bonded terms of each bead
// in gather form and the next step's cell key in the epilogue), organised for
// the VALU issue rate that bounds the kernel:
//   - positions are staged as 32-bit fixed-point periodic coordinates
//     X = floor(x 2^32 / L) mod 2^32.  The difference of two such words,
//     read as a signed integer, is the minimum-image separation in units of
//     L / 2^32, so every wavefront (seam tiles included) walks its window
//     without image shifts, and the separation of (j, i) is the exact
//     negation of that of (i, j): f_ji = -f_ij bit for bit.  A position
//     moves by less than the quantum L / 2^32 (1.4e-7 at L = 591, 5.5e-7 at
//     L = 2365); the pair arithmetic after the cutoff test scales the
//     separation in FP64 and rounds it once to FP32, as k_pair does.
//   - the per-gid factors of the pair hash's first multiplication are staged
//     with each bead (P(g) = g phi_lo, Q(g) = g phi_hi + mulhi(g, phi_lo),
//     mod 2^32, for the (min gid << 32 | max gid) * phi of mix13's input), so a
//     pair pays six 32-bit multiplies instead of nine and the same bits come
//     out;
//   - each lane first runs the cutoff test over its window and lists its
//     interacting candidates (16-bit staged indices) in LDS, then evaluates
//     the listed pairs: the wavefront runs the force arithmetic ~max over
//     lanes of the interaction count (~17 at rho = 3) instead of ~max over
//     lanes of the window length (~31).  A lane with more than KL
//     interactions evaluates the rest in place during the test.
// Forces are FP64 sums of FP32 pair terms per bead, no atomics:
// deterministic, and the summation order of a bead is fixed by its window.
#pragma once
#include "oc_kernels.hpp"

namespace oc {

constexpr uint32_t kPhiLo = 0x7f4a7c15u, kPhiHi = 0x9e3779b9u;  // 0x9e3779b97f4a7c15
#ifndef OC_PAIR_LIST
#define OC_PAIR_LIST 16
#endif
constexpr int kPairList = OC_PAIR_LIST;  // listed interactions per lane (16-bit entries)

// fixed-point periodic coordinate of x in [0, L): sc = 2^32 / L
__device__ __forceinline__ uint32_t fix_coord(double x, double sc) { return (uint32_t)(uint64_t)(x * sc); }

// factors of the pair hash's first multiplication of gid g
__device__ __forceinline__ uint2 hash_pre(uint32_t g) {
    return make_uint2(g * kPhiLo, g * kPhiHi + __umulhi(g, kPhiLo));
}

// theta word of pair_hash(key, gi, gj) >> 32 from the staged factors:
// (lo << 32 | hi) * phi = (P(lo) + Q(hi)) << 32 | P(hi)  (mod 2^64)
__device__ __forceinline__ int32_t pair_theta_pre(uint64_t key, uint32_t gi, uint2 hi_, uint32_t gj, uint2 hj) {
    const bool ilo = gi < gj;
    const uint32_t plo = ilo ? hi_.x : hj.x, phi_ = ilo ? hj.x : hi_.x, qhi = ilo ? hj.y : hi_.y;
    uint64_t z = (((uint64_t)(plo + qhi) << 32) | phi_) ^ key;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return (int32_t)((uint32_t)(z >> 32) ^ (uint32_t)(z >> 63));
}

struct FxBead {
    uint2 q;    // fixed-point X, Y
    float2 v;   // predicted velocity
    uint32_t g; // gid | type << 29
    uint2 h;    // hash factors of the gid
};

// FP32 force of the pair (i, j) from the fixed-point separation (dxi, dyi)
// (quantum units, 0 < r < 1): F_ij = fm (dx, dy), fm = w (a_ij - gamma w
// r.v + sigma/sqrt(dt) theta_ij) / r, w = 1 - r, the arithmetic of pair_core.
// The separation is scaled in FP64 and rounded once to FP32 (as k_pair rounds
// its FP64 difference), so near the cutoff w = 1 - r carries no extra
// rounding of the scale.
__device__ __forceinline__ void pair_fx(int32_t dxi, int32_t dyi, double isc, float2 vi, uint32_t gi, uint2 hi_,
                                        const float *arow, const FxBead &J, float gam, float sdtk, uint64_t sseed,
                                        double &ax, double &ay) {
    const float dx = (float)((double)dxi * isc), dy = (float)((double)dyi * isc);
    const float r2 = __builtin_fmaf(dx, dx, dy * dy);
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ir;
    const float w = 1.0f - r;
    const float dvx = vi.x - J.v.x, dvy = vi.y - J.v.y;
    const float rv = __builtin_fmaf(dx, dvx, dy * dvy) * ir;
    const float th = (float)pair_theta_pre(sseed, gi, hi_, J.g & kGidMask, J.h);
    const float a = arow[J.g >> kGidBits];
    const float fm = w * __builtin_fmaf(sdtk, th, __builtin_fmaf(-gam * w, rv, a)) * ir;
    ax += (double)(fm * dx);
    ay += (double)(fm * dy);
}

template <int BLOCK, int CAP, int KL, bool ANG>
__global__ __launch_bounds__(BLOCK) void k_pair_fx(PairArgs A) {
    static_assert(CAP <= 65536, "list entries are 16-bit staged indices");
    constexpr int kMaxRows = pair_max_rows(BLOCK);
    __shared__ uint2 s_q[CAP];
    __shared__ float2 s_vt[CAP];
    __shared__ uint32_t s_gt[CAP];
    __shared__ uint2 s_h[CAP];
    __shared__ uint16_t s_lst[KL * BLOCK];
    __shared__ int32_t s_row[3][kMaxRows + 1];
    __shared__ float s_a[64];
    __shared__ int32_t s_piece[3][3][3];  // [col][piece A/B/C] = {first slot, length, lds start}
    __shared__ int32_t s_colbase[4];
    __shared__ int32_t s_col[3];

    // XCD-aware tile order, as k_pair
    const int t = (blockIdx.x & 7) * A.per_xcd + (blockIdx.x >> 3);
    if (t >= A.ntiles) return;
    const Tile T = A.tiles[t];
    const RepInfo &R = A.ri[T.rep];
    const RepStep &S = A.rs[T.rep];
    const int ncy = R.ncy, ncx = R.ncx;
    const int M = (int)ceil(R.icsy) + 1;  // staging margin in rows
    const int kb0 = T.k0 - M;             // first staged (unwrapped) row
    const int nrow = T.k1 - T.k0 + 2 * M + 1;
    const int32_t *cs = A.cs;
    const double sc = 4294967296.0 / R.L;

    // pieces of the staged rows (lanes q < 9 of wavefront 0), as k_pair
    if (threadIdx.x < 64) {
        const float av = S.a_eff[threadIdx.x];
        const int q = threadIdx.x;
        const int lo = kb0, hi = T.k1 - 1 + M;
        int f = 0, n = 0;
        if (q < 9) {
            const int c = q / 3, pc = q - 3 * c;
            int col = T.cx - 1 + c;
            col = col < 0 ? col + ncx : (col >= ncx ? col - ncx : col);
            const int cb = R.coff + col * ncy;
            int ra = 0, rb = -1;
            if (pc == 0) {
                if (lo < 0) { ra = lo + ncy; rb = ncy - 1; }
            } else if (pc == 1) {
                ra = max(lo, 0); rb = min(hi, ncy - 1);
            } else if (hi >= ncy) {
                ra = 0; rb = hi - ncy;
            }
            if (rb >= ra) {
                f = cs[cb + ra];
                n = cs[cb + rb + 1] - f;
            }
            if (pc == 0) s_col[c] = col;
        }
        int incl = n;
        for (int o = 1; o < 16; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (q >= o) incl += v;
        }
        if (q < 9) {
            const int c = q / 3;
            s_piece[c][q - 3 * c][0] = f;
            s_piece[c][q - 3 * c][1] = n;
            s_piece[c][q - 3 * c][2] = incl - n;
            if (q == 3 * c) s_colbase[c] = incl - n;
            if (q == 8) s_colbase[3] = incl;
        }
        s_a[q] = av;
    }
    __syncthreads();
    const int total = s_colbase[3];
    const bool staged = total <= CAP && nrow <= kMaxRows;

    // row offsets and the staging (every load of a thread before its first
    // LDS store)
    {
        constexpr int kRowPer = (3 * (kMaxRows + 1) + BLOCK - 1) / BLOCK;
        constexpr int kStgPer = (CAP + BLOCK - 1) / BLOCK;
        const int nrow3 = nrow <= kMaxRows ? 3 * nrow : 0;
        int pb[9], pf[9];
#pragma unroll
        for (int w = 0; w < 9; w++) {
            pb[w] = s_piece[w / 3][w % 3][2];
            pf[w] = s_piece[w / 3][w % 3][0];
        }
        int rcs[kRowPer];
#pragma unroll
        for (int u = 0; u < kRowPer; u++) {
            const int e = (int)threadIdx.x + u * BLOCK;
            int idx = R.coff;
            if (e < nrow3) {
                const int c = e >= 2 * nrow ? 2 : (e >= nrow ? 1 : 0);
                const int k = kb0 + (e - c * nrow);
                idx = R.coff + s_col[c] * ncy + (k < 0 ? k + ncy : (k < ncy ? k : k - ncy));
            }
            rcs[u] = cs[idx];
        }
        double2 sx[kStgPer];
        float2 sv[kStgPer];
        uint32_t sg[kStgPer];
        int sd[kStgPer];
#pragma unroll
        for (int u = 0; u < kStgPer; u++) {
            const int e = (int)threadIdx.x + u * BLOCK;
            int src = R.off;
            sd[u] = -1;
            if (staged && e < total) {
                int q = 0;
#pragma unroll
                for (int w = 1; w < 9; w++) q = e >= pb[w] ? w : q;
                int fq = pf[0], bq = pb[0];
#pragma unroll
                for (int w = 1; w < 9; w++) {
                    fq = q == w ? pf[w] : fq;
                    bq = q == w ? pb[w] : bq;
                }
                src = fq + (e - bq);
                sd[u] = e;
            }
            sx[u] = A.xy[src];
            sv[u] = A.vt[src];
            sg[u] = A.gt[src];
        }
#pragma unroll
        for (int u = 0; u < kRowPer; u++) {
            const int e = (int)threadIdx.x + u * BLOCK;
            if (e < nrow3) {
                const int c = e >= 2 * nrow ? 2 : (e >= nrow ? 1 : 0);
                const int k = kb0 + (e - c * nrow);
                const int pc = k < 0 ? 0 : (k < ncy ? 1 : 2);
                s_row[c][e - c * nrow] = pb[3 * c + pc] + (rcs[u] - pf[3 * c + pc]);
            }
        }
#pragma unroll
        for (int u = 0; u < kStgPer; u++)
            if (sd[u] >= 0) {
                s_q[sd[u]] = make_uint2(fix_coord(sx[u].x, sc), fix_coord(sx[u].y, sc));
                s_vt[sd[u]] = sv[u];
                s_gt[sd[u]] = sg[u];
                s_h[sd[u]] = hash_pre(sg[u] & kGidMask);
            }
    }
    __syncthreads();
    if (nrow > kMaxRows) {  // the host sizes tiles so that this cannot happen
        if (threadIdx.x == 0) atomicOr(A.err, ERR_TILE);
        return;
    }

    const float gam = S.gamma, sdtk = S.sdt * kTheta32;
    const uint64_t sseed = S.sseed;
    const double isc = R.L / 4294967296.0;
    const float rq2 = (float)(sc * sc);  // r < 1 in quantum units
    const int ibeg = s_row[1][T.k0 - kb0], iend = s_row[1][T.k1 - kb0];
    const int slot0 = s_piece[1][1][0] + (ibeg - s_piece[1][1][2]);  // T.k0 lies in piece B of column 1
    double mb = 0.0;  // largest bond length seen by this lane

    // bead j of the window: staged (LDS index) or, for a tile that overflows
    // the buffer, its slot in global memory
    auto bead = [&](int j, bool stg) {
        FxBead B;
        if (stg) {
            B.q = s_q[j];
            B.v = s_vt[j];
            B.g = s_gt[j];
            B.h = s_h[j];
        } else {
            const double2 p = A.xy[j];
            B.q = make_uint2(fix_coord(p.x, sc), fix_coord(p.y, sc));
            B.v = A.vt[j];
            B.g = A.gt[j];
            B.h = hash_pre(B.g & kGidMask);
        }
        return B;
    };

    auto run = [&](const bool stg) {
        for (int ib = ibeg; ib < iend; ib += BLOCK) {
            const int ii = ib + (int)threadIdx.x;
            const bool act = ii < iend;
            const int slot = slot0 + ((act ? ii : ibeg) - ibeg);
            BondPre Bd;
            const bool bond = A.bonded && act;
            if (bond) bond_links(slot, A.mp, A.lnk, A.prv, Bd);
            const double2 vh = A.key && act ? A.v[slot] : make_double2(0.0, 0.0);
            const double2 pi = A.xy[slot];
            const FxBead I = bead(stg ? (act ? ii : ibeg) : slot, stg);
            const uint32_t gi = I.g & kGidMask;
            const float *arow = s_a + 8 * (I.g >> kGidBits);
            int khi = (int)((pi.y + 1.0) * R.icsy);
            int klo = (int)((pi.y - 1.0) * R.icsy + ncy) - ncy;
            if (klo < kb0 || khi > T.k1 - 1 + M) {  // cannot happen for rows k0..k1-1 (margin M); guard
                atomicOr(A.err, ERR_TILE);
                klo = max(klo, kb0);
                khi = min(khi, T.k1 - 1 + M);
            }
            double ax = 0.0, ay = 0.0;
            // cutoff test of candidate j; an interacting one goes to the list
            // (or, past KL entries, is evaluated at once)
            auto test = [&](int j, int &cnt) {
                const FxBead J = bead(j, stg);
                const int32_t dxi = (int32_t)(I.q.x - J.q.x), dyi = (int32_t)(I.q.y - J.q.y);
                const float fx = (float)dxi, fy = (float)dyi;
                const float r2q = __builtin_fmaf(fx, fx, fy * fy);
                if (r2q < rq2 && r2q > 0.0f) {
                    if (stg && cnt < KL) s_lst[cnt * BLOCK + threadIdx.x] = (uint16_t)j;
                    else pair_fx(dxi, dyi, isc, I.v, gi, I.h, arow, J, gam, sdtk, sseed, ax, ay);
                    ++cnt;
                }
            };
            int cnt = 0;
            if (stg) {
                // the three column ranges of the window as one sequence
                const int st0 = s_row[0][klo - kb0], en0 = s_row[0][khi + 1 - kb0];
                const int st1 = s_row[1][klo - kb0], en1 = s_row[1][khi + 1 - kb0];
                const int st2 = s_row[2][klo - kb0], en2 = s_row[2][khi + 1 - kb0];
                int j = st0 == en0 ? st1 : st0;
                j = j == en1 ? st2 : j;
                const int nj = act ? max(en0 - st0, 0) + max(en1 - st1, 0) + max(en2 - st2, 0) : 0;
                if (act && (st1 < en0 || st2 < en1)) atomicOr(A.err, ERR_TILE);
                // the next candidate's coordinates are loaded before the
                // current one is tested (indices clamped into the buffer)
                uint2 qn = s_q[min(j, CAP - 1)];
                for (int it = 0; it < nj; it++) {
                    const uint2 q = qn;
                    const int jc = j;
                    ++j;
                    j = j == en0 ? st1 : j;
                    j = j == en1 ? st2 : j;
                    qn = s_q[min(j, CAP - 1)];
                    const int32_t dxi = (int32_t)(I.q.x - q.x), dyi = (int32_t)(I.q.y - q.y);
                    const float fx = (float)dxi, fy = (float)dyi;
                    const float r2q = __builtin_fmaf(fx, fx, fy * fy);
                    if (r2q < rq2 && r2q > 0.0f) {
                        if (cnt < KL) s_lst[cnt * BLOCK + threadIdx.x] = (uint16_t)jc;
                        else pair_fx(dxi, dyi, isc, I.v, gi, I.h, arow, bead(jc, true), gam, sdtk, sseed, ax, ay);
                        ++cnt;
                    }
                }
                // the listed pairs, the next entry's bead loaded before the
                // current pair is evaluated
                const int nl = min(cnt, KL);
                FxBead Jn = bead(nl > 0 ? min((int)s_lst[threadIdx.x], CAP - 1) : 0, true);
                for (int k = 0; k < nl; k++) {
                    const FxBead J = Jn;
                    const int kn = k + 1 < nl ? k + 1 : k;
                    Jn = bead(min((int)s_lst[kn * BLOCK + threadIdx.x], CAP - 1), true);
                    pair_fx((int32_t)(I.q.x - J.q.x), (int32_t)(I.q.y - J.q.y), isc, I.v, gi, I.h, arow, J, gam,
                            sdtk, sseed, ax, ay);
                }
            } else if (act) {
                // per column and periodic piece, slots from global memory
                for (int c = 0; c < 3; c++)
                    for (int q = 0; q < 3; q++) {
                        int pa, pb2;
                        if (q == 0) { pa = klo; pb2 = min(khi, -1); }
                        else if (q == 1) { pa = max(klo, 0); pb2 = min(khi, ncy - 1); }
                        else { pa = max(klo, ncy); pb2 = khi; }
                        if (pa > pb2) continue;
                        const int adj = s_piece[c][q][0] - s_piece[c][q][2];
                        const int ja = s_row[c][pa - kb0] + adj, jb = s_row[c][pb2 + 1 - kb0] + adj;
                        for (int jj = ja; jj < jb; jj++) test(jj, cnt);
                    }
            }
            if (act) {
                const bool formed = bond && chem_pending(A, slot, Bd);
                if (bond) bond_partners(A.xy, A.gt, Bd);
                if (formed) chem_formed(A, S, T.rep, gi, Bd);
                double bx = 0.0;
                if (A.stir) bx = A.stir[slot] - S.stir_mean;
                double fxs = bx + ax, fys = ay;
                // a branched lipid's terms come from k_mol (chemistry bonds stay here)
                if (A.bonded && (R.arch == 0 || Bd.m < 0))
                    bonded_gather<ANG>(R, pi, I.g, Bd, A.xy, A.lnk, A.prv, fxs, fys, mb);
                A.f[slot] = make_double2(fxs, fys);
                if (A.key) {
                    double2 q, vo;
                    float2 vtn;
                    integrate1(R, pi, vh, make_double2(fxs, fys), q, vo, vtn);
                    new_key(R, T.rep, slot, pi, q, A.key, A.rmax, A.moved);
                }
            }
        }
    };
    if (staged) run(true);
    else run(false);
    if (!A.bonded) return;
    // running max bond length of the replica (as k_pair)
    uint64_t bits = (uint64_t)__double_as_longlong(mb);
    bits = wave_max_u64(bits);
    __shared__ uint64_t s_mb[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) s_mb[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++) bits = max(bits, s_mb[w]);
        unsigned long long *mbw = (unsigned long long *)&A.rsw[T.rep].maxbond_bits;
        if (bits > *(volatile unsigned long long *)mbw) atomicMax(mbw, (unsigned long long)bits);
    }
}

}  // namespace oc
