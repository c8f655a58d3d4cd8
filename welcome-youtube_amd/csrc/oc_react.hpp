// oc_react.hpp -- the reference chemistry's catalyst reactions and feed cap
// on gfx950, in the hash-decided form of the reference GPU port
// (cel/engine_gpu/oercel_gpu.cu:1645-1709 chem_react, 1746-1990 chemistry):
//
//   k_integ        the step's integrate in place (cel/engine/oercel.c:2303-2334),
//                  for engines that run the reference chemistry between the
//                  integrate and the sort, with the largest displacement
//   k_react_list   catalysts that drew a synthesis or a replication
//   k_react        one block per replica: the active catalysts sorted by
//                  their hash priority, then executed in rounds -- a catalyst
//                  runs once every lower-priority-rank catalyst whose reach
//                  overlaps its own has run -- so the result is the serial loop
//                  in priority order, one thread per catalyst
//   k_rfeed_count  feed cap: the N count after decay and the candidates' keys
//   k_rfeed_thr    feed cap: the largest key that is kept (radix select)
//
// The per-catalyst body is the reference CPU engine's (cel/engine/oercel.c:
// 1353-1397 find_type_near, 1423-1468 chain_overlap_energy, 1474-1556
// relax_new_chain, 1575-1653 the reactions) on the cell index of the last
// sort: the integrate moved every bead by less than the chemistry margin 0.5
// since (checked; a larger move is an error, where the reference sorts again).
#pragma once

#include "oc_kernels.hpp"

namespace oc {

constexpr int kCatCap = 4096;   // active catalysts per replica and check
constexpr int kReactThr = 256;  // threads of k_react
constexpr int kRelaxNb = 512;   // relaxation neighbours (the reference's MAXNB)
constexpr double kChemMargin = 0.5;

// wrap1, cel/engine/oercel.c:761-771
__device__ __forceinline__ double wrap_box(double v, double L) {
    if (v < 0.0) {
        v += L;
        if (v >= L) v -= L;
    } else if (v >= L) {
        v -= L;
        if (v >= L || v < 0.0) v = fmod(v, L);
    }
    return v;
}

// Integrate of every bead, in slot order and in place, with its new cell key
// (against the cell of the last sort) and the largest per-axis displacement
// of each replica (cel/engine/oercel.c:2313-2317).
__global__ __launch_bounds__(kBeadBlock) void k_integ(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                      const int32_t *__restrict__ bblk, double2 *__restrict__ xy,
                                                      double2 *__restrict__ v, float2 *__restrict__ vt,
                                                      const double2 *__restrict__ f, int32_t *__restrict__ key,
                                                      int2 *__restrict__ rmax, uint32_t *__restrict__ moved) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    uint64_t dm = 0;
    if (loc < R.N) {
        const int s = R.off + loc;
        const double2 p = xy[s], vv = v[s], ff = f[s];
        const double ddx = R.dt * vv.x + R.dt2 * ff.x, ddy = R.dt * vv.y + R.dt2 * ff.y;
        dm = (uint64_t)__double_as_longlong(fmax(fabs(ddx), fabs(ddy)));
        double2 po, vo;
        float2 t;
        integrate1(R, p, vv, ff, po, vo, t);
        xy[s] = po;
        v[s] = vo;
        vt[s] = t;
        new_key(R, r, s, p, po, key, rmax, moved);
    }
    dm = wave_max_u64(dm);
    if ((threadIdx.x & 63) == 0 && dm) atomicMax((unsigned long long *)&rs[r].dmax_bits, (unsigned long long)dm);
}

// Active catalysts of the replicas whose check it is: a Ka / Kb whose draws
// hash_u01(ksyn, gid) < p_syn or hash_u01(krep, gid) < p_rep (of its kind),
// appended to the replica's list (any order: k_react sorts them).
__global__ __launch_bounds__(kBeadBlock) void k_react_list(const RepInfo *__restrict__ ri,
                                                           const RepStep *__restrict__ rs,
                                                           const int32_t *__restrict__ bblk,
                                                           const uint32_t *__restrict__ gt, int32_t *__restrict__ cat,
                                                           int32_t *__restrict__ catn) {
    const int r = bblk[blockIdx.x];
    const RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.react) return;
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const uint32_t w = gt[s], ty = w >> kGidBits, g = w & kGidMask;
    if (ty != OC_KA && ty != OC_KB) return;
    const int var = ty == OC_KB;
    if (!(hash_u01(S.ksyn, g) < R.p_syn[var] || hash_u01(S.krep, g) < R.p_rep[var])) return;
    const int k = atomicAdd(&catn[r], 1);
    if (k < kCatCap) cat[(size_t)r * kCatCap + k] = s;
}

struct ReactBufs {
    double2 *xy, *v;
    float2 *vt;
    uint32_t *gt;
    int8_t *mp;
    int32_t *lnk, *prv;
    const int32_t *cs;  // cell starts of the last sort
    int32_t *key;       // next sort: cell keys (a synthesis re-keys the beads it moves)
    int2 *rmax;
    uint32_t *moved;
};

struct Near {
    int slot, cell;  // cell: where the last sort listed the bead
    uint32_t g;
    double d2;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// The K nearest beads of `type` within r_cat of slot k, by (d2, gid); returns
// how many lie within r_cat, at most 256 (find_type_near).
__device__ int near_of_type(const RepInfo &R, const ReactBufs &B, int k, uint32_t type, Near *out, int K) {
    const double2 pk = B.xy[k];
    const int cx = cell_x(R, pk.x), cy = cell_y(R, pk.y);
    int n = 0, nk = 0;
    for (int ox = -R.rwx; ox <= R.rwx; ox++) {
        const int xx = wrapi(cx + ox, R.ncx);
        for (int oy = -R.rwy; oy <= R.rwy; oy++) {
            const int cc = R.coff + xx * R.ncy + wrapi(cy + oy, R.ncy);
            const int j1 = B.cs[cc + 1];
            for (int j = B.cs[cc]; j < j1; j++) {
                const uint32_t w = B.gt[j];
                if ((w >> kGidBits) != type || j == k) continue;
                const double2 pj = B.xy[j];
                const double dx = minimg(pj.x - pk.x, R.L, R.halfL), dy = minimg(pj.y - pk.y, R.L, R.halfL);
                const double d2 = dx * dx + dy * dy;
                if (!(d2 < R.r2cat)) continue;
                if (n < 256) n++;
                if (K <= 0) continue;
                const uint32_t gj = w & kGidMask;
                if (nk == K && !(d2 < out[K - 1].d2 || (d2 == out[K - 1].d2 && gj < out[K - 1].g))) continue;
                int b = nk < K ? nk++ : K - 1;
                while (b > 0 && (out[b - 1].d2 > d2 || (out[b - 1].d2 == d2 && out[b - 1].g > gj))) {
                    out[b] = out[b - 1];
                    b--;
                }
                out[b].slot = j;
                out[b].cell = cc - R.coff;
                out[b].g = gj;
                out[b].d2 = d2;
            }
        }
    }
    return n;
}

// bead b of a new lipid centred at (cx0, cy0) along (ux, uy), wrapped
__device__ __forceinline__ double2 chain_pos(const RepInfo &R, int b, double cx0, double cy0, double ux, double uy) {
    double px, py;
    if (R.arch == OC_ARCH_LINEAR) {
        const double off = (b - 0.5 * (R.mlen - 1)) * R.r0;
        px = cx0 + off * ux;
        py = cy0 + off * uy;
    } else {
        px = cx0 + (R.tpx[b] * ux - R.tpy[b] * uy);
        py = cy0 + (R.tpx[b] * uy + R.tpy[b] * ux);
    }
    return make_double2(wrap_box(px, R.L), wrap_box(py, R.L));
}

__device__ __forceinline__ bool in_list(int j, const int *sl, int n) {
    bool hit = false;
    for (int q = 0; q < n; q++) hit |= j == sl[q];
    return hit;
}

// conservative overlap energy of a would-be lipid with every other bead
// within 1 (chain_overlap_energy; the raw a_ij)
__device__ double overlap_energy(const RepInfo &R, const ReactBufs &B, double cx0, double cy0, double ux, double uy,
                                 const int *sl) {
    double U = 0.0;
    for (int b = 0; b < R.mlen; b++) {
        const double2 p = chain_pos(R, b, cx0, cy0, ux, uy);
        const int tb = b < R.nH ? OC_H : OC_T;
        const int cx = cell_x(R, p.x), cy = cell_y(R, p.y);
        for (int ox = -R.owx; ox <= R.owx; ox++) {
            const int xx = wrapi(cx + ox, R.ncx);
            for (int oy = -R.owy; oy <= R.owy; oy++) {
                const int cc = R.coff + xx * R.ncy + wrapi(cy + oy, R.ncy);
                const int j1 = B.cs[cc + 1];
                for (int j = B.cs[cc]; j < j1; j++) {
                    if (in_list(j, sl, R.mlen)) continue;
                    const double2 pj = B.xy[j];
                    const double dx = minimg(pj.x - p.x, R.L, R.halfL), dy = minimg(pj.y - p.y, R.L, R.halfL);
                    const double r2 = dx * dx + dy * dy;
                    if (r2 >= 1.0) continue;
                    const double w = 1.0 - sqrt(r2);
                    U += 0.5 * R.a[tb * 8 + (B.gt[j] >> kGidBits)] * w * w;
                }
            }
        }
    }
    return U;
}

// steepest descent of a new lipid in the frozen field of its neighbours
// (relax_new_chain): neighbours within 1 + 0.05 syn_relax of any of its beads
// collected once (per-bead windows, first visit order, at most 512), then
// syn_relax moves of 0.002 F capped at 0.05 per bead
__device__ void relax_chain(const RepInfo &R, const ReactBufs &B, const int *sl) {
    const int ml = R.mlen;
    const double cap = 0.05, alpha = 0.002;
    int nbj[kRelaxNb];
    int nnb = 0;
    const double reach = 1.0 + cap * R.syn_relax, reach2 = reach * reach;
    for (int b = 0; b < ml && nnb < kRelaxNb; b++) {
        const double2 p = B.xy[sl[b]];
        const int cx = cell_x(R, p.x), cy = cell_y(R, p.y);
        for (int ox = -R.lwx; ox <= R.lwx; ox++) {
            const int xx = wrapi(cx + ox, R.ncx);
            for (int oy = -R.lwy; oy <= R.lwy; oy++) {
                const int cc = R.coff + xx * R.ncy + wrapi(cy + oy, R.ncy);
                const int j1 = B.cs[cc + 1];
                for (int j = B.cs[cc]; j < j1; j++) {
                    if (in_list(j, sl, ml) || in_list(j, nbj, nnb)) continue;
                    const double2 pj = B.xy[j];
                    const double dx = minimg(pj.x - p.x, R.L, R.halfL), dy = minimg(pj.y - p.y, R.L, R.halfL);
                    if (dx * dx + dy * dy < reach2 && nnb < kRelaxNb) nbj[nnb++] = j;
                }
            }
        }
    }
    for (int it = 0; it < R.syn_relax; it++) {
        double fx[OC_MAXMLEN], fy[OC_MAXMLEN];
        double2 q[OC_MAXMLEN];
        for (int b = 0; b < ml; b++) {
            fx[b] = 0.0;
            fy[b] = 0.0;
            q[b] = B.xy[sl[b]];
        }
        for (int b = 0; b < ml; b++) {
            const uint32_t ti = B.gt[sl[b]] >> kGidBits;
            for (int u = 0; u < nnb; u++) {
                const int j = nbj[u];
                const double2 pj = B.xy[j];
                const double dx = minimg(q[b].x - pj.x, R.L, R.halfL), dy = minimg(q[b].y - pj.y, R.L, R.halfL);
                const double r2 = dx * dx + dy * dy;
                if (r2 >= 1.0 || r2 < 1e-24) continue;
                const double rr = sqrt(r2);
                const double f = R.a[ti * 8 + (B.gt[j] >> kGidBits)] * (1.0 - rr) / rr;
                fx[b] += f * dx;
                fy[b] += f * dy;
            }
        }
        if (R.arch == OC_ARCH_LINEAR) {
            double bx[OC_MAXMLEN], by[OC_MAXMLEN], bl[OC_MAXMLEN];
            for (int b = 0; b < ml - 1; b++) {
                bx[b] = minimg(q[b + 1].x - q[b].x, R.L, R.halfL);
                by[b] = minimg(q[b + 1].y - q[b].y, R.L, R.halfL);
                bl[b] = sqrt(bx[b] * bx[b] + by[b] * by[b]);
                const double f = -R.kb * (bl[b] - R.r0) / bl[b];
                fx[b + 1] += f * bx[b];
                fy[b + 1] += f * by[b];
                fx[b] -= f * bx[b];
                fy[b] -= f * by[b];
            }
            const double ka = R.ka;
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
            // template bonds, then angle terms, by template position (mol_forces)
            for (int e = 0; e < R.tnb; e++) {
                const int bi = R.bi[e], bj = R.bj[e];
                const double dx = minimg(q[bj].x - q[bi].x, R.L, R.halfL), dy = minimg(q[bj].y - q[bi].y, R.L, R.halfL);
                const double rr = sqrt(dx * dx + dy * dy);
                const double f = -R.kb * (rr - R.br0[e]) / rr;
                fx[bj] += f * dx;
                fy[bj] += f * dy;
                fx[bi] -= f * dx;
                fy[bi] -= f * dy;
            }
            for (int e = 0; e < R.tna; e++) {
                const int pi_ = R.ai[e], pj = R.aj[e], pk = R.ak[e];
                const double ax = minimg(q[pj].x - q[pi_].x, R.L, R.halfL), ay = minimg(q[pj].y - q[pi_].y, R.L, R.halfL);
                const double cxv = minimg(q[pk].x - q[pj].x, R.L, R.halfL), cyv = minimg(q[pk].y - q[pj].y, R.L, R.halfL);
                const double la = sqrt(ax * ax + ay * ay), lc = sqrt(cxv * cxv + cyv * cyv);
                const double ilac = 1.0 / (la * lc);
                const double cs = (ax * cxv + ay * cyv) * ilac;
                const double gg = R.akind[e] == 0 ? -R.aK[e] : 2.0 * R.aK[e] * (cs - R.ac0[e]);
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
            B.xy[sl[b]] = make_double2(wrap_box(q[b].x + mx, R.L), wrap_box(q[b].y + my, R.L));
        }
    }
}

// One active catalyst's reactions: synthesis, then replication.
__device__ void react_one(const RepInfo &R, const RepStep &S, const ReactBufs &B, int k, int r, Event *ev,
                          unsigned long long *evn, int64_t evcap, int record, uint32_t *err) {
    const uint32_t wk = B.gt[k], tk = wk >> kGidBits, g = wk & kGidMask;
    if (tk != OC_KA && tk != OC_KB) return;
    const int var = tk == OC_KB;
    const double us = hash_u01(S.ksyn, g), ur = hash_u01(S.krep, g);
    const int ml = R.mlen, nsynN = R.nsynN;
    Near nb[OC_MAXMLEN + 2], nbw[OC_MAXMLEN];
    const int nn = near_of_type(R, B, k, OC_N, nb, nsynN + 2);
    if (nn < 2) return;
    int used = 0;
    const double2 pk = B.xy[k];
    int nw = 0;
    if (nn >= nsynN && nsynN < ml) nw = near_of_type(R, B, k, OC_W, nbw, ml - nsynN);
    if (nn >= nsynN && nsynN + nw >= ml && us < R.p_syn[var]) {
        double cxm = 0, cym = 0, mvx = 0, mvy = 0, mtx = 0, mty = 0;
        int sl[OC_MAXMLEN];
        for (int b = 0; b < ml; b++) {
            const int j = b < nsynN ? nb[b].slot : nbw[b - nsynN].slot;
            sl[b] = j;
            const double2 pj = B.xy[j], vj = B.v[j];
            const float2 tj = B.vt[j];
            cxm += minimg(pj.x - pk.x, R.L, R.halfL);
            cym += minimg(pj.y - pk.y, R.L, R.halfL);
            mvx += vj.x;
            mvy += vj.y;
            mtx += (double)tj.x;
            mty += (double)tj.y;
        }
        cxm = pk.x + cxm / ml;
        cym = pk.y + cym / ml;
        mvx /= ml;
        mvy /= ml;
        mtx /= ml;
        mty /= ml;
        double ux = 1.0, uy = 0.0, ubest = INFINITY;
        for (int tr = 0; tr < R.syn_tries; tr++) {
            const double phi = 2.0 * M_PI * hash_u01(hash_key(R.seed, S.step, 40 + (uint64_t)tr), g);
            const double cx_ = cos(phi), cy_ = sin(phi);
            const double U = R.syn_tries > 1 ? overlap_energy(R, B, cxm, cym, cx_, cy_, sl) : 0.0;
            if (U < ubest) {
                ubest = U;
                ux = cx_;
                uy = cy_;
            }
        }
        for (int b = 0; b < ml; b++) {
            const int j = sl[b];
            B.xy[j] = chain_pos(R, b, cxm, cym, ux, uy);
            B.v[j] = make_double2(mvx, mvy);
            B.vt[j] = make_float2((float)mtx, (float)mty);
            B.gt[j] = (B.gt[j] & kGidMask) | ((uint32_t)(b < R.nH ? OC_H : OC_T) << kGidBits);
            B.mp[j] = (int8_t)b;
            B.lnk[j] = b + 1 < ml ? sl[b + 1] : -1;
            B.prv[j] = b > 0 ? sl[b - 1] : -1;
        }
        if (R.syn_relax > 0) relax_chain(R, B, sl);
        for (int b = 0; b < ml; b++) {
            const int oc = b < nsynN ? nb[b].cell : nbw[b - nsynN].cell;
            new_key_from(R, r, sl[b], oc / R.ncy, oc % R.ncy, B.xy[sl[b]], B.key, B.rmax, B.moved);
        }
        if (record) push_event(ev, evn, evcap, err, S.step, OC_EV_SYN, g, (uint32_t)var, r, pk.x, pk.y);
        used = nsynN;
    }
    if (nn - used >= 2 && ur < R.p_rep[var]) {
        int nv = var, mut = 0;
        if (hash_u01(S.kmu, g) < R.mu) {
            nv = 1 - var;
            mut = 1;
        }
        const int j1 = nb[used].slot, j2 = nb[used + 1].slot;
        B.gt[j1] = (B.gt[j1] & kGidMask) | ((uint32_t)(nv ? OC_KB : OC_KA) << kGidBits);
        B.gt[j2] = (B.gt[j2] & kGidMask) | ((uint32_t)OC_W << kGidBits);
        if (record)
            push_event(ev, evn, evcap, err, S.step, mut ? OC_EV_MUT : OC_EV_REP, B.gt[j1] & kGidMask, (uint32_t)nv, r,
                       pk.x, pk.y);
    }
}

// two catalysts whose cells are closer than (cfx, cfy) may read or write each
// other's beads (periodic cell distance)
__device__ __forceinline__ bool reach_overlap(const RepInfo &R, int ax, int ay, int bx, int by) {
    int dx = abs(ax - bx), dy = abs(ay - by);
    dx = min(dx, R.ncx - dx);
    dy = min(dy, R.ncy - dy);
    return dx <= R.cfx && dy <= R.cfy;
}

__device__ __forceinline__ bool prio_less(uint64_t ka, uint32_t ga, uint64_t kb, uint32_t gb) {
    return ka != kb ? ka < kb : ga < gb;
}

// The reactions of one replica's check (one block of kReactThr threads).
__global__ __launch_bounds__(kReactThr) void k_react(const RepInfo *__restrict__ ri, const RepStep *__restrict__ rs,
                                                     ReactBufs B, const int32_t *__restrict__ cat,
                                                     int32_t *__restrict__ catn, Event *__restrict__ ev,
                                                     unsigned long long *__restrict__ evn, int64_t evcap, int record,
                                                     uint32_t *__restrict__ err) {
    const int r = blockIdx.x, t = threadIdx.x;
    const RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.react) return;
    __shared__ uint64_t s_key[kCatCap];  // priority keys; after the sort: cells (cx, cy)
    __shared__ int32_t s_slot[kCatCap];
    __shared__ uint8_t s_done[kCatCap], s_run[kCatCap];
    __shared__ int s_left;
    const int n = catn[r];
    const bool bad_n = n > kCatCap, bad_d = __longlong_as_double((long long)S.dmax_bits) >= kChemMargin;
    __syncthreads();  // every thread has read catn
    if (bad_n || bad_d) {
        if (t == 0) {
            atomicOr(err, bad_n ? ERR_CATCAP : ERR_DISP);
            catn[r] = 0;
        }
        return;
    }
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = t; i < P; i += kReactThr) {
        if (i < n) {
            const int s = cat[(size_t)r * kCatCap + i];
            s_slot[i] = s;
            s_key[i] = mix13(S.kpri ^ ((uint64_t)(B.gt[s] & kGidMask) * 0xd1b54a32d192ed03ULL));
        } else {
            s_slot[i] = -1;
            s_key[i] = ~0ull;
        }
    }
    __syncthreads();
    // bitonic sort by (key, gid); padding (slot -1) counts as the largest gid
    for (int kk = 2; kk <= P; kk <<= 1)
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
            for (int i = t; i < P; i += kReactThr) {
                const int l = i ^ jj;
                if (l <= i) continue;
                const int si = s_slot[i], sl_ = s_slot[l];
                const uint32_t gi = si < 0 ? 0xffffffffu : (B.gt[si] & kGidMask);
                const uint32_t gl = sl_ < 0 ? 0xffffffffu : (B.gt[sl_] & kGidMask);
                const bool up = (i & kk) == 0;
                if (prio_less(s_key[l], gl, s_key[i], gi) == up) {
                    const uint64_t tk = s_key[i];
                    s_key[i] = s_key[l];
                    s_key[l] = tk;
                    s_slot[i] = sl_;
                    s_slot[l] = si;
                }
            }
            __syncthreads();
        }
    int32_t *s_cx = (int32_t *)s_key, *s_cy = s_cx + kCatCap;
    for (int i = t; i < n; i += kReactThr) {
        const double2 p = B.xy[s_slot[i]];
        s_cx[i] = cell_x(R, p.x);
        s_cy[i] = cell_y(R, p.y);
        s_done[i] = 0;
    }
    __syncthreads();
    // rounds: a catalyst runs when no lower-ranked catalyst within reach is
    // still waiting; the lowest waiting one always runs, so every round
    // makes progress
    for (;;) {
        for (int i = t; i < n; i += kReactThr) {
            s_run[i] = 0;
            if (s_done[i]) continue;
            bool ready = true;
            for (int j = 0; j < i && ready; j++)
                if (!s_done[j] && reach_overlap(R, s_cx[i], s_cy[i], s_cx[j], s_cy[j])) ready = false;
            s_run[i] = ready ? 1 : 0;
        }
        __syncthreads();
        for (int i = t; i < n; i += kReactThr)
            if (s_run[i]) react_one(R, S, B, s_slot[i], r, ev, evn, evcap, record, err);
        if (t == 0) s_left = 0;
        __syncthreads();
        for (int i = t; i < n; i += kReactThr) {
            if (s_run[i]) s_done[i] = 1;
            if (!s_done[i]) atomicOr(&s_left, 1);
        }
        __syncthreads();
        if (!s_left) break;
        __syncthreads();
    }
    if (t == 0) catn[r] = 0;
}

// Feed cap, first pass: the number of N after the decay pass and the keys of
// the feed candidates (the replica's key list, any order).
__global__ __launch_bounds__(kBeadBlock) void k_rfeed_count(const RepInfo *__restrict__ ri,
                                                            const RepStep *__restrict__ rs,
                                                            const int32_t *__restrict__ bblk,
                                                            const double2 *__restrict__ xy,
                                                            const uint32_t *__restrict__ gt,
                                                            uint64_t *__restrict__ fkey, int2 *__restrict__ fcnt) {
    const int r = bblk[blockIdx.x];
    const RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.fcap_on) return;
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    bool isN = false, cand = false;
    uint32_t g = 0;
    if (loc < R.N) {
        const int s = R.off + loc;
        const uint32_t w = gt[s];
        g = w & kGidMask;
        isN = (w >> kGidBits) == OC_N;
        cand = (w >> kGidBits) == OC_W && feed_candidate(R, S, w, xy[s]);
    }
    const int nN = __popcll(__ballot(isN));
    if ((threadIdx.x & 63) == 0 && nN) atomicAdd(&fcnt[r].x, nN);
    if (cand) {
        const int k = atomicAdd(&fcnt[r].y, 1);
        fkey[(size_t)R.off + k] = feed_key(S, g);
    }
}

// Feed cap, second pass (one block per replica): with room = floor(feed_cap N
// + 1e-9) - (number of N), all candidates are kept when they fit, else the
// room smallest keys: the largest kept key by an 8-bit radix select over the
// (unique) keys.  Clears the counters for the next check.
__global__ __launch_bounds__(256) void k_rfeed_thr(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                   const uint64_t *__restrict__ fkey, int2 *__restrict__ fcnt) {
    const int r = blockIdx.x, t = threadIdx.x;
    RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.fcap_on) return;
    __shared__ unsigned int hist[256];
    __shared__ uint64_t s_pre;
    __shared__ int s_rank;
    const int2 c = fcnt[r];
    const int nc = c.y, room = max(0, R.fcap_n - c.x);
    __syncthreads();  // every thread has read the counters
    if (nc <= room || room == 0) {
        if (t == 0) {
            S.feed_add = nc <= room ? nc : 0;
            S.feed_thr = ~0ull;
            fcnt[r] = make_int2(0, 0);
        }
        return;
    }
    const uint64_t *K = fkey + R.off;
    uint64_t pre = 0, mask = 0;
    int rank = room - 1;
    for (int sh = 56; sh >= 0; sh -= 8) {
        hist[t] = 0;
        __syncthreads();
        for (int q = t; q < nc; q += 256) {
            const uint64_t k = K[q];
            if ((k & mask) == pre) atomicAdd(&hist[(k >> sh) & 255], 1u);
        }
        __syncthreads();
        if (t == 0) {
            int cum = 0;
            for (int b = 0; b < 256; b++) {
                if (cum + (int)hist[b] > rank) {
                    s_pre = pre | ((uint64_t)b << sh);
                    s_rank = rank - cum;
                    break;
                }
                cum += (int)hist[b];
            }
        }
        __syncthreads();
        pre = s_pre;
        rank = s_rank;
        mask |= 0xffull << sh;
        __syncthreads();
    }
    if (t == 0) {
        S.feed_add = room;
        S.feed_thr = pre;
        fcnt[r] = make_int2(0, 0);
    }
}

}  // namespace oc
