// oc_react.hpp -- the reference chemistry's catalyst reactions and feed cap
// on gfx950, in the hash-decided form of the reference GPU port
// (cel/engine_gpu/oercel_gpu.cu:1645-1709 chem_react, 1746-1990 chemistry):
//
//   k_integ        the step's integrate in place (cel/engine/oercel.c:2303-2334),
//                  for engines that run the reference chemistry between the
//                  integrate and the sort; at a check it also lists the
//                  catalysts that drew a synthesis or a replication
//   k_react_list   the same list for a check on a loaded state
//   k_react_prep   one block per replica: the active catalysts sorted by their
//                  hash priority (LDS bitonic sort), with their cells
//   k_react_free   grid-wide, one wavefront per catalyst: every catalyst that
//                  no lower-ranked catalyst can reach reacts at once
//   k_react_rest   one block per replica: the others, in rounds -- a catalyst
//                  runs once every lower-ranked catalyst within its reach has
//                  run -- so the result is the serial loop in priority order
//   k_rfeed_count  feed cap: the N count after decay and the candidates' keys
//   k_rfeed_thr    feed cap: the largest key that is kept (radix select)
//
// A reaction runs on one wavefront: it gathers the beads within its reach
// once (listed cells and distance), then each query of the CPU engine's body
// (K nearest, overlap energy per orientation, relaxation neighbours) filters
// that list with the query's own window of listed cells and its distance
// test -- the same sets as the CPU engine's cell scans -- with lanes over
// beads and wavefront sums.
//
// The per-catalyst body is the reference CPU engine's (cel/engine/oercel.c:
// 1353-1397 find_type_near, 1423-1468 chain_overlap_energy, 1474-1556
// relax_new_chain, 1575-1653 the reactions) on the cell index of the last
// sort: the integrate moved every bead by less than the chemistry margin 0.5
// since (checked; a larger move is an error, where the reference sorts again).
#pragma once

#include "oc_kernels.hpp"

namespace oc {

constexpr int kCatCap = 4096;     // active catalysts per replica and check
constexpr int kPrepThr = 1024;    // threads of k_react_prep
constexpr int kRestThr = 128;     // threads of k_react_rest (2 wavefronts)
constexpr int kFreeBlocks = 256;  // blocks of k_react_free (4 wavefronts each)
constexpr int kRelaxNb = 512;     // relaxation neighbours (the reference's MAXNB)
constexpr int kCand = 128;        // beads of one kind within r_cat of a catalyst
constexpr int kOv = 192;          // beads within reach of a new lipid's overlap energy (LDS copy)
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
// (against the cell of the last sort); flags a replica in which a bead moved
// by the chemistry margin or more (cel/engine/oercel.c:2313-2317, 2337-2343).
// At a check, also lists the active catalysts: a Ka / Kb whose draws
// hash_u01(ksyn, gid) < p_syn or hash_u01(krep, gid) < p_rep (of its kind),
// appended to the replica's list (any order: k_react_prep sorts them).
__global__ __launch_bounds__(kBeadBlock) void k_integ(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                      const int32_t *__restrict__ bblk, double2 *__restrict__ xy,
                                                      double2 *__restrict__ v, float2 *__restrict__ vt,
                                                      const double2 *__restrict__ f, int32_t *__restrict__ key,
                                                      int2 *__restrict__ rmax, uint32_t *__restrict__ moved,
                                                      const uint32_t *__restrict__ gt, int32_t *__restrict__ cat,
                                                      int32_t *__restrict__ catn) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const double2 p = xy[s], vv = v[s], ff = f[s];
    const double ddx = R.dt * vv.x + R.dt2 * ff.x, ddy = R.dt * vv.y + R.dt2 * ff.y;
    if (!(fmax(fabs(ddx), fabs(ddy)) < kChemMargin)) atomicOr((unsigned long long *)&rs[r].disp, 1ull);
    double2 po, vo;
    float2 t;
    integrate1(R, p, vv, ff, po, vo, t);
    xy[s] = po;
    v[s] = vo;
    vt[s] = t;
    new_key(R, r, s, p, po, key, rmax, moved);
    const RepStep &S = rs[r];
    if (!cat || !S.rchem_now || !R.react) return;
    const uint32_t w = gt[s], ty = w >> kGidBits, g = w & kGidMask;
    if (ty != OC_KA && ty != OC_KB) return;
    const int var = ty == OC_KB;
    if (!(hash_u01(S.ksyn, g) < R.p_syn[var] || hash_u01(S.krep, g) < R.p_rep[var])) return;
    const int k = atomicAdd(&catn[r], 1);
    if (k < kCatCap) cat[(size_t)r * kCatCap + k] = s;
}

// The active catalysts of a check on a loaded state (oc_chem), as k_integ
// lists them in a step.
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

// a bead within a catalyst's reach: position, slot, the cell the last sort
// listed it in, its gt word; skip marks the beads of the lipid being made
struct GBead {
    double x, y;
    int32_t slot, cell;
    uint32_t w, skip;
};

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
    int32_t *cat, *catn;  // per replica: active catalysts (priority order after k_react_prep), count
    int2 *ccell;          // their cells
    uint8_t *cdone;       // reacted
    GBead *gs;            // per-wavefront gather lists, gcap beads each
    int gcap;
    Event *ev;
    unsigned long long *evn;
    int64_t evcap;
    int record;
    uint32_t *err;
};

// per-wavefront LDS of a reaction
struct WaveLds {
    double cd2[kCand];                    // K-nearest candidates: d2, gid, gather index
    uint32_t cg[kCand];
    int32_t cq[kCand];
    int32_t nb[OC_MAXMLEN + 2], nbw[OC_MAXMLEN];  // the K nearest N / W (gather indices)
    int32_t nbi[kRelaxNb];                // relaxation neighbours (gather indices)
    double ox[kOv], oy[kOv];              // overlap-energy partners: position, gt word, listed cell
    uint32_t ow[kOv];
    int32_t ocell[kOv];
    double px[OC_MAXMLEN], py[OC_MAXMLEN], fx[OC_MAXMLEN], fy[OC_MAXMLEN];  // the new lipid
    int32_t pcx[OC_MAXMLEN], pcy[OC_MAXMLEN], sl[OC_MAXMLEN], scell[OC_MAXMLEN];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// butterfly sums: every lane ends with the same bits
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << (threadIdx.x & 63)) - 1ull; }

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// is (local) cell `cell` within (wx, wy) cells of (cx, cy)?  (periodic)
__device__ __forceinline__ bool in_win(const RepInfo &R, int cell, int cx, int cy, int wx, int wy) {
    const int jx = cell / R.ncy, jy = cell - jx * R.ncy;
    int dx = abs(jx - cx), dy = abs(jy - cy);
    dx = min(dx, R.ncx - dx);
    dy = min(dy, R.ncy - dy);
    return dx <= wx && dy <= wy;
}

// Every bead listed within (rdx, rdy) cells of the catalyst's cell (the whole
// axis when that wraps onto itself) and within rg of it, in a fixed order;
// returns the count, -1 beyond the list's capacity.
__device__ int gather(const RepInfo &R, const ReactBufs &B, double2 pk, int kcx, int kcy, GBead *G) {
    const int lane = threadIdx.x & 63;
    const bool allx = 2 * R.rdx + 1 >= R.ncx, ally = 2 * R.rdy + 1 >= R.ncy;
    const int nxg = allx ? R.ncx : 2 * R.rdx + 1, nyg = ally ? R.ncy : 2 * R.rdy + 1, nc = nxg * nyg;
    int n = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        const int c = c0 + lane;
        int a = 0, e = 0, cell = 0;
        if (c < nc) {
            const int ix = c / nyg, iy = c - ix * nyg;
            const int xx = allx ? ix : wrapi(kcx + ix - R.rdx, R.ncx), yy = ally ? iy : wrapi(kcy + iy - R.rdy, R.ncy);
            cell = xx * R.ncy + yy;
            a = B.cs[R.coff + cell];
            e = B.cs[R.coff + cell + 1];
        }
        const int len = e - a, mx = wave_max_i(len);
        for (int t = 0; t < mx; t++) {
            bool keep = false;
            GBead gb;
            if (t < len) {
                const int j = a + t;
                const double2 p = B.xy[j];
                const double dx = minimg(p.x - pk.x, R.L, R.halfL), dy = minimg(p.y - pk.y, R.L, R.halfL);
                keep = dx * dx + dy * dy < R.rg2;
                gb.x = p.x;
                gb.y = p.y;
                gb.slot = j;
                gb.cell = cell;
                gb.w = B.gt[j];
                gb.skip = 0;
            }
            const uint64_t m = __ballot(keep);
            const int at = n + __popcll(m & lanes_below());
            if (keep && at < B.gcap) G[at] = gb;
            n += __popcll(m);
        }
    }
    return n <= B.gcap ? n : -1;
}

// The K nearest gathered beads of `type` within r_cat of slot k (listed in
// the r_cat window, by (d2, gid)) into out[] (gather indices); returns how
// many lie within r_cat (at most 256; find_type_near), -1 beyond kCand.
__device__ int near_wave(const RepInfo &R, const GBead *G, int n, int k, double2 pk, int kcx, int kcy,
                         uint32_t type, WaveLds &L, int32_t *out, int K) {
    const int lane = threadIdx.x & 63;
    int m = 0;
    for (int q0 = 0; q0 < n; q0 += 64) {
        const int q = q0 + lane;
        bool keep = false;
        double d2 = 0.0;
        uint32_t g = 0;
        if (q < n) {
            const GBead &b = G[q];
            if ((b.w >> kGidBits) == type && b.slot != k && in_win(R, b.cell, kcx, kcy, R.rwx, R.rwy)) {
                const double dx = minimg(b.x - pk.x, R.L, R.halfL), dy = minimg(b.y - pk.y, R.L, R.halfL);
                d2 = dx * dx + dy * dy;
                keep = d2 < R.r2cat;
                g = b.w & kGidMask;
            }
        }
        const uint64_t mk = __ballot(keep);
        const int at = m + __popcll(mk & lanes_below());
        if (keep && at < kCand) {
            L.cd2[at] = d2;
            L.cg[at] = g;
            L.cq[at] = q;
        }
        m += __popcll(mk);
    }
    if (m > kCand) return -1;
    wave_sync();
    for (int i = lane; i < m; i += 64) {
        int rank = 0;
        for (int c = 0; c < m; c++) rank += L.cd2[c] < L.cd2[i] || (L.cd2[c] == L.cd2[i] && L.cg[c] < L.cg[i]);
        if (rank < K) out[rank] = L.cq[i];
    }
    wave_sync();
    return m < 256 ? m : 256;
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

// conservative overlap energy of a would-be lipid with every other bead
// within 1 listed in the overlap window of each of its beads
// (chain_overlap_energy; the raw a_ij); partners from the LDS copy when it
// holds them (no > 0), else from the gather list
__device__ double overlap_wave(const RepInfo &R, const GBead *G, int n, const WaveLds &L, int no, double cx0,
                               double cy0, double ux, double uy) {
    const int lane = threadIdx.x & 63;
    double u = 0.0;
    for (int b = 0; b < R.mlen; b++) {
        const double2 p = chain_pos(R, b, cx0, cy0, ux, uy);
        const int tb = b < R.nH ? OC_H : OC_T;
        const int pcx = cell_x(R, p.x), pcy = cell_y(R, p.y);
        if (no > 0) {
            for (int q = lane; q < no; q += 64) {
                if (!in_win(R, L.ocell[q], pcx, pcy, R.owx, R.owy)) continue;
                const double dx = minimg(L.ox[q] - p.x, R.L, R.halfL), dy = minimg(L.oy[q] - p.y, R.L, R.halfL);
                const double r2 = dx * dx + dy * dy;
                if (r2 >= 1.0) continue;
                const double w = 1.0 - sqrt(r2);
                u += 0.5 * R.a[tb * 8 + (L.ow[q] >> kGidBits)] * w * w;
            }
        } else {
            for (int q = lane; q < n; q += 64) {
                const GBead &g = G[q];
                if (g.skip || !in_win(R, g.cell, pcx, pcy, R.owx, R.owy)) continue;
                const double dx = minimg(g.x - p.x, R.L, R.halfL), dy = minimg(g.y - p.y, R.L, R.halfL);
                const double r2 = dx * dx + dy * dy;
                if (r2 >= 1.0) continue;
                const double w = 1.0 - sqrt(r2);
                u += 0.5 * R.a[tb * 8 + (g.w >> kGidBits)] * w * w;
            }
        }
    }
    return wave_sum_d(u);
}

constexpr int kNbLane = kRelaxNb / 64;  // relaxation neighbours held per lane

// steepest descent of the new lipid (L.px / L.py, cells L.pcx / L.pcy) in
// the frozen field of its neighbours (relax_new_chain): neighbours listed in
// the relaxation window of one of its beads and within 1 + 0.05 syn_relax of
// it, at most 512; syn_relax moves of 0.002 F capped at 0.05 per bead
__device__ void relax_wave(const RepInfo &R, const GBead *G, int n, WaveLds &L) {
    const int lane = threadIdx.x & 63, ml = R.mlen;
    const double cap = 0.05, alpha = 0.002, reach = 1.0 + cap * R.syn_relax, reach2 = reach * reach;
    int nnb = 0;
    for (int q0 = 0; q0 < n && nnb < kRelaxNb; q0 += 64) {
        const int q = q0 + lane;
        bool keep = false;
        if (q < n && !G[q].skip) {
            const GBead &g = G[q];
            for (int b = 0; b < ml && !keep; b++) {
                if (!in_win(R, g.cell, L.pcx[b], L.pcy[b], R.lwx, R.lwy)) continue;
                const double dx = minimg(g.x - L.px[b], R.L, R.halfL), dy = minimg(g.y - L.py[b], R.L, R.halfL);
                keep = dx * dx + dy * dy < reach2;
            }
        }
        const uint64_t mk = __ballot(keep);
        const int at = nnb + __popcll(mk & lanes_below());
        if (keep && at < kRelaxNb) L.nbi[at] = q;
        nnb = min(nnb + __popcll(mk), kRelaxNb);
    }
    wave_sync();
    double nx[kNbLane], ny[kNbLane];
    uint32_t nt[kNbLane];
#pragma unroll
    for (int t = 0; t < kNbLane; t++) {
        const int m = t * 64 + lane;
        nt[t] = 0xffffffffu;
        nx[t] = ny[t] = 0.0;
        if (m < nnb) {
            const GBead &g = G[L.nbi[m]];
            nx[t] = g.x;
            ny[t] = g.y;
            nt[t] = g.w >> kGidBits;
        }
    }
    const bool lin = R.arch == OC_ARCH_LINEAR;
    for (int it = 0; it < R.syn_relax; it++) {
        double myfx = 0.0, myfy = 0.0;  // lane b: the pair force on bead b
        for (int b = 0; b < ml; b++) {
            const double bx = L.px[b], by = L.py[b];
            const int tb = b < R.nH ? OC_H : OC_T;
            double fxl = 0.0, fyl = 0.0;
#pragma unroll
            for (int t = 0; t < kNbLane; t++) {
                if (nt[t] == 0xffffffffu) continue;
                const double dx = minimg(bx - nx[t], R.L, R.halfL), dy = minimg(by - ny[t], R.L, R.halfL);
                const double r2 = dx * dx + dy * dy;
                if (r2 >= 1.0 || r2 < 1e-24) continue;
                const double rr = sqrt(r2);
                const double f = R.a[tb * 8 + nt[t]] * (1.0 - rr) / rr;
                fxl += f * dx;
                fyl += f * dy;
            }
            const double sx = wave_sum_d(fxl), sy = wave_sum_d(fyl);
            if (lane == b) {
                myfx = sx;
                myfy = sy;
            }
        }
        if (lin) {
            // lane b: bond b (beads b, b + 1) and angle b (beads b, b + 1, b + 2);
            // each bead adds its terms in the CPU engine's order (bonds, then
            // angles, by index)
            double bx = 0.0, by = 0.0, bl = 1.0, fb = 0.0;
            if (lane < ml - 1) {
                bx = minimg(L.px[lane + 1] - L.px[lane], R.L, R.halfL);
                by = minimg(L.py[lane + 1] - L.py[lane], R.L, R.halfL);
                bl = sqrt(bx * bx + by * by);
                fb = -R.kb * (bl - R.r0) / bl;
            }
            const double ka = R.ka;
            const double cxv = __shfl(bx, lane + 1, 64), cyv = __shfl(by, lane + 1, 64), lc = __shfl(bl, lane + 1, 64);
            double dUax = 0.0, dUay = 0.0, dUcx = 0.0, dUcy = 0.0;
            if (lane < ml - 2 && ka != 0.0) {
                const double ax = bx, ay = by, la = bl;
                const double ilac = 1.0 / (la * lc), cs = (ax * cxv + ay * cyv) * ilac;
                dUax = -ka * (cxv * ilac - cs * ax / (la * la));
                dUay = -ka * (cyv * ilac - cs * ay / (la * la));
                dUcx = -ka * (ax * ilac - cs * cxv / (lc * lc));
                dUcy = -ka * (ay * ilac - cs * cyv / (lc * lc));
            }
            const double pbx = __shfl(bx, lane - 1, 64), pby = __shfl(by, lane - 1, 64), pfb = __shfl(fb, lane - 1, 64);
            const double a1x = __shfl(dUax, lane - 1, 64), a1y = __shfl(dUay, lane - 1, 64);
            const double c1x = __shfl(dUcx, lane - 1, 64), c1y = __shfl(dUcy, lane - 1, 64);
            const double c2x = __shfl(dUcx, lane - 2, 64), c2y = __shfl(dUcy, lane - 2, 64);
            if (lane < ml) {
                if (lane >= 1) {
                    myfx += pfb * pbx;
                    myfy += pfb * pby;
                }
                if (lane <= ml - 2) {
                    myfx -= fb * bx;
                    myfy -= fb * by;
                }
                if (ka != 0.0) {
                    if (lane >= 2) {
                        myfx -= c2x;
                        myfy -= c2y;
                    }
                    if (lane >= 1 && lane <= ml - 2) {
                        myfx += -a1x + c1x;
                        myfy += -a1y + c1y;
                    }
                    if (lane <= ml - 3) {
                        myfx += dUax;
                        myfy += dUay;
                    }
                }
            }
        } else {
            if (lane < ml) {
                L.fx[lane] = myfx;
                L.fy[lane] = myfy;
            }
            wave_sync();
            if (lane == 0) {
                // template bonds, then angle terms, by template position (mol_forces)
                for (int e = 0; e < R.tnb; e++) {
                    const int bi = R.bi[e], bj = R.bj[e];
                    const double dx = minimg(L.px[bj] - L.px[bi], R.L, R.halfL),
                                 dy = minimg(L.py[bj] - L.py[bi], R.L, R.halfL);
                    const double rr = sqrt(dx * dx + dy * dy);
                    const double f = -R.kb * (rr - R.br0[e]) / rr;
                    L.fx[bj] += f * dx;
                    L.fy[bj] += f * dy;
                    L.fx[bi] -= f * dx;
                    L.fy[bi] -= f * dy;
                }
                for (int e = 0; e < R.tna; e++) {
                    const int pi_ = R.ai[e], pj = R.aj[e], pk = R.ak[e];
                    const double ax = minimg(L.px[pj] - L.px[pi_], R.L, R.halfL),
                                 ay = minimg(L.py[pj] - L.py[pi_], R.L, R.halfL);
                    const double cxv = minimg(L.px[pk] - L.px[pj], R.L, R.halfL),
                                 cyv = minimg(L.py[pk] - L.py[pj], R.L, R.halfL);
                    const double la = sqrt(ax * ax + ay * ay), lc = sqrt(cxv * cxv + cyv * cyv);
                    const double ilac = 1.0 / (la * lc);
                    const double cs = (ax * cxv + ay * cyv) * ilac;
                    const double gg = R.akind[e] == 0 ? -R.aK[e] : 2.0 * R.aK[e] * (cs - R.ac0[e]);
                    const double dUax = gg * (cxv * ilac - cs * ax / (la * la));
                    const double dUay = gg * (cyv * ilac - cs * ay / (la * la));
                    const double dUcx = gg * (ax * ilac - cs * cxv / (lc * lc));
                    const double dUcy = gg * (ay * ilac - cs * cyv / (lc * lc));
                    L.fx[pi_] += dUax;
                    L.fy[pi_] += dUay;
                    L.fx[pj] += -dUax + dUcx;
                    L.fy[pj] += -dUay + dUcy;
                    L.fx[pk] -= dUcx;
                    L.fy[pk] -= dUcy;
                }
            }
            wave_sync();
            if (lane < ml) {
                myfx = L.fx[lane];
                myfy = L.fy[lane];
            }
        }
        if (lane < ml) {
            double mx = alpha * myfx, my = alpha * myfy;
            const double m = sqrt(mx * mx + my * my);
            if (m > cap) {
                mx *= cap / m;
                my *= cap / m;
            }
            L.px[lane] = wrap_box(L.px[lane] + mx, R.L);
            L.py[lane] = wrap_box(L.py[lane] + my, R.L);
        }
        wave_sync();
    }
}

// One active catalyst's reactions (synthesis, then replication) on one
// wavefront; G is the wavefront's gather list.
__device__ void react_wave(const RepInfo &R, const RepStep &S, const ReactBufs &B, GBead *G, WaveLds &L, int k,
                           int r) {
    const int lane = threadIdx.x & 63;
    const uint32_t wk = B.gt[k], tk = wk >> kGidBits, g = wk & kGidMask;
    if (tk != OC_KA && tk != OC_KB) return;
    const int var = tk == OC_KB;
    const double us = hash_u01(S.ksyn, g), ur = hash_u01(S.krep, g);
    const int ml = R.mlen, nsynN = R.nsynN;
    const double2 pk = B.xy[k];
    const int kcx = cell_x(R, pk.x), kcy = cell_y(R, pk.y);
    const int n = gather(R, B, pk, kcx, kcy, G);
    if (n < 0) {
        if (lane == 0) atomicOr(B.err, ERR_CATCAP);
        return;
    }
    wave_sync();
    const int nn = near_wave(R, G, n, k, pk, kcx, kcy, OC_N, L, L.nb, nsynN + 2);
    if (nn < 0) {
        if (lane == 0) atomicOr(B.err, ERR_CATCAP);
        return;
    }
    if (nn < 2) return;
    int used = 0;
    int nw = 0;
    if (nn >= nsynN && nsynN < ml) nw = near_wave(R, G, n, k, pk, kcx, kcy, OC_W, L, L.nbw, ml - nsynN);
    if (nw < 0) {
        if (lane == 0) atomicOr(B.err, ERR_CATCAP);
        return;
    }
    if (nn >= nsynN && nsynN + nw >= ml && us < R.p_syn[var]) {
        // the beads of the new lipid, in the CPU engine's order
        if (lane < ml) {
            const int q = lane < nsynN ? L.nb[lane] : L.nbw[lane - nsynN];
            L.sl[lane] = G[q].slot;
            L.scell[lane] = G[q].cell;
            G[q].skip = 1;
        }
        wave_sync();
        double cxm = 0, cym = 0, mvx = 0, mvy = 0, mtx = 0, mty = 0;
        for (int b = 0; b < ml; b++) {
            const int j = L.sl[b];
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
        // the overlap partners of every orientation lie within d + 1 of the
        // catalyst (d: r_cat plus the template's extent): an LDS copy
        int no = 0;
        if (R.syn_tries > 1) {
            const double ro2 = R.ro2;
            for (int q0 = 0; q0 < n; q0 += 64) {
                const int q = q0 + lane;
                bool keep = false;
                if (q < n && !G[q].skip) {
                    const double dx = minimg(G[q].x - pk.x, R.L, R.halfL), dy = minimg(G[q].y - pk.y, R.L, R.halfL);
                    keep = dx * dx + dy * dy < ro2;
                }
                const uint64_t mk = __ballot(keep);
                const int at = no + __popcll(mk & lanes_below());
                if (keep && at < kOv) {
                    L.ox[at] = G[q].x;
                    L.oy[at] = G[q].y;
                    L.ow[at] = G[q].w;
                    L.ocell[at] = G[q].cell;
                }
                no += __popcll(mk);
            }
            if (no > kOv) no = 0;  // too many: the tries read the gather list
            wave_sync();
        }
        double ux = 1.0, uy = 0.0, ubest = INFINITY;
        for (int tr = 0; tr < R.syn_tries; tr++) {
            const double phi = 2.0 * M_PI * hash_u01(hash_key(R.seed, S.step, 40 + (uint64_t)tr), g);
            const double cx_ = cos(phi), cy_ = sin(phi);
            const double U = R.syn_tries > 1 ? overlap_wave(R, G, n, L, no, cxm, cym, cx_, cy_) : 0.0;
            if (U < ubest) {
                ubest = U;
                ux = cx_;
                uy = cy_;
            }
        }
        if (lane < ml) {
            const int j = L.sl[lane];
            const double2 p = chain_pos(R, lane, cxm, cym, ux, uy);
            L.px[lane] = p.x;
            L.py[lane] = p.y;
            L.pcx[lane] = cell_x(R, p.x);
            L.pcy[lane] = cell_y(R, p.y);
            B.v[j] = make_double2(mvx, mvy);
            B.vt[j] = make_float2((float)mtx, (float)mty);
            B.gt[j] = (B.gt[j] & kGidMask) | ((uint32_t)(lane < R.nH ? OC_H : OC_T) << kGidBits);
            B.mp[j] = (int8_t)lane;
            B.lnk[j] = lane + 1 < ml ? L.sl[lane + 1] : -1;
            B.prv[j] = lane > 0 ? L.sl[lane - 1] : -1;
        }
        wave_sync();
        if (R.syn_relax > 0) relax_wave(R, G, n, L);
        if (lane < ml) {
            const int j = L.sl[lane], oc = L.scell[lane];
            const double2 p = make_double2(L.px[lane], L.py[lane]);
            B.xy[j] = p;
            new_key_from(R, r, j, oc / R.ncy, oc % R.ncy, p, B.key, B.rmax, B.moved);
        }
        if (lane == 0 && B.record)
            push_event(B.ev, B.evn, B.evcap, B.err, S.step, OC_EV_SYN, g, (uint32_t)var, r, pk.x, pk.y);
        used = nsynN;
    }
    if (nn - used >= 2 && ur < R.p_rep[var]) {
        int nv = var, mut = 0;
        if (hash_u01(S.kmu, g) < R.mu) {
            nv = 1 - var;
            mut = 1;
        }
        if (lane == 0) {
            const int j1 = G[L.nb[used]].slot, j2 = G[L.nb[used + 1]].slot;
            B.gt[j1] = (B.gt[j1] & kGidMask) | ((uint32_t)(nv ? OC_KB : OC_KA) << kGidBits);
            B.gt[j2] = (B.gt[j2] & kGidMask) | ((uint32_t)OC_W << kGidBits);
            if (B.record)
                push_event(B.ev, B.evn, B.evcap, B.err, S.step, mut ? OC_EV_MUT : OC_EV_REP, B.gt[j1] & kGidMask,
                           (uint32_t)nv, r, pk.x, pk.y);
        }
    }
    wave_sync();
}

// two catalysts whose cells are closer than (cfx, cfy) may read or write each
// other's beads (periodic cell distance)
__device__ __forceinline__ bool reach_overlap(const RepInfo &R, int2 a, int2 b) {
    int dx = abs(a.x - b.x), dy = abs(a.y - b.y);
    dx = min(dx, R.ncx - dx);
    dy = min(dy, R.ncy - dy);
    return dx <= R.cfx && dy <= R.cfy;
}

__device__ __forceinline__ bool prio_less(uint64_t ka, uint32_t ga, uint64_t kb, uint32_t gb) {
    return ka != kb ? ka < kb : ga < gb;
}

// The replica's active catalysts in priority order (key hash_u64(kpri, gid),
// then gid), with their cells; a list over capacity or a step that moved a
// bead by the chemistry margin is an error (no reaction runs).
__global__ __launch_bounds__(kPrepThr) void k_react_prep(const RepInfo *__restrict__ ri,
                                                         const RepStep *__restrict__ rs, ReactBufs B) {
    const int r = blockIdx.x, t = threadIdx.x;
    const RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.react) return;
    __shared__ uint64_t s_key[kCatCap];
    __shared__ int32_t s_slot[kCatCap];
    const int n = B.catn[r];
    const bool bad_n = n > kCatCap, bad_d = S.disp != 0;
    if (n == 0 && !bad_d) return;
    __syncthreads();  // every thread has read catn
    if (bad_n || bad_d) {
        if (t == 0) {
            atomicOr(B.err, bad_n ? ERR_CATCAP : ERR_DISP);
            B.catn[r] = 0;
        }
        return;
    }
    int P = 1;
    while (P < n) P <<= 1;
    int32_t *cat = B.cat + (size_t)r * kCatCap;
    for (int i = t; i < P; i += kPrepThr) {
        if (i < n) {
            const int s = cat[i];
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
            for (int i = t; i < P; i += kPrepThr) {
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
    for (int i = t; i < n; i += kPrepThr) {
        const int s = s_slot[i];
        const double2 p = B.xy[s];
        cat[i] = s;
        B.ccell[(size_t)r * kCatCap + i] = make_int2(cell_x(R, p.x), cell_y(R, p.y));
        B.cdone[(size_t)r * kCatCap + i] = 0;
    }
}

// Catalysts with no lower-ranked catalyst within reach: all react at once,
// one wavefront each (they touch disjoint beads, and nothing before them in
// the priority order touches theirs).
__global__ __launch_bounds__(256) void k_react_free(const RepInfo *__restrict__ ri, const RepStep *__restrict__ rs,
                                                    ReactBufs B, int nrep) {
    __shared__ WaveLds s_w[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + wv, W = gridDim.x * 4;
    GBead *G = B.gs + (size_t)gw * B.gcap;
    for (int r = 0; r < nrep; r++) {
        const RepStep &S = rs[r];
        const RepInfo &R = ri[r];
        if (!S.rchem_now || !R.react) continue;
        const int n = B.catn[r];
        const int32_t *cat = B.cat + (size_t)r * kCatCap;
        const int2 *cc = B.ccell + (size_t)r * kCatCap;
        for (int i = gw; i < n; i += W) {
            const int2 ci = cc[i];
            bool hit = false;
            for (int j0 = 0; j0 < i && !hit; j0 += 64) {
                const int j = j0 + lane;
                hit = __ballot(j < i && reach_overlap(R, ci, cc[j])) != 0;
            }
            if (hit) continue;
            react_wave(R, S, B, G, s_w[wv], cat[i], r);
            if (lane == 0) B.cdone[(size_t)r * kCatCap + i] = 1;
        }
    }
}

// The other catalysts of a replica, in rounds (one block of two
// wavefronts): a catalyst runs in a round when no lower-ranked catalyst
// within its reach is still waiting; the lowest waiting one always runs, so
// the rounds end.  Clears the replica's list for the next check.
__global__ __launch_bounds__(kRestThr) void k_react_rest(const RepInfo *__restrict__ ri,
                                                         const RepStep *__restrict__ rs, ReactBufs B, int wbase) {
    const int r = blockIdx.x, t = threadIdx.x, wv = t >> 6;
    const RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.react) return;
    __shared__ int2 s_c[kCatCap];
    __shared__ uint8_t s_done[kCatCap], s_run[kCatCap];
    __shared__ WaveLds s_w[kRestThr / 64];
    __shared__ int s_left;
    const int n = B.catn[r];
    __syncthreads();  // every thread has read catn
    GBead *G = B.gs + (size_t)(wbase + r * (kRestThr / 64) + wv) * B.gcap;
    const int32_t *cat = B.cat + (size_t)r * kCatCap;
    for (int i = t; i < n; i += kRestThr) {
        s_c[i] = B.ccell[(size_t)r * kCatCap + i];
        s_done[i] = B.cdone[(size_t)r * kCatCap + i];
    }
    __syncthreads();
    for (;;) {
        for (int i = t; i < n; i += kRestThr) {
            s_run[i] = 0;
            if (s_done[i]) continue;
            bool ready = true;
            for (int j = 0; j < i && ready; j++)
                if (!s_done[j] && reach_overlap(R, s_c[i], s_c[j])) ready = false;
            s_run[i] = ready ? 1 : 0;
        }
        if (t == 0) s_left = 0;
        __syncthreads();
        // wavefront w takes the ready catalysts w, w + kRestThr / 64, ... (in list order)
        int seen = 0;
        for (int i0 = 0; i0 < n; i0 += 64) {
            const int i = i0 + (t & 63);
            const uint64_t m = __ballot(i < n && s_run[i]);
            uint64_t mm = m;
            while (mm) {
                const int b = __ffsll((unsigned long long)mm) - 1;
                mm &= mm - 1;
                if (seen % (kRestThr / 64) == wv) react_wave(R, S, B, G, s_w[wv], cat[i0 + b], r);
                seen++;
            }
        }
        __syncthreads();
        for (int i = t; i < n; i += kRestThr) {
            if (s_run[i]) s_done[i] = 1;
            if (!s_done[i]) atomicOr(&s_left, 1);
        }
        __syncthreads();
        if (!s_left) break;
        __syncthreads();
    }
    if (t == 0) B.catn[r] = 0;
}

// Feed cap, first pass: the number of N after the decay pass (per bead
// block) and the keys of the feed candidates (the replica's key list, any
// order).
__global__ __launch_bounds__(kBeadBlock) void k_rfeed_count(const RepInfo *__restrict__ ri,
                                                            const RepStep *__restrict__ rs,
                                                            const int32_t *__restrict__ bblk,
                                                            const double2 *__restrict__ xy,
                                                            const uint32_t *__restrict__ gt,
                                                            uint64_t *__restrict__ fkey, int2 *__restrict__ fcnt,
                                                            int32_t *__restrict__ npart) {
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
    __shared__ int s_n[kBeadBlock / 64];
    const int nN = __popcll(__ballot(isN));
    if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = nN;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sN = 0;
        for (int w = 0; w < kBeadBlock / 64; w++) sN += s_n[w];
        npart[blockIdx.x] = sN;
    }
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
                                                   const uint64_t *__restrict__ fkey, int2 *__restrict__ fcnt,
                                                   const int32_t *__restrict__ npart) {
    const int r = blockIdx.x, t = threadIdx.x;
    RepStep &S = rs[r];
    const RepInfo &R = ri[r];
    if (!S.rchem_now || !R.fcap_on) return;
    __shared__ unsigned int hist[256];
    __shared__ uint64_t s_pre;
    __shared__ int s_rank, s_nN;
    __shared__ bool s_one;
    if (t == 0) s_nN = 0;
    __syncthreads();
    {
        const int nb = (R.N + kBeadBlock - 1) / kBeadBlock;
        int sN = 0;
        for (int b = t; b < nb; b += 256) sN += npart[R.bblk0 + b];
        atomicAdd(&s_nN, sN);
    }
    __syncthreads();
    const int2 c = fcnt[r];
    const int nc = c.y, room = max(0, R.fcap_n - s_nN);
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
                    s_one = hist[b] == 1u;
                    break;
                }
                cum += (int)hist[b];
            }
        }
        __syncthreads();
        pre = s_pre;
        rank = s_rank;
        mask |= 0xffull << sh;
        const bool one = s_one;
        __syncthreads();
        if (one) {
            // a single key has these leading digits: it is the threshold
            for (int q = t; q < nc; q += 256) {
                const uint64_t k = K[q];
                if ((k & mask) == pre) s_pre = k;
            }
            __syncthreads();
            pre = s_pre;
            break;
        }
    }
    if (t == 0) {
        S.feed_add = room;
        S.feed_thr = pre;
        fcnt[r] = make_int2(0, 0);
    }
}

}  // namespace oc
