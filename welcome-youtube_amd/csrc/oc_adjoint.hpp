// oc_adjoint.hpp -- gfx950 kernels of the window gradient (no counterpart in
// the reference): d loss / d (a_WW, a_WH, a_WT, a_HH, a_HT, a_TT, gamma) for a
// smooth radial-distribution loss, the noise stream held fixed.
//
// Reverse sweep over the Groot-Warren step (f_n = F(x_n, v~_n)):
//   X_n += Jx^T F_n ; T_n = Jv^T F_n ; theta_bar += Jtheta^T F_n
//   X_{n-1} = X_n + dL/dx_{n-1} ; V_{n-1} = dt X_n + T_n + V_n ;
//   F_{n-1} = dt^2 X_n + (1/2 + lambda) dt T_n + dt V_n
// carried back through each step's sort permutation.  Pair arithmetic is FP32
// (as the forward), per-bead sums and the parameter gradients FP64.
#pragma once
#include "oc_kernels.hpp"

namespace oc {

struct RdfArgs {
    int32_t ta, tb, nbins;
    float rmax, width, dr, cut;
    float idr;  // 1 / dr: the bin range of a pair; a bin at its ends has |t| >= 1 and adds nothing
    int32_t Rc;  // columns searched on each side
};

constexpr double kFix = 4294967296.0;  // fixed-point scale of the RDF histogram (2^32)

constexpr int kRdfChunk = 64;  // beads per block of the bead-block reverse step
constexpr int kRdfSub = kBeadBlock / kRdfChunk;
constexpr int kAdjG = 4;  // lanes per bead of the bead-block reverse step
constexpr int kRdfMaxCol = 9;  // columns of the window (Rc <= 4)
constexpr int kTlG = 4;   // lanes per listed bead of the RDF kernels

// One recorded state of the window gradient's tape, in that state's slot
// order: cell starts, positions, gt words, the RDF gradient (output), and the
// type list of the RDF: for every column, its beads of type ta or tb in slot
// order, stored from the column's first slot (lxy, lgt, lidx = slot; lidx = -1
// past the column's list), with lcs[col0 + col * (ncy + 1) + k] the list
// position of row k's first listed bead (k = ncy: the column's list end).
struct TapeRef {
    const int32_t *cs;
    const double2 *xy;
    const uint32_t *gt;
    double2 *rg;
    double2 *lxy;
    uint32_t *lgt;
    int32_t *lidx, *lcs;
};

// The tape's list arrays as global-memory pointers for the kernels that read
// them: TapeRef's pointers are loaded from memory, so the compiler cannot tell
// their address space and emits flat loads, which also count against the LDS
// counter, so that every wait for a histogram atomic waits for the loads too.
template <class T>
using gptr = const __attribute__((address_space(1))) T *;
template <class T>
using gwptr = __attribute__((address_space(1))) T *;
typedef double dvec2 __attribute__((ext_vector_type(2)));  // a position as a plain vector (one 16-byte load)
struct TapeG {
    gptr<int32_t> lcs, lidx;
    gptr<dvec2> lxy;
    gptr<uint32_t> lgt;
};
__device__ __forceinline__ double2 ldxy(gptr<dvec2> p, int i) {
    const dvec2 v = p[i];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ TapeG tape_g(const TapeRef &S) {
    TapeG G;
    G.lcs = (gptr<int32_t>)S.lcs;
    G.lidx = (gptr<int32_t>)S.lidx;
    G.lxy = (gptr<dvec2>)S.lxy;
    G.lgt = (gptr<uint32_t>)S.lgt;
    return G;
}

__device__ __forceinline__ int block_excl_scan256(int v, int *tmp, int &total) {
    // exclusive scan over a 256-thread block; tmp: 4 ints of LDS
    const int inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) tmp[w] = inc;
    __syncthreads();
    int off = 0;
    for (int k = 0; k < w; k++) off += tmp[k];
    total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();
    return off + inc - v;
}

// Type list of every tape state of a segment: block (column, state).  Rows
// are taken 256 at a time, one thread each: count the row's listed beads, scan,
// then write them in slot order.  With zero_rg the other beads' RDF gradient
// is set to zero.
__global__ __launch_bounds__(256) void k_tlist(const RepInfo *__restrict__ ri, const int2 *__restrict__ cols,
                                               const TapeRef *__restrict__ tape, RdfArgs A, int zero_rg) {
    __shared__ int s_tmp[4];
    const TapeRef S = tape[blockIdx.y];
    // the tape's arrays as global-memory pointers (no flat accesses)
    const gptr<int32_t> cs = (gptr<int32_t>)S.cs;
    const gptr<uint32_t> gt = (gptr<uint32_t>)S.gt;
    const gptr<dvec2> xy = (gptr<dvec2>)S.xy;
    const gwptr<int32_t> lcs = (gwptr<int32_t>)S.lcs, lidx = (gwptr<int32_t>)S.lidx;
    const gwptr<uint32_t> lgt = (gwptr<uint32_t>)S.lgt;
    const gwptr<dvec2> lxy = (gwptr<dvec2>)S.lxy, rg = (gwptr<dvec2>)S.rg;
    const int2 rc = cols[blockIdx.x];  // (replica, column)
    const RepInfo &R = ri[rc.x];
    const int cb = R.coff + rc.y * R.ncy, lb = R.coff + R.col0 + rc.y * (R.ncy + 1);
    const int s1 = cs[cb + R.ncy];
    int run = cs[cb];
    for (int k0 = 0; k0 < R.ncy; k0 += 256) {
        const int k = k0 + (int)threadIdx.x;
        int a = 0, e = 0, cnt = 0;
        if (k < R.ncy) {
            a = cs[cb + k];
            e = cs[cb + k + 1];
            for (int d = a; d < e; d++) {
                const int t = (int)(gt[d] >> kGidBits);
                cnt += t == A.ta || t == A.tb;
            }
        }
        int total;
        const int base = run + block_excl_scan256(cnt, s_tmp, total);
        if (k < R.ncy) {
            lcs[lb + k] = base;
            int q = base;
            for (int d = a; d < e; d++) {
                const uint32_t g = gt[d];
                const int t = (int)(g >> kGidBits);
                if (t == A.ta || t == A.tb) {
                    lxy[q] = xy[d];
                    lgt[q] = g;
                    lidx[q] = d;
                    q++;
                } else if (zero_rg) {
                    rg[d] = dvec2{0.0, 0.0};
                }
            }
        }
        run += total;
    }
    if (threadIdx.x == 0) lcs[lb + R.ncy] = run;
    for (int q = run + (int)threadIdx.x; q < s1; q += 256) lidx[q] = -1;
}

// The listed beads within `cut` of listed bead i (cut < L/2 - 1): columns
// cx-Rc..cx+Rc and the rows of [y - cut, y + cut], each column as up to three
// contiguous list ranges (rows wrapped below, in the box, wrapped above).  The
// kTlG lanes of a group deal the candidates among themselves: a pass takes
// offset gl + u*kTlG of every column, loading the gt words of all columns,
// then the positions (lanes without a candidate re-read bead i), so a pass
// costs two memory latencies.  pre(j, gt) selects a candidate and fn(j, dx, dy)
// gets the FP32 image difference.
template <typename Pre, typename Fn>
__device__ __forceinline__ void visit_list(const RepInfo &R, const TapeRef &S, int i, double2 p, float cut, int Rc,
                                           int gl, Pre pre, Fn fn) {
    const TapeG G = tape_g(S);
    const int cx = (int)(p.x * R.icsx) >= R.ncx ? R.ncx - 1 : (int)(p.x * R.icsx);
    const int khi = (int)((p.y + cut) * R.icsy);
    const int klo = (int)((p.y - cut) * R.icsy + 4 * R.ncy) - 4 * R.ncy;
    const int pa1 = max(klo, 0), pb1 = min(khi, R.ncy - 1);
    int j0[kRdfMaxCol], j1[kRdfMaxCol], cbs[kRdfMaxCol];
    double sxq[kRdfMaxCol];
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) {
        const int colu = cx + q - Rc;
        const int col = colu < 0 ? colu + R.ncx : (colu >= R.ncx ? colu - R.ncx : colu);
        sxq[q] = colu < 0 ? -R.L : (colu >= R.ncx ? R.L : 0.0);
        cbs[q] = R.coff + R.col0 + col * (R.ncy + 1);
        const bool act = q <= 2 * Rc;  // columns past 2 Rc + 1 are empty
        j0[q] = act ? G.lcs[cbs[q] + pa1] : 0;
        j1[q] = act ? G.lcs[cbs[q] + pb1 + 1] : 0;
    }
    int lmax = 0;
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) lmax = max(lmax, j1[q] - j0[q]);
    for (int off = gl; off < lmax; off += kTlG) {
        int jq[kRdfMaxCol];
        uint32_t gq[kRdfMaxCol];
#pragma unroll
        for (int q = 0; q < kRdfMaxCol; q++) {
            const int j = j0[q] + off;
            jq[q] = j < j1[q] ? j : i;
            gq[q] = G.lgt[jq[q]];
        }
        double2 pq[kRdfMaxCol];
#pragma unroll
        for (int q = 0; q < kRdfMaxCol; q++) {
            if (jq[q] == i || !pre(gq[q])) jq[q] = -1;
            pq[q] = ldxy(G.lxy, jq[q] >= 0 ? jq[q] : i);
        }
#pragma unroll
        for (int q = 0; q < kRdfMaxCol; q++)
            if (jq[q] >= 0) fn((float)((p.x - pq[q].x) - sxq[q]), (float)(p.y - pq[q].y));
    }
    // rows wrapped at the periodic seam (beads within cut of y = 0 or L)
    if (klo >= 0 && khi < R.ncy) return;
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) {
        if (q > 2 * Rc) continue;
        for (int piece = 0; piece < 2; piece++) {
            int a, e;
            double sy;
            if (piece == 0) {
                if (klo >= 0) continue;
                a = G.lcs[cbs[q] + klo + R.ncy]; e = G.lcs[cbs[q] + R.ncy]; sy = -R.L;  // rows wrapped from the top
            } else {
                if (khi < R.ncy) continue;
                a = G.lcs[cbs[q]]; e = G.lcs[cbs[q] + khi - R.ncy + 1]; sy = R.L;       // rows wrapped from the bottom
            }
            for (int j = a + gl; j < e; j += kTlG)
                if (j != i && pre(G.lgt[j])) {
                    const double2 qq = ldxy(G.lxy, j);
                    fn((float)((p.x - qq.x) - sxq[q]), (float)((p.y - qq.y) - sy));
                }
        }
    }
}

// list positions of block blockIdx.x (kBeadBlock slots of one replica) that
// hold a listed bead, into s_list; returns the replica
__device__ __forceinline__ int list_block(const RepInfo *ri, const int32_t *bblk, const TapeRef &S, int *s_list,
                                          int *s_n) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc < R.N && tape_g(S).lidx[R.off + loc] >= 0) s_list[atomicAdd(s_n, 1)] = R.off + loc;
    __syncthreads();
    return r;
}

// Smooth RDF histogram of every tape state of a segment: block (bead block,
// state).  Each listed bead is handled by a group of kTlG lanes;
// contributions go to per-wavefront LDS fixed-point histograms (integer
// atomics: order independent), summed into the block's partial
// part[(bin * states + state) * blocks + block].
__global__ __launch_bounds__(kBeadBlock) void k_rdf(const RepInfo *__restrict__ ri, const int32_t *__restrict__ bblk,
                                                    const TapeRef *__restrict__ tape, RdfArgs A,
                                                    unsigned long long *__restrict__ part) {
    __shared__ unsigned long long h[kBeadBlock / 64][256];
    __shared__ int s_list[kBeadBlock];
    __shared__ int s_n;
    const TapeRef S = tape[blockIdx.y];
    for (int b = threadIdx.x; b < (kBeadBlock / 64) * 256; b += kBeadBlock) h[b >> 8][b & 255] = 0ull;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    list_block(ri, bblk, S, s_list, &s_n);
    const RepInfo &R = ri[bblk[blockIdx.x]];
    unsigned long long *hw = h[threadIdx.x >> 6];
    const int gl = threadIdx.x & (kTlG - 1);
    const float pref = 35.0f / (32.0f * A.width), iw = 1.0f / A.width;
    for (int e = threadIdx.x / kTlG; e < s_n; e += kBeadBlock / kTlG) {
        const int i = s_list[e];
        const uint32_t gti = tape_g(S).lgt[i], ti = gti >> kGidBits, gi = gti & kGidMask;
        visit_list(
            R, S, i, ldxy(tape_g(S).lxy, i), A.cut, A.Rc, gl,
            [&](uint32_t gtj) {
                const uint32_t tj = gtj >> kGidBits;
                if (A.ta == A.tb) return gi < (gtj & kGidMask);  // each pair once
                return (int)ti == A.ta && (int)tj == A.tb;
            },
            [&](float dx, float dy) {
                const float r2 = dx * dx + dy * dy;
                if (r2 >= A.cut * A.cut || r2 == 0.0f) return;
                const float rr = sqrtf(r2);
                const int b0 = max(0, (int)floorf((rr - A.width) * A.idr));
                const int b1 = min(A.nbins - 1, (int)floorf((rr + A.width) * A.idr));
                for (int b = b0; b <= b1; b++) {
                    const float t = (rr - (b + 0.5f) * A.dr) * iw;
                    if (fabsf(t) >= 1.0f) continue;
                    const float q = 1.0f - t * t;
                    atomicAdd(&hw[b], (unsigned long long)((double)(pref * q * q * q) * kFix + 0.5));
                }
            });
    }
    __syncthreads();
    for (int b = threadIdx.x; b < A.nbins; b += blockDim.x) {
        unsigned long long v = 0;
        for (int w = 0; w < kBeadBlock / 64; w++) v += h[w][b];
        part[((size_t)b * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = v;  // the reduction reads coalesced
    }
}

// window accumulator += sum of the block partials of each replica over nst
// states: one block per (bin, replica), a strided sum then an LDS tree
// (integer: exact)
__global__ __launch_bounds__(256) void k_rdf_reduce(const RepInfo *__restrict__ ri, int nblk, int nst,
                                                    const unsigned long long *__restrict__ part,
                                                    unsigned long long *__restrict__ acc, int nbins) {
    __shared__ unsigned long long red[256];
    const int b = blockIdx.x, r = blockIdx.y;
    const RepInfo &R = ri[r];
    const int k0 = R.bblk0, nb = (R.N + kBeadBlock - 1) / kBeadBlock;
    unsigned long long v = 0;
    for (int y = 0; y < nst; y++)
        for (int k = threadIdx.x; k < nb; k += 256) v += part[((size_t)b * nst + y) * nblk + k0 + k];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[(size_t)r * nbins + b] += red[0];
}

// rg_i = d loss / d x_i of the RDF at tape state blockIdx.y for the listed
// beads (the others were zeroed with the list); gw[r][b] = dL/dg_b * C_b.  All
// states of a recomputed segment go in one launch.  Each listed bead is
// handled by kTlG lanes; the group's partial sums are combined by a fixed
// shuffle tree.
__global__ __launch_bounds__(kBeadBlock) void k_rdf_adj(const RepInfo *__restrict__ ri,
                                                        const int32_t *__restrict__ bblk,
                                                        const TapeRef *__restrict__ tape, RdfArgs A,
                                                        const double *__restrict__ gw) {
    __shared__ float w_s[256];
    __shared__ int s_list[kBeadBlock];
    __shared__ int s_n;
    const TapeRef S = tape[blockIdx.y];
    const int r = bblk[blockIdx.x];
    for (int b = threadIdx.x; b < A.nbins; b += blockDim.x) w_s[b] = (float)gw[(size_t)r * A.nbins + b];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    list_block(ri, bblk, S, s_list, &s_n);
    const RepInfo &R = ri[r];
    const int gl = threadIdx.x & (kTlG - 1);
    const float pref = 35.0f / (32.0f * A.width), iw = 1.0f / A.width;
    // rounds over the list; every lane of a group takes part in the shuffles
    const int ng = kBeadBlock / kTlG;
    for (int e0 = 0; e0 < s_n; e0 += ng) {
        const int e = e0 + (int)threadIdx.x / kTlG;
        double ax = 0.0, ay = 0.0;
        int i = -1;
        if (e < s_n) {
            i = s_list[e];
            const uint32_t ti = tape_g(S).lgt[i] >> kGidBits;
            visit_list(
                R, S, i, ldxy(tape_g(S).lxy, i), A.cut, A.Rc, gl,
                [&](uint32_t gtj) {
                    const uint32_t tj = gtj >> kGidBits;
                    return ((int)ti == A.ta && (int)tj == A.tb) || ((int)ti == A.tb && (int)tj == A.ta);
                },
                [&](float dx, float dy) {
                    const float r2 = dx * dx + dy * dy;
                    if (r2 >= A.cut * A.cut || r2 == 0.0f) return;
                    const float ir = __builtin_amdgcn_rsqf(r2), rr = r2 * ir;
                    const int b0 = max(0, (int)floorf((rr - A.width) * A.idr));
                    const int b1 = min(A.nbins - 1, (int)floorf((rr + A.width) * A.idr));
                    float ds = 0.0f;
                    for (int b = b0; b <= b1; b++) {
                        const float t = (rr - (b + 0.5f) * A.dr) * iw;
                        if (fabsf(t) >= 1.0f) continue;
                        const float q = 1.0f - t * t;
                        ds += w_s[b] * pref * 3.0f * q * q * (-2.0f * t * iw);
                    }
                    ax += (double)(ds * dx * ir);
                    ay += (double)(ds * dy * ir);
                });
        }
        for (int o = kTlG / 2; o > 0; o >>= 1) {
            ax += __shfl_xor(ax, o, kTlG);
            ay += __shfl_xor(ay, o, kTlG);
        }
        if (gl == 0 && i >= 0) S.rg[S.lidx[i]] = make_double2(ax, ay);
    }
}

// Several device-to-device copies in one launch (the snapshots of the window
// gradient): 16-byte words where source and destination are aligned, bytes
// for the rest.
constexpr int kCopyMax = 12;
struct CopyJobs {
    char *dst[kCopyMax];
    const char *src[kCopyMax];
    int64_t bytes[kCopyMax];
    int n;
};
__global__ __launch_bounds__(256) void k_copy_multi(CopyJobs J) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, nth = (int64_t)gridDim.x * blockDim.x;
    for (int k = 0; k < J.n; k++) {
        char *d = J.dst[k];
        const char *sp = J.src[k];
        const int64_t nb = J.bytes[k];
        const bool al = (((uintptr_t)d | (uintptr_t)sp) & 15) == 0;
        const int64_t n16 = al ? nb >> 4 : 0;
        for (int64_t w = tid; w < n16; w += nth) ((uint4 *)d)[w] = ((const uint4 *)sp)[w];
        for (int64_t b = n16 * 16 + tid; b < nb; b += nth) d[b] = sp[b];
    }
}

__device__ __forceinline__ int theta_index(uint32_t a, uint32_t b) {
    if (a > b) { const uint32_t t = a; a = b; b = t; }
    if (b > 2) return -1;
    return a == 0 ? (int)b : (a == 1 ? 2 + (int)b : 5);
}

// Step parameters of steps base + first + y (y = blockIdx.y) into
// rsa[y * nrep + r]: every reverse step of a segment in one launch.
__global__ __launch_bounds__(64) void k_step_params_seg(const RepInfo *__restrict__ ri, RepStep *__restrict__ rsa,
                                                        int64_t first, const int64_t *__restrict__ base) {
    const int r = blockIdx.x, y = blockIdx.y;
    step_params_wave(ri[r], rsa[(size_t)y * gridDim.x + r], r, threadIdx.x, first + y, -1, base);
}

// One reverse step at state n (the tape's arrays of that state): with the
// force adjoint Fa,
//   X_n = X + Jx^T Fa (pair and bond terms) + G_n, T_n = Jv^T Fa,
//   theta_bar += Jtheta^T Fa (half of each pair: every pair is visited from
//   both beads; summed per tile in a fixed order),
// then X_{n-1}, V_{n-1}, F_{n-1} scattered to the order of step n-1 through
// the sort map (new slot i <- old slot src[i]).  Every term of bead i is
// summed by its own lane (gather form, no atomics).  Xo == nullptr: theta
// gradient only (the window's first force).
struct AdjStep {
    const RepStep *rs;  // step parameters of step n, per replica
    const int32_t *cs, *src, *lnk, *prv;
    const double2 *xy, *G;
    const float2 *vt;
    const uint32_t *gt;
    const int8_t *mp;
    const double2 *Fa, *X, *V;
    double2 *Xo, *Vo, *Fo;
    double *thp;  // per pair tile: running theta-gradient sums (8 doubles)
};

// One reverse step with kAdjG lanes per bead, bead blocks of kRdfChunk beads
// (the launch for small systems, whose pair tiles are too few to fill the
// chip): candidates read from global memory; theta summed per block.
__global__ __launch_bounds__(kBeadBlock) void k_adj_rstep(const RepInfo *__restrict__ ri,
                                                          const int32_t *__restrict__ bblk, AdjStep A) {
    // block b: beads [(b % kRdfSub) * kRdfChunk, +kRdfChunk) of bead block
    // b / kRdfSub, kAdjG lanes per bead
    const int bb = blockIdx.x / kRdfSub;
    const int r = bblk[bb];
    const RepInfo &R = ri[r];
    const RepStep &S = A.rs[r];
    const int32_t *__restrict__ cs = A.cs;
    const double2 *__restrict__ xy = A.xy;
    const float2 *__restrict__ vt = A.vt;
    const uint32_t *__restrict__ gt = A.gt;
    const double2 *__restrict__ Fa = A.Fa;
    const int loc = (bb - R.bblk0) * kBeadBlock + (blockIdx.x % kRdfSub) * kRdfChunk + (int)threadIdx.x / kAdjG;
    const int gl = threadIdx.x & (kAdjG - 1);
    const bool live = loc < R.N;
    const int i = R.off + (live ? loc : 0);
    const double2 pi = xy[i];
    const float2 vi = vt[i];
    const uint32_t gti = gt[i], ti = gti >> kGidBits, gi = gti & kGidMask;
    const double2 Fi = Fa[i];
    // the step-back inputs of bead i, loaded now so that they arrive during
    // the pair loop
    const bool step_back = A.Xo != nullptr;
    int pn = -1, pp = -1, od = 0;
    double2 xin = make_double2(0.0, 0.0), gin = xin, vin = xin;
    if (step_back) {
        pn = A.lnk[i];
        pp = A.mp[i] > 0 ? A.prv[i] : -1;
        od = A.src[i];
        xin = A.X[i];
        gin = A.G[i];
        vin = A.V[i];
    }
    const float gam = S.gamma, sdt = S.sdt, sdg = S.sdt / (2.0f * S.gamma);
    const int ncy = R.ncy, ncx = R.ncx;
    const int cx = (int)(pi.x * R.icsx) >= ncx ? ncx - 1 : (int)(pi.x * R.icsx);
    const int khi = (int)((pi.y + 1.0) * R.icsy);
    const int klo = (int)((pi.y - 1.0) * R.icsy + ncy) - ncy;
    double xa = 0.0, ya = 0.0, ta = 0.0, tb = 0.0;
    double th[7] = {0, 0, 0, 0, 0, 0, 0};
    auto one = [&](int j, double sx, double sy) {
        const double2 pj = xy[j];
        const float dx = (float)((pi.x - pj.x) - sx), dy = (float)((pi.y - pj.y) - sy);
        const float r2 = dx * dx + dy * dy;
        if (!(r2 < 1.0f && r2 > 1e-24f)) return;
        const float ir = __builtin_amdgcn_rsqf(r2);
        const float rr = r2 * ir, w = 1.0f - rr;
        const float ex = dx * ir, ey = dy * ir;
        const float2 vj = vt[j];
        const float ux = vi.x - vj.x, uy = vi.y - vj.y;
        const float eu = ex * ux + ey * uy;
        const uint32_t gtj = gt[j], tj = gtj >> kGidBits;
        const uint64_t hsh = pair_hash(S.sseed, gi, gtj & kGidMask);
        const float thn = (float)(int32_t)(uint32_t)(hsh >> 32) * kTheta32;
        const float a = S.a_eff[8 * ti + tj];
        const float m = w * (a - gam * w * eu + sdt * thn);
        const float dmr = -(a + sdt * thn) + 2.0f * gam * w * eu;
        const float dmu = -gam * w * w;
        const double2 Fj = Fa[j];
        const float lx = (float)(Fi.x - Fj.x), ly = (float)(Fi.y - Fj.y);
        const float el = ex * lx + ey * ly;
        const float pux = ux - eu * ex, puy = uy - eu * ey;
        const float plx = lx - el * ex, ply = ly - el * ey;
        xa += (double)((dmr * ex + dmu * pux * ir) * el + m * plx * ir);
        ya += (double)((dmr * ey + dmu * puy * ir) * el + m * ply * ir);
        ta += (double)(dmu * el * ex);
        tb += (double)(dmu * el * ey);
        const double ha = (double)(0.5f * w * el);
        const int q = theta_index(ti, tj);
        if (q == 0) th[0] += ha;
        else if (q == 1) th[1] += ha;
        else if (q == 2) th[2] += ha;
        else if (q == 3) th[3] += ha;
        else if (q == 4) th[4] += ha;
        else if (q == 5) th[5] += ha;
        th[6] += (double)(0.5f * el * (-w * w * eu + sdg * w * thn));
    };
    if (live) {
        // the in-box row range of each column first (loads issued together), then
        // rows wrapped at the periodic seam; candidates dealt to the group's lanes
        const int pa1 = max(klo, 0), pb1 = min(khi, ncy - 1);
        int j0[3], j1[3], cb[3];
        double sxs[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            int col = cx + c - 1;
            sxs[c] = col < 0 ? -R.L : (col >= ncx ? R.L : 0.0);
            col = col < 0 ? col + ncx : (col >= ncx ? col - ncx : col);
            cb[c] = R.coff + col * ncy;
            j0[c] = cs[cb[c] + pa1];
            j1[c] = cs[cb[c] + pb1 + 1];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (klo < 0) {
                const int e1 = cs[cb[c] + ncy];
                for (int j = cs[cb[c] + klo + ncy] + gl; j < e1; j += kAdjG) one(j, sxs[c], -R.L);
            }
            for (int j = j0[c] + gl; j < j1[c]; j += kAdjG) one(j, sxs[c], 0.0);
            if (khi >= ncy) {
                const int e1 = cs[cb[c] + khi - ncy + 1];
                for (int j = cs[cb[c]] + gl; j < e1; j += kAdjG) one(j, sxs[c], R.L);
            }
        }
    }
    // sums over the group (fixed shuffle tree), finished by its first lane
    for (int o = kAdjG / 2; o > 0; o >>= 1) {
        xa += __shfl_xor(xa, o, kAdjG);
        ya += __shfl_xor(ya, o, kAdjG);
        ta += __shfl_xor(ta, o, kAdjG);
        tb += __shfl_xor(tb, o, kAdjG);
#pragma unroll
        for (int q = 0; q < 7; q++) th[q] += __shfl_xor(th[q], o, kAdjG);
    }
    // theta gradient: the block's beads in a fixed tree, added to the block's
    // running sum thp[block] (steps in the order of the sweep)
    __shared__ double s_th[7][kRdfChunk];
    const int bead = (int)threadIdx.x / kAdjG;
    if (gl == 0)
#pragma unroll
        for (int q = 0; q < 7; q++) s_th[q][bead] = live ? th[q] : 0.0;
    __syncthreads();
    for (int w = kRdfChunk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < 7 * w) {
            const int q = threadIdx.x / w, k = threadIdx.x - q * w;
            s_th[q][k] += s_th[q][k + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) A.thp[(size_t)blockIdx.x * 8 + threadIdx.x] += s_th[threadIdx.x][0];
    if (!live || gl != 0 || !step_back) return;
    // harmonic bonds of bead i (FP64): the chain neighbours of a lipid bead,
    // the chemistry partner of another bead; with d = x_i - x_p and
    // l = Fa_i - Fa_p, X_i += g l + g' r (e.l) e, g = -kb (r - r0) / r,
    // g' = -kb r0 / r^2
    double2 qn = pi, fn = Fi, qp = pi, fp = Fi;
    if (pn >= 0) {
        qn = xy[pn];
        fn = Fa[pn];
    }
    if (pp >= 0) {
        qp = xy[pp];
        fp = Fa[pp];
    }
    auto bond = [&](double2 q, double2 Fp) {
        const double dx = minimg(pi.x - q.x, R.L, R.halfL), dy = minimg(pi.y - q.y, R.L, R.halfL);
        const double rr = sqrt(dx * dx + dy * dy);
        const double g = -R.kb * (rr - R.r0) / rr, gp = -R.kb * R.r0 / (rr * rr);
        const double ex = dx / rr, ey = dy / rr;
        const double lx = Fi.x - Fp.x, ly = Fi.y - Fp.y;
        const double el = ex * lx + ey * ly;
        xa += g * lx + gp * rr * el * ex;
        ya += g * ly + gp * rr * el * ey;
    };
    if (pp >= 0) bond(qp, fp);
    if (pn >= 0) bond(qn, fn);
    // + d loss / d x_n of the RDF, then the step back to n-1
    const double xx = xin.x + xa + gin.x, xy_ = xin.y + ya + gin.y;
    A.Xo[od] = make_double2(xx, xy_);
    A.Vo[od] = make_double2(R.dt * xx + ta + vin.x, R.dt * xy_ + tb + vin.y);
    A.Fo[od] = make_double2(R.dt2 * xx + R.cvt * ta + R.dt * vin.x, R.dt2 * xy_ + R.cvt * tb + R.dt * vin.y);
}

// One reverse step per pair tile (the force kernel's tiles: a column x a
// range of rows): the candidates of the tile's three columns -- position, v~,
// gid|type and the force adjoint Fa of state n -- are staged in LDS once, each
// lane takes one bead of the tile and sums the pair terms above over its
// window (per column and periodic piece, with the image shifts) from LDS; a
// tile whose candidates do not fit runs from global memory.  The bond terms,
// the RDF gradient and the step back through the sort map follow per bead.
// theta: each lane's sum over its beads, a fixed shuffle tree per wavefront,
// the wavefronts in order, added to the tile's running sum thp[tile].
template <int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) void k_adj_tile(const RepInfo *__restrict__ ri, const Tile *__restrict__ tiles,
                                                    int ntiles, int per_xcd, AdjStep A, uint32_t *__restrict__ err) {
    constexpr int kMaxRows = pair_max_rows(BLOCK);
    __shared__ double2 s_xy[CAP];
    __shared__ double2 s_fa[CAP];
    __shared__ float2 s_vt[CAP];
    __shared__ uint32_t s_gt[CAP];
    __shared__ int32_t s_row[3][kMaxRows + 1];
    __shared__ int32_t s_piece[3][3][3];  // [col][piece] = {first slot, length, staged start}
    __shared__ int32_t s_col[3];
    __shared__ int32_t s_tot;
    __shared__ double s_th[BLOCK / 64][8];
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    const Tile T = tiles[t];
    const RepInfo &R = ri[T.rep];
    const RepStep &S = A.rs[T.rep];
    const int ncy = R.ncy, ncx = R.ncx;
    const int M = (int)ceil(R.icsy) + 1;  // staging margin in rows
    const int kb0 = T.k0 - M;
    const int nrow = T.k1 - T.k0 + 2 * M + 1;
    const int32_t *__restrict__ cs = A.cs;
    if (nrow > kMaxRows) {  // the host sizes tiles so that this cannot happen
        if (threadIdx.x == 0) atomicOr(err, ERR_TILE);
        return;
    }
    // pieces of the staged rows lo..hi of each column: A = rows < 0 (wrapped
    // from the top), B = rows in the box, C = rows >= ncy (wrapped from the
    // bottom); staged in (column, piece) order
    if (threadIdx.x < 64) {
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
        }
        if (q == 8) s_tot = incl;
    }
    __syncthreads();
    const int total = s_tot;
    const bool stg = total <= CAP;
    // s_row[c][k - kb0]: staged (or piece-relative) index of the first bead of
    // unwrapped row k of column c
    for (int e = threadIdx.x; e < 3 * nrow; e += BLOCK) {
        const int c = e / nrow, k = kb0 + (e - c * nrow);
        const int pc = k < 0 ? 0 : (k < ncy ? 1 : 2);
        const int kk = k < 0 ? k + ncy : (k < ncy ? k : k - ncy);
        s_row[c][e - c * nrow] = s_piece[c][pc][2] + (cs[R.coff + s_col[c] * ncy + kk] - s_piece[c][pc][0]);
    }
    if (stg)
        for (int e = threadIdx.x; e < total; e += BLOCK) {
            int w = 0;  // the last piece starting at or before e
#pragma unroll
            for (int u = 1; u < 9; u++) w = e >= s_piece[u / 3][u % 3][2] ? u : w;
            const int src = s_piece[w / 3][w % 3][0] + (e - s_piece[w / 3][w % 3][2]);
            s_xy[e] = A.xy[src];
            s_vt[e] = A.vt[src];
            s_gt[e] = A.gt[src];
            s_fa[e] = A.Fa[src];
        }
    __syncthreads();

    const double L = R.L;
    const double sxc[3] = {T.cx == 0 ? -L : 0.0, 0.0, T.cx == ncx - 1 ? L : 0.0};
    const float gam = S.gamma, sdt = S.sdt, sdg = S.sdt / (2.0f * S.gamma);
    const int ibeg = s_row[1][T.k0 - kb0], iend = s_row[1][T.k1 - kb0];
    const int slot0 = s_piece[1][1][0] + (ibeg - s_piece[1][1][2]);  // T.k0 lies in piece B of column 1
    const bool step_back = A.Xo != nullptr;
    double th[7] = {0, 0, 0, 0, 0, 0, 0};
    auto run = [&](const double2 *XY, const float2 *VT, const uint32_t *GT, const double2 *FA) {
        for (int ii = ibeg + (int)threadIdx.x; ii < iend; ii += BLOCK) {
            const int i = slot0 + (ii - ibeg);
            const int is = stg ? ii : i;
            // the step-back inputs of bead i, loaded now so that they arrive
            // during the pair loop
            int pn = -1, pp = -1, od = 0;
            double2 xin = make_double2(0.0, 0.0), gin = xin, vin = xin;
            if (step_back) {
                pn = A.lnk[i];
                pp = A.mp[i] > 0 ? A.prv[i] : -1;
                od = A.src[i];
                xin = A.X[i];
                gin = A.G[i];
                vin = A.V[i];
            }
            const double2 pi = XY[is];
            const float2 vi = VT[is];
            const uint32_t gti = GT[is], ti = gti >> kGidBits, gi = gti & kGidMask;
            const double2 Fi = FA[is];
            int khi = (int)((pi.y + 1.0) * R.icsy);
            int klo = (int)((pi.y - 1.0) * R.icsy + ncy) - ncy;
            if (klo < kb0 || khi > T.k1 - 1 + M) {  // cannot happen for rows k0..k1-1 (margin M); guard
                atomicOr(err, ERR_TILE);
                klo = max(klo, kb0);
                khi = min(khi, T.k1 - 1 + M);
            }
            double xa = 0.0, ya = 0.0, ta = 0.0, tb = 0.0;
            auto one = [&](int j, double sx, double sy) {
                const double2 pj = XY[j];
                const float dx = (float)((pi.x - pj.x) - sx), dy = (float)((pi.y - pj.y) - sy);
                const float r2 = dx * dx + dy * dy;
                if (!(r2 < 1.0f && r2 > 1e-24f)) return;
                const float ir = __builtin_amdgcn_rsqf(r2);
                const float rr = r2 * ir, w = 1.0f - rr;
                const float ex = dx * ir, ey = dy * ir;
                const float2 vj = VT[j];
                const float ux = vi.x - vj.x, uy = vi.y - vj.y;
                const float eu = ex * ux + ey * uy;
                const uint32_t gtj = GT[j], tj = gtj >> kGidBits;
                const uint64_t hsh = pair_hash(S.sseed, gi, gtj & kGidMask);
                const float thn = (float)(int32_t)(uint32_t)(hsh >> 32) * kTheta32;
                const float a = S.a_eff[8 * ti + tj];
                const float m = w * (a - gam * w * eu + sdt * thn);
                const float dmr = -(a + sdt * thn) + 2.0f * gam * w * eu;
                const float dmu = -gam * w * w;
                const double2 Fj = FA[j];
                const float lx = (float)(Fi.x - Fj.x), ly = (float)(Fi.y - Fj.y);
                const float el = ex * lx + ey * ly;
                const float pux = ux - eu * ex, puy = uy - eu * ey;
                const float plx = lx - el * ex, ply = ly - el * ey;
                xa += (double)((dmr * ex + dmu * pux * ir) * el + m * plx * ir);
                ya += (double)((dmr * ey + dmu * puy * ir) * el + m * ply * ir);
                ta += (double)(dmu * el * ex);
                tb += (double)(dmu * el * ey);
                const double ha = (double)(0.5f * w * el);
                const int q = theta_index(ti, tj);
                if (q == 0) th[0] += ha;
                else if (q == 1) th[1] += ha;
                else if (q == 2) th[2] += ha;
                else if (q == 3) th[3] += ha;
                else if (q == 4) th[4] += ha;
                else if (q == 5) th[5] += ha;
                th[6] += (double)(0.5f * el * (-w * w * eu + sdg * w * thn));
            };
            for (int c = 0; c < 3; c++) {
                for (int q = 0; q < 3; q++) {
                    int pa, pb;
                    double sy;
                    if (q == 0) { pa = klo; pb = min(khi, -1); sy = -L; }
                    else if (q == 1) { pa = max(klo, 0); pb = min(khi, ncy - 1); sy = 0.0; }
                    else { pa = max(klo, ncy); pb = khi; sy = L; }
                    if (pa > pb) continue;
                    const int adj = stg ? 0 : s_piece[c][q][0] - s_piece[c][q][2];
                    const int ja = s_row[c][pa - kb0] + adj, jb = s_row[c][pb + 1 - kb0] + adj;
                    for (int j = ja; j < jb; j++) one(j, sxc[c], sy);
                }
            }
            if (!step_back) continue;
            // harmonic bonds of bead i (FP64): the chain neighbours of a lipid
            // bead, the chemistry partner of another bead; with d = x_i - x_p
            // and l = Fa_i - Fa_p, X_i += g l + g' r (e.l) e,
            // g = -kb (r - r0) / r, g' = -kb r0 / r^2
            double2 qn = pi, fn = Fi, qp = pi, fp = Fi;
            if (pn >= 0) {
                qn = A.xy[pn];
                fn = A.Fa[pn];
            }
            if (pp >= 0) {
                qp = A.xy[pp];
                fp = A.Fa[pp];
            }
            auto bond = [&](double2 qq, double2 Fp) {
                const double dx = minimg(pi.x - qq.x, R.L, R.halfL), dy = minimg(pi.y - qq.y, R.L, R.halfL);
                const double rr = sqrt(dx * dx + dy * dy);
                const double g = -R.kb * (rr - R.r0) / rr, gp = -R.kb * R.r0 / (rr * rr);
                const double ex = dx / rr, ey = dy / rr;
                const double lx = Fi.x - Fp.x, ly = Fi.y - Fp.y;
                const double el = ex * lx + ey * ly;
                xa += g * lx + gp * rr * el * ex;
                ya += g * ly + gp * rr * el * ey;
            };
            if (pp >= 0) bond(qp, fp);
            if (pn >= 0) bond(qn, fn);
            // + d loss / d x_n of the RDF, then the step back to n-1
            const double xx = xin.x + xa + gin.x, xy_ = xin.y + ya + gin.y;
            A.Xo[od] = make_double2(xx, xy_);
            A.Vo[od] = make_double2(R.dt * xx + ta + vin.x, R.dt * xy_ + tb + vin.y);
            A.Fo[od] = make_double2(R.dt2 * xx + R.cvt * ta + R.dt * vin.x, R.dt2 * xy_ + R.cvt * tb + R.dt * vin.y);
        }
    };
    if (stg) run(s_xy, s_vt, s_gt, s_fa);
    else run(A.xy, A.vt, A.gt, A.Fa);
    // theta: fixed shuffle tree per wavefront, then the wavefronts in order
#pragma unroll
    for (int q = 0; q < 7; q++) {
        double v = th[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        th[q] = v;
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 7; q++) s_th[threadIdx.x >> 6][q] = th[q];
    __syncthreads();
    if (threadIdx.x < 7) {
        double v = 0.0;
        for (int w = 0; w < BLOCK / 64; w++) v += s_th[w][threadIdx.x];
        A.thp[(size_t)t * 8 + threadIdx.x] += v;
    }
}

// per-replica theta gradient: the sum of the tile partials of the replica's
// tiles (trng[r] = first tile, count), strided then an LDS tree (fixed order,
// FP64)
__global__ __launch_bounds__(256) void k_theta_reduce(const int2 *__restrict__ trng, const double *__restrict__ thp,
                                                      double *__restrict__ out) {
    __shared__ double red[7][256];
    const int k0 = trng[blockIdx.x].x, nb = trng[blockIdx.x].y;  // the replica's tiles
    double sq[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < nb; k += 256) {
        const double *t = thp + (size_t)(k0 + k) * 8;
#pragma unroll
        for (int q = 0; q < 7; q++) sq[q] += t[q];
    }
#pragma unroll
    for (int q = 0; q < 7; q++) red[q][threadIdx.x] = sq[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 7; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 7) out[(size_t)blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_zero_f64(double *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace oc
