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
    int32_t Rc;  // columns searched on each side
};

constexpr double kFix = 4294967296.0;  // fixed-point scale of the RDF histogram (2^32)

// The beads within `cut` of bead i (cut < L/2 - 1): columns cx-Rc..cx+Rc and
// the rows of [y - cut, y + cut], each column as up to three contiguous slot
// ranges (rows wrapped below, in the box, wrapped above).  The kRdfG lanes of a
// group deal the candidates of each range among themselves; fn(j, dx, dy)
// gets the FP32 image difference.  Types are tested (tfilt) before the
// position is loaded.
constexpr int kRdfG = 16;
constexpr int kRdfMaxCol = 9;  // columns of the window (Rc <= 4)
constexpr int kRdfChunk = 64;  // beads per block of the RDF kernels (256 threads: 16 groups)
constexpr int kRdfSub = kBeadBlock / kRdfChunk;
constexpr int kAdjG = 4;  // lanes per bead of the pair adjoint (blocks of kRdfChunk beads)
template <typename Pre, typename Fn>
__device__ __forceinline__ void visit_wide(const RepInfo &R, const int32_t *cs, const double2 *xy, int i, double2 p,
                                           float cut, int Rc, int gl, Pre pre, Fn fn) {
    const int cx = (int)(p.x * R.icsx) >= R.ncx ? R.ncx - 1 : (int)(p.x * R.icsx);
    const int khi = (int)((p.y + cut) * R.icsy);
    const int klo = (int)((p.y - cut) * R.icsy + 4 * R.ncy) - 4 * R.ncy;
    // the in-box row range of every column first (all loads issued together)
    const int pa1 = max(klo, 0), pb1 = min(khi, R.ncy - 1);
    int j0[kRdfMaxCol], j1[kRdfMaxCol], cbs[kRdfMaxCol];
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) {
        const int dc = q - Rc;
        if (dc > Rc) break;
        int col = cx + dc;
        col = col < 0 ? col + R.ncx : (col >= R.ncx ? col - R.ncx : col);
        cbs[q] = R.coff + col * R.ncy;
        j0[q] = cs[cbs[q] + pa1];
        j1[q] = cs[cbs[q] + pb1 + 1];
    }
    double sxq[kRdfMaxCol];
    int lmax = 0;
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) {
        const int dc = q - Rc;
        if (dc > Rc) break;
        const int colu = cx + dc;
        sxq[q] = colu < 0 ? -R.L : (colu >= R.ncx ? R.L : 0.0);
        lmax = max(lmax, j1[q] - j0[q]);
    }
    // in-box rows: pass u takes offset gl + u*kRdfG of every column; a pass
    // issues the type tests of all columns, then the positions of the matches
    for (int off = gl; off < lmax; off += kRdfG) {
        int jq[kRdfMaxCol];
#pragma unroll
        for (int q = 0; q < kRdfMaxCol; q++) {
            jq[q] = -1;
            if (q - Rc > Rc) continue;
            const int j = j0[q] + off;
            if (j < j1[q] && j != i && pre(j)) jq[q] = j;
        }
#pragma unroll
        for (int q = 0; q < kRdfMaxCol; q++)
            if (jq[q] >= 0) {
                const double2 qq = xy[jq[q]];
                fn(jq[q], (float)((p.x - qq.x) - sxq[q]), (float)((p.y - qq.y)));
            }
    }
    // rows wrapped at the periodic seam (beads within cut of y = 0 or L)
    if (klo >= 0 && khi < R.ncy) return;
#pragma unroll
    for (int q = 0; q < kRdfMaxCol; q++) {
        if (q - Rc > Rc) break;
        for (int piece = 0; piece < 2; piece++) {
            int a, b;
            double sy;
            if (piece == 0) {
                if (klo >= 0) continue;
                a = cs[cbs[q] + klo + R.ncy]; b = cs[cbs[q] + R.ncy]; sy = -R.L;  // rows wrapped from the top
            } else {
                if (khi < R.ncy) continue;
                a = cs[cbs[q]]; b = cs[cbs[q] + khi - R.ncy + 1]; sy = R.L;       // rows wrapped from the bottom
            }
            for (int j = a + gl; j < b; j += kRdfG)
                if (j != i && pre(j)) {
                    const double2 qq = xy[j];
                    fn(j, (float)((p.x - qq.x) - sxq[q]), (float)((p.y - qq.y) - sy));
                }
        }
    }
}

// Smooth RDF histogram of the current state.  Per block (kRdfChunk beads):
// the (ta or tb)-type beads are listed in LDS and each is handled by a group
// of kRdfG lanes; contributions go to per-wavefront LDS fixed-point
// histograms (integer atomics: order independent), summed into the block's
// partial.
__global__ __launch_bounds__(kBeadBlock) void k_rdf(const RepInfo *__restrict__ ri, const int32_t *__restrict__ bblk,
                                                    const int32_t *__restrict__ cs, const double2 *__restrict__ xy,
                                                    const uint32_t *__restrict__ gt, RdfArgs A,
                                                    unsigned long long *__restrict__ part) {
    __shared__ unsigned long long h[kBeadBlock / 64][256];
    __shared__ int s_list[kBeadBlock];
    __shared__ int s_n;
    for (int b = threadIdx.x; b < (kBeadBlock / 64) * 256; b += kBeadBlock) h[b >> 8][b & 255] = 0ull;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // block b: beads [(b % kRdfSub) * kRdfChunk, +kRdfChunk) of bead block b / kRdfSub
    const int bb = blockIdx.x / kRdfSub;
    const int r = bblk[bb];
    const RepInfo &R = ri[r];
    const int loc = (bb - R.bblk0) * kBeadBlock + (blockIdx.x % kRdfSub) * kRdfChunk + threadIdx.x;
    if (threadIdx.x < kRdfChunk && loc < R.N) {
        const int i = R.off + loc;
        const int ti = (int)(gt[i] >> kGidBits);
        if (ti == A.ta || ti == A.tb) s_list[atomicAdd(&s_n, 1)] = i;
    }
    __syncthreads();
    unsigned long long *hw = h[threadIdx.x >> 6];
    const int gl = threadIdx.x & (kRdfG - 1);
    const float pref = 35.0f / (32.0f * A.width), iw = 1.0f / A.width;
    for (int e = threadIdx.x / kRdfG; e < s_n; e += kBeadBlock / kRdfG) {
        const int i = s_list[e];
        const uint32_t gti = gt[i], ti = gti >> kGidBits, gi = gti & kGidMask;
        visit_wide(
            R, cs, xy, i, xy[i], A.cut, A.Rc, gl,
            [&](int j) {
                const uint32_t gtj = gt[j], tj = gtj >> kGidBits;
                if (A.ta == A.tb) return (int)tj == A.ta && gi < (gtj & kGidMask);
                return (int)ti == A.ta && (int)tj == A.tb;
            },
            [&](int j, float dx, float dy) {
                const float r2 = dx * dx + dy * dy;
                if (r2 >= A.cut * A.cut || r2 == 0.0f) return;
                const float rr = sqrtf(r2);
                const int b0 = max(0, (int)floorf((rr - A.width) / A.dr));
                const int b1 = min(A.nbins - 1, (int)floorf((rr + A.width) / A.dr));
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
        part[(size_t)b * gridDim.x + blockIdx.x] = v;  // bin-major: the reduction reads coalesced
    }
}

// window accumulator += sum of the block partials of each replica: one block
// per (bin, replica), a strided sum then an LDS tree (integer: exact)
__global__ __launch_bounds__(256) void k_rdf_reduce(const RepInfo *__restrict__ ri, int nblk,
                                                    const unsigned long long *__restrict__ part,
                                                    unsigned long long *__restrict__ acc, int nbins) {
    __shared__ unsigned long long red[256];
    const int b = blockIdx.x, r = blockIdx.y;
    const RepInfo &R = ri[r];
    const int k0 = R.bblk0 * kRdfSub, nb = (R.N + kBeadBlock - 1) / kBeadBlock * kRdfSub;
    unsigned long long v = 0;
    for (int k = threadIdx.x; k < nb; k += 256) v += part[(size_t)b * nblk + k0 + k];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[(size_t)r * nbins + b] += red[0];
}

// X_i += d loss / d x_i of the RDF at the current state; gw[r][b] = dL/dg_b * C_b.
// Beads of either type are listed per block and handled by kRdfG lanes each;
// the group's partial sums are combined by a fixed shuffle tree.
__global__ __launch_bounds__(kBeadBlock) void k_rdf_adj(const RepInfo *__restrict__ ri,
                                                        const int32_t *__restrict__ bblk,
                                                        const int32_t *__restrict__ cs, const double2 *__restrict__ xy,
                                                        const uint32_t *__restrict__ gt, RdfArgs A,
                                                        const double *__restrict__ gw, double2 *__restrict__ X) {
    __shared__ float w_s[256];
    __shared__ int s_list[kBeadBlock];
    __shared__ int s_n;
    const int bb = blockIdx.x / kRdfSub;
    const int r = bblk[bb];
    for (int b = threadIdx.x; b < A.nbins; b += blockDim.x) w_s[b] = (float)gw[(size_t)r * A.nbins + b];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const RepInfo &R = ri[r];
    const int loc = (bb - R.bblk0) * kBeadBlock + (blockIdx.x % kRdfSub) * kRdfChunk + threadIdx.x;
    if (threadIdx.x < kRdfChunk && loc < R.N) {
        const int i = R.off + loc;
        const int ti = (int)(gt[i] >> kGidBits);
        if (ti == A.ta || ti == A.tb) s_list[atomicAdd(&s_n, 1)] = i;
    }
    __syncthreads();
    const int gl = threadIdx.x & (kRdfG - 1);
    const float pref = 35.0f / (32.0f * A.width), iw = 1.0f / A.width;
    // rounds over the list; every lane of a group takes part in the shuffles
    const int ng = kBeadBlock / kRdfG;
    for (int e0 = 0; e0 < s_n; e0 += ng) {
        const int e = e0 + (int)threadIdx.x / kRdfG;
        double ax = 0.0, ay = 0.0;
        int i = -1;
        if (e < s_n) {
            i = s_list[e];
            const uint32_t ti = gt[i] >> kGidBits;
            visit_wide(
                R, cs, xy, i, xy[i], A.cut, A.Rc, gl,
                [&](int j) {
                    const uint32_t tj = gt[j] >> kGidBits;
                    return ((int)ti == A.ta && (int)tj == A.tb) || ((int)ti == A.tb && (int)tj == A.ta);
                },
                [&](int j, float dx, float dy) {
                    const float r2 = dx * dx + dy * dy;
                    if (r2 >= A.cut * A.cut || r2 == 0.0f) return;
                    const float rr = sqrtf(r2);
                    const int b0 = max(0, (int)floorf((rr - A.width) / A.dr));
                    const int b1 = min(A.nbins - 1, (int)floorf((rr + A.width) / A.dr));
                    float ds = 0.0f;
                    for (int b = b0; b <= b1; b++) {
                        const float t = (rr - (b + 0.5f) * A.dr) * iw;
                        if (fabsf(t) >= 1.0f) continue;
                        const float q = 1.0f - t * t;
                        ds += w_s[b] * pref * 3.0f * q * q * (-2.0f * t * iw);
                    }
                    ax += (double)(ds * dx / rr);
                    ay += (double)(ds * dy / rr);
                });
        }
        for (int o = kRdfG / 2; o > 0; o >>= 1) {
            ax += __shfl_xor(ax, o, kRdfG);
            ay += __shfl_xor(ay, o, kRdfG);
        }
        if (gl == 0 && i >= 0) {
            X[i].x += ax;
            X[i].y += ay;
        }
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

// J^T products of the pair forces at state n for the force adjoint Fa:
// X += Jx^T Fa (FP64 sums), T = Jv^T Fa, per-gid theta gradient += (half of
// each pair: every pair is visited from both beads).
__global__ __launch_bounds__(kBeadBlock) void k_pair_adj(const RepInfo *__restrict__ ri,
                                                         const RepStep *__restrict__ rs,
                                                         const int32_t *__restrict__ bblk,
                                                         const int32_t *__restrict__ cs,
                                                         const double2 *__restrict__ xy, const float2 *__restrict__ vt,
                                                         const uint32_t *__restrict__ gt,
                                                         const double2 *__restrict__ Fa, double2 *__restrict__ X,
                                                         double2 *__restrict__ T, double *__restrict__ thb) {
    // block b: beads [(b % kRdfSub) * kRdfChunk, +kRdfChunk) of bead block
    // b / kRdfSub, kAdjG lanes per bead
    const int bb = blockIdx.x / kRdfSub;
    const int r = bblk[bb];
    const RepInfo &R = ri[r];
    const RepStep &S = rs[r];
    const int loc = (bb - R.bblk0) * kBeadBlock + (blockIdx.x % kRdfSub) * kRdfChunk + (int)threadIdx.x / kAdjG;
    const int gl = threadIdx.x & (kAdjG - 1);
    const bool live = loc < R.N;
    const int i = R.off + (live ? loc : 0);
    const double2 pi = xy[i];
    const float2 vi = vt[i];
    const uint32_t gti = gt[i], ti = gti >> kGidBits, gi = gti & kGidMask;
    const double2 Fi = Fa[i];
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
    // sums over the group (fixed shuffle tree), written by its first lane
    for (int o = kAdjG / 2; o > 0; o >>= 1) {
        xa += __shfl_xor(xa, o, kAdjG);
        ya += __shfl_xor(ya, o, kAdjG);
        ta += __shfl_xor(ta, o, kAdjG);
        tb += __shfl_xor(tb, o, kAdjG);
#pragma unroll
        for (int q = 0; q < 7; q++) th[q] += __shfl_xor(th[q], o, kAdjG);
    }
    if (!live || gl != 0) return;
    X[i].x += xa;
    X[i].y += ya;
    T[i] = make_double2(ta, tb);
    double *o = thb + (size_t)(R.off + gi) * 8;
    for (int q = 0; q < 7; q++) o[q] += th[q];
}

// bond terms of the force adjoint (FP64): lipid chains from their head, a
// chemistry bond from its smaller-gid bead
__global__ __launch_bounds__(kBeadBlock) void k_bond_adj(const RepInfo *__restrict__ ri,
                                                         const int32_t *__restrict__ bblk,
                                                         const double2 *__restrict__ xy,
                                                         const uint32_t *__restrict__ gt,
                                                         const int8_t *__restrict__ mp, const int32_t *__restrict__ lnk,
                                                         const double2 *__restrict__ Fa, double2 *__restrict__ X) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    auto adj = [&](int i, int j) {
        const double2 p = xy[i], q = xy[j];
        const double dx = minimg(q.x - p.x, R.L, R.halfL), dy = minimg(q.y - p.y, R.L, R.halfL);
        const double rr = sqrt(dx * dx + dy * dy);
        const double g = -R.kb * (rr - R.r0) / rr, gp = -R.kb * R.r0 / (rr * rr);
        const double ex = dx / rr, ey = dy / rr;
        const double lx = Fa[j].x - Fa[i].x, ly = Fa[j].y - Fa[i].y;
        const double el = ex * lx + ey * ly;
        const double bx = g * lx + gp * rr * el * ex, by = g * ly + gp * rr * el * ey;
        X[j].x += bx;
        X[j].y += by;
        X[i].x -= bx;
        X[i].y -= by;
    };
    const int m = mp[s];
    if (m == 0) {
        int prev = s, n = 1;
        for (int d = lnk[s]; d >= 0 && n < R.mlen; d = lnk[d], n++) {
            adj(prev, d);
            prev = d;
        }
    } else if (m < 0) {
        const int p = lnk[s];
        if (p >= 0 && (gt[s] & kGidMask) < (gt[p] & kGidMask)) adj(s, p);
    }
}

// one reverse step: new adjoints in the order of step n, scattered to the
// order of step n-1 through the sort map src (new slot d <- old slot src[d])
__global__ __launch_bounds__(kBeadBlock) void k_adj_step(const RepInfo *__restrict__ ri,
                                                         const int32_t *__restrict__ bblk,
                                                         const int32_t *__restrict__ src,
                                                         const double2 *__restrict__ X, const double2 *__restrict__ V,
                                                         const double2 *__restrict__ F, const double2 *__restrict__ T,
                                                         double2 *__restrict__ Xo, double2 *__restrict__ Vo,
                                                         double2 *__restrict__ Fo) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int d = R.off + loc;
    const int o = src[d];
    const double2 x = X[d], v = V[d], f = F[d], t = T[d];
    Xo[o] = x;
    Vo[o] = make_double2(R.dt * x.x + t.x + v.x, R.dt * x.y + t.y + v.y);
    Fo[o] = make_double2(R.dt2 * x.x + R.cvt * t.x + R.dt * v.x, R.dt2 * x.y + R.cvt * t.y + R.dt * v.y);
}

// per-replica sum of the per-gid theta gradients (fixed partition, FP64)
__global__ __launch_bounds__(256) void k_theta_reduce(const RepInfo *__restrict__ ri, const double *__restrict__ thb,
                                                      double *__restrict__ out) {
    __shared__ double red[256];
    const RepInfo &R = ri[blockIdx.x];
    for (int q = 0; q < 7; q++) {
        double s = 0.0;
        for (int g = threadIdx.x; g < R.N; g += 256) s += thb[(size_t)(R.off + g) * 8 + q];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * 8 + q] = red[0];
        __syncthreads();
    }
}

__global__ void k_zero_f64(double *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace oc
