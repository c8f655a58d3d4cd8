// oc_kernels.hip -- gfx950 kernels of the 2D DPD step.
//
// Step order (one launch sequence per step for all replicas at once):
//   k_step_params  per-replica step counter, noise key, warm-up ramp, sigma(t)
//   k_keys         integrated position -> new cell key, block maxima of the
//                  cell displacement
//   k_tsort        per sort tile: the beads of each new cell, collected from
//                  the old cells within the radius in old-slot order (the
//                  stable order), numbered within the tile
//   k_tscan        tile offsets (+ the step's per-replica values)
//   k_place        new cell starts, and the gather of the state into the new
//                  order, integrating on the way (the integrate of
//                  cel/engine/oercel.c:2303-2334 fused into the scatter of
//                  sort_cells, :897-921)
//   k_chem_*       precursor chemistry (own definition)
//   k_pair         full-shell DPD pair forces, LDS-staged tiles, with the
//                  bonded terms of each bead (gather form)
#pragma once
#include "oc_common.hpp"

namespace oc {

// --------------------------------------------------------------------------- helpers

__device__ __forceinline__ int cell_x(const RepInfo &R, double x) {
    int c = (int)(x * R.icsx);
    return c >= R.ncx ? R.ncx - 1 : c;
}
__device__ __forceinline__ int cell_y(const RepInfo &R, double y) {
    int c = (int)(y * R.icsy);
    return c >= R.ncy ? R.ncy - 1 : c;
}

// Groot-Warren update of one bead, cel/engine/oercel.c:2319-2332 (FP64,
// unfused: identical bits to the CPU restatement)
__device__ __forceinline__ void integrate1(const RepInfo &R, double2 p, double2 v, double2 f, double2 &po, double2 &vo,
                                           float2 &vt) {
    const double ddx = R.dt * v.x + R.dt2 * f.x, ddy = R.dt * v.y + R.dt2 * f.y;
    double xi = p.x + ddx, yi = p.y + ddy;
    xi = xi < 0.0 ? xi + R.L : xi;
    yi = yi < 0.0 ? yi + R.L : yi;
    xi = xi >= R.L ? xi - R.L : xi;
    yi = yi >= R.L ? yi - R.L : yi;
    po = make_double2(xi, yi);
    vt = make_float2((float)(v.x + R.cvt * f.x), (float)(v.y + R.cvt * f.y));
    vo = make_double2(v.x + R.dt * f.x, v.y + R.dt * f.y);
}

__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}

// append an event (a dropped one when the buffer is full raises ERR_EVOVF)
__device__ __forceinline__ void push_event(Event *ev, unsigned long long *evn, int64_t cap, uint32_t *err, int64_t step,
                                           int kind, uint32_t a, uint32_t b, int rep, double x = 0.0,
                                           double y = 0.0) {
    const unsigned long long k = atomicAdd(evn, 1ull);
    if ((int64_t)k < cap) {
        Event e;
        e.step = step; e.kind = kind; e.gid_a = a; e.gid_b = b; e.rep = rep; e.x = x; e.y = y;
        ev[k] = e;
    } else {
        atomicOr(err, ERR_EVOVF);
    }
}

// --------------------------------------------------------------------------- per-step parameters

// Per-step values of replica r, computed by one wavefront (lane k: a_eff[k]).
// step = base[r] + delta (per-replica base), or set_step, or the current + delta.
__device__ __forceinline__ void step_params_wave(const RepInfo &R, RepStep &S, int r, int k, int64_t delta,
                                                 int64_t set_step, const int64_t *base) {
    const int64_t st = base ? base[r] + delta : (set_step >= 0 ? set_step : S.step + delta);
    // a_eff, cel/engine/oercel.c:1122-1129: one matrix entry per lane
    const double a = R.a[k];
    double ae = a;
    if (!(st >= R.warm_end || R.warmup <= 0 || a <= 25.0)) {
        const double fr = 1.0 - (double)(R.warm_end - st) / (double)R.warmup;
        ae = 25.0 + (a - 25.0) * (fr < 0 ? 0 : fr);
    }
    S.a_eff[k] = (float)ae;
    // the step's keys, one per lane: lane 1 the pair-noise key, lanes 2-13
    // the keys sseed + 1 .. kcap of RepStep (consecutive fields)
    if (k >= 1 && k <= 13) {
        uint64_t salt = 0;
        switch (k) {
        case 2: salt = SALT_CONV; break;
        case 3: salt = SALT_FORM; break;
        case 4: salt = SALT_BREAK; break;
        case 5: salt = 11; break;  // decay and feed keys, cel/engine/oercel.c:1664-1665, 1724
        case 6: salt = 13; break;
        case 7: salt = 14; break;
        case 8: salt = 12; break;  // catalyst decay and reactions, feed cap: cel/engine_gpu/oercel_gpu.cu:1765-1772
        case 9: salt = 21; break;
        case 10: salt = 22; break;
        case 11: salt = 23; break;
        case 12: salt = 24; break;
        default: salt = 25; break;
        }
        (&S.sseed)[k - 1] = k == 1 ? step_seed(R.seed, st) : hash_key(R.seed, st, salt);
    }
    if (k != 0) return;
    S.step = st;
    const double t = R.t_anchor + (double)(st - R.step_anchor) * R.dt;
    S.t = t;
    // kT_at, cel/engine/oercel.c:622-624; sdt as in pair_forces, :1204
    const double kTt = R.kT_amp == 0.0 ? R.kT : R.kT + R.kT_amp * sin(2.0 * M_PI * t / R.kT_period);
    S.sdt = (float)(sqrt(2.0 * R.gamma * kTt) / sqrt(R.dt));
    S.gamma = (float)R.gamma;
    S.chem_now = (R.chem && R.n_chem > 0 && (st % R.n_chem) == 0) ? 1 : 0;
    S.rchem_now = (R.rchem && R.n_chem > 0 && (st % R.n_chem) == 0) ? 1 : 0;
    S.disp = 0;
}

__global__ __launch_bounds__(64) void k_step_params(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                    int64_t delta, int64_t set_step,
                                                    const int64_t *__restrict__ base) {
    const int r = blockIdx.x;
    step_params_wave(ri[r], rs[r], r, threadIdx.x, delta, set_step, base);
}

// --------------------------------------------------------------------------- sort

// Sort tile k of a column split into ntc tiles: rows [k ncy / ntc, (k + 1)
// ncy / ntc) hold row.
__device__ __forceinline__ int sort_tile_row(const RepInfo &R, int row) {
    int k = (int)(((int64_t)row * R.ntc) / R.ncy);
    while (k + 1 < R.ntc && ((int64_t)(k + 1) * R.ncy) / R.ntc <= row) k++;
    while (k > 0 && ((int64_t)k * R.ncy) / R.ntc > row) k--;
    return k;
}

// New cell key of every bead (of its integrated position when INTEG; bit 31
// set when the bead changed cells).  A bead that moved by more than one cell
// (a rare event) sets *moved and raises its replica's largest periodic cell
// displacement rmax[r] (x, y); the sort then widens its window to it.  The
// force kernel computes the same keys in its epilogue (new_key), so a step
// whose forces came from it skips this launch.
// key of bead s at q that sits in (local) cell (ox, oy) of the last sort.
// moved[0]: some bead moved by more than one cell; moved[1 + tile]: such a
// bead arrives in that sort tile (which then collects its beads over the
// replica's largest displacement)
__device__ __forceinline__ void new_key_from(const RepInfo &R, int r, int s, int ox, int oy, double2 q, int32_t *key,
                                             int2 *rmax, uint32_t *moved) {
    const int cx = cell_x(R, q.x), cy = cell_y(R, q.y);
    // bit 31: the bead changed cells
    key[s] = (R.coff + cx * R.ncy + cy) | (cx != ox || cy != oy ? (int32_t)0x80000000 : 0);
    int dx = abs(cx - ox), dy = abs(cy - oy);
    dx = min(dx, R.ncx - dx);
    dy = min(dy, R.ncy - dy);
    if (dx > 1 || dy > 1) {
        atomicMax(&rmax[r].x, dx);
        atomicMax(&rmax[r].y, dy);
        atomicOr(moved, 1u);
        atomicOr(&moved[1 + R.stile0 + cx * R.ntc + sort_tile_row(R, cy)], 1u);
    }
}
__device__ __forceinline__ void new_key(const RepInfo &R, int r, int s, double2 p, double2 q, int32_t *key,
                                        int2 *rmax, uint32_t *moved) {
    new_key_from(R, r, s, cell_x(R, p.x), cell_y(R, p.y), q, key, rmax, moved);
}

template <bool INTEG>
__global__ __launch_bounds__(kBeadBlock) void k_keys(const RepInfo *__restrict__ ri, const int32_t *__restrict__ bblk,
                                                     const double2 *__restrict__ xy, const double2 *__restrict__ v,
                                                     const double2 *__restrict__ f, int32_t *__restrict__ key,
                                                     int2 *__restrict__ rmax, uint32_t *__restrict__ moved) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const double2 p = xy[s];
    double2 q = p;
    if (INTEG) {
        double2 vo;
        float2 vt;
        integrate1(R, p, v[s], f[s], q, vo, vt);
    }
    new_key(R, r, s, p, q, key, rmax, moved);
}

// Chemistry role of a bead at a check (own definition): bit 0 -- the bead
// breaks its chemistry bond (the bond's smaller gid draws hash_u01(kbreak) <
// p_break); bit 1 -- it searches its neighbourhood (a P whose draw
// hash_u01(kconv) < p_conv looks for a lipid tail, a free T -- type T, in no
// lipid, unbonded -- for a partner).  wp: the gid word of the chemistry
// partner (read only for a bonded bead outside lipids).
constexpr int kRoleShift = 30;
// Caller-supplied draws of a check (oc_chem_uniforms, one replica): the
// conversion and break uniforms per gid, the formation uniform per pair (keys
// lo << 32 | hi, strictly increasing; a pair missing from it raises ERR_INJ
// and does not bond).
struct ChemInj {
    const double *uconv, *ubreak;
    const uint64_t *fk;
    const double *fv;
    int64_t nf;
    uint32_t *err;
};
__device__ __forceinline__ int chem_role(const RepInfo &R, const RepStep &S, uint32_t w, int m, int l, uint32_t wp,
                                         const ChemInj *inj = nullptr) {
    const uint32_t g = w & kGidMask, ty = w >> kGidBits;
    const bool brk = m < 0 && l >= 0 && g < (wp & kGidMask) &&
                     (inj ? inj->ubreak[g] : hash_u01(S.kbreak, g)) < R.p_break;
    const bool conv = ty == OC_N && (inj ? inj->uconv[g] : hash_u01(S.kconv, g)) < R.p_conv;
    const bool form = ty == OC_T && m < 0 && l < 0;
    return (brk ? 1 : 0) | (conv || form ? 2 : 0);
}

// Class byte of every bead at a check (written with the state of the check's
// start: by the placement, or by k_chem_list for a loaded state): a lipid tail
// (type T in a lipid) or a free T (type T, in no lipid, unbonded) -- the only
// candidates a search reads further.
enum : uint8_t { CC_LIPT = 1, CC_FREET = 2 };
__device__ __forceinline__ uint8_t chem_class(uint32_t w, int m, int l) {
    if ((w >> kGidBits) != OC_T) return 0;
    return m >= 0 ? CC_LIPT : (l < 0 ? CC_FREET : 0);
}
// A free T's formation proposal to slot b is held in its link word as
// -2 - b until the force pass of the step (or k_chem_resolve) settles it: the
// bond forms when the proposal is mutual.
__device__ __forceinline__ int32_t chem_prop_code(int b) { return -2 - b; }
// The partner of free T s whose link word l holds a proposal (l < -1): b =
// -2 - l when b's link word proposes s, or already holds s (b settled
// first); else -1 (no bond).
__device__ __forceinline__ int32_t chem_settle(int s, int32_t l, const int32_t *lnk) {
    const int b = -2 - l;
    const int32_t lb = lnk[b];
    return lb == chem_prop_code(s) || lb == s ? b : -1;
}

// The check's list is kept in segments: a producer block (a placement tile, or
// a bead block of k_chem_list) numbers the entries of its beads with a role
// (an LDS counter, in any order: every decision is per bead) and writes entry
// e to the segment's inline record when e < kChemInl (the decision loads it
// together with the header) or to spill[base + e], base = the block's first
// slot; the header seg = (base, cnt, replica).  No global atomics.
constexpr int kChemInl = 16;
struct ChemSeg {
    int32_t base, cnt, rep, pad;
};
// One listed bead: slot | role << kRoleShift, gid word and position.
struct ChemEnt {
    int32_t sr;
    uint32_t w;
    double x, y;
};
struct ChemList {
    ChemSeg *seg;
    ChemEnt *inl;    // kChemInl entries per segment
    ChemEnt *spill;  // indexed by slot
};
__device__ __forceinline__ ChemEnt *chem_ent(const ChemList &C, int k, int base, int e) {
    return e < kChemInl ? C.inl + (size_t)k * kChemInl + e : C.spill + base + e;
}

// The search window of a point p (rw <= 1 + 1e-6 and ncy >= 8 rows: at most
// one seam) as up to six contiguous slot ranges: per column x-1, x, x+1 the
// rows of [y - rw, y + rw], split at the periodic seam; csf(col, row) is the
// first slot of cell (col, row) for row in [0, ncy] (row ncy: the end of the
// column).  Every cell-start load is issued before the first is used.
struct Win {
    int first[6], cnt[6], total;
};
template <class CS>
__device__ __forceinline__ Win chem_window(const RepInfo &R, double2 p, double rw, CS csf) {
    const int cx = cell_x(R, p.x);
    const int khi = (int)((p.y + rw) * R.icsy);
    const int klo = (int)((p.y - rw) * R.icsy + R.ncy) - R.ncy;
    int a0, b0, a1 = 0, b1 = -1;
    if (klo < 0) { a0 = klo + R.ncy; b0 = R.ncy - 1; a1 = 0; b1 = khi; }
    else if (khi >= R.ncy) { a0 = klo; b0 = R.ncy - 1; a1 = 0; b1 = khi - R.ncy; }
    else { a0 = klo; b0 = khi; }
    // all twelve loads unconditional (without a second piece a1 = 0, b1 = -1:
    // row 0, in bounds), so that they are in flight together
    int v[12];
    const bool two = b1 >= a1;
#pragma unroll
    for (int dc = 0; dc < 3; dc++) {
        int col = cx + dc - 1;
        col = col < 0 ? col + R.ncx : (col >= R.ncx ? col - R.ncx : col);
        v[4 * dc] = csf(col, a0);
        v[4 * dc + 1] = csf(col, b0 + 1);
        v[4 * dc + 2] = csf(col, a1);
        v[4 * dc + 3] = csf(col, b1 + 1);
    }
#pragma unroll
    for (int dc = 0; dc < 3; dc++) {
        v[4 * dc + 2] = two ? v[4 * dc + 2] : 0;
        v[4 * dc + 3] = two ? v[4 * dc + 3] : 0;
    }
    Win W;
    W.total = 0;
#pragma unroll
    for (int dc = 0; dc < 3; dc++) {
        W.first[2 * dc] = v[4 * dc];
        W.cnt[2 * dc] = v[4 * dc + 1] - v[4 * dc];
        W.first[2 * dc + 1] = v[4 * dc + 2];
        W.cnt[2 * dc + 1] = v[4 * dc + 3] - v[4 * dc + 2];
        W.total += W.cnt[2 * dc] + W.cnt[2 * dc + 1];
    }
    return W;
}

// Piece list of the periodic window [c - R, c + R] of 0..n-1 in ascending order.
__device__ __forceinline__ int win_pieces(int c, int R, int n, int *a, int *b) {
    if (2 * R + 1 >= n) { a[0] = 0; b[0] = n - 1; return 1; }
    if (c - R < 0) { a[0] = 0; b[0] = c + R; a[1] = c - R + n; b[1] = n - 1; return 2; }
    if (c + R >= n) { a[0] = 0; b[0] = c + R - n; a[1] = c - R; b[1] = n - 1; return 2; }
    a[0] = c - R; b[0] = c + R;
    return 1;
}

// Sort tiles: a column's rows [k0, k1) of new cells (up to kSortRows rows,
// kSortThr threads, two rows each).  The stable counting sort of
// cel/engine/oercel.c:828-933 orders the beads of a new cell by old slot.
// With a displacement radius of one cell (the normal case) the tile stages
// the keys of the old rows k0-1 .. k1 of its three columns in LDS in old-slot
// order (columns ascending; within a column the row pieces of the periodic
// range ascending), so LDS order = old-slot order.  The beads of new cell c in
// the stable order are the arrivals from lower cells, the stayers of c, the
// arrivals from higher cells: a stayer's rank is (arrivals from lower cells)
// + (its index in old cell c) - (departures from c before it); a bead that
// changed cells counts the beads of c before it in c's window (one lane per
// such bead).  Counts and arrival counts are LDS integer atomics (order
// independent).  A larger radius, a box narrower than three columns or an
// overflowing stage takes the general path: each row visits the old cells
// within the radius in increasing cell index (= old slot) from global memory
// and keeps the beads whose key is the cell.  k_tscan turns the tile counts
// into slot offsets; k_place writes the new cell starts and gathers the state
// into the new order.  No global atomics, no histogram to clear.
constexpr int kSortRows = 256;                // rows per sort tile
constexpr int kSortThr = 128;                 // threads per sort tile (two rows each)
constexpr int kSortCand = 768;                // staged candidates (~580 at rho 3, nsub 4; more: the general path)
constexpr int kSortMv = 64;                   // listed beads that changed cells (~10; more: ranked in place)
constexpr int kSortRowTab = kSortRows + 3;    // staged rows per column (+ end)
constexpr int kSortCapLog = 10;
constexpr int kSortCap = 1 << kSortCapLog;    // beads per sort tile
constexpr int kSortGrid = 256 * 12;           // workgroups walking the sort tiles

// Tile walk of a grid over n tiles that keeps neighbouring tiles on one XCD.
// Workgroups are dispatched to the eight XCDs round robin (workgroup b runs
// on XCD b % 8), and each XCD has its own L2: XCD x takes the contiguous band
// [x per, (x + 1) per) of tiles, its gridDim / 8 workgroups walking the band
// with that stride, so the columns that share staged keys, state lines and
// the (tile, rank) entries of lipid partners are read through one L2.  A grid
// that is not a multiple of 8 walks the tiles with stride gridDim.
struct TileWalk {
    int first, end, stride;
};
__device__ __forceinline__ TileWalk tile_walk(int n) {
    TileWalk w;
    if ((gridDim.x & 7) == 0) {
        const int per = (n + 7) >> 3, x = (int)(blockIdx.x & 7);
        w.first = x * per + (int)(blockIdx.x >> 3);
        w.end = min((x + 1) * per, n);
        w.stride = (int)(gridDim.x >> 3);
    } else {
        w.first = (int)blockIdx.x;
        w.end = n;
        w.stride = (int)gridDim.x;
    }
    return w;
}

__device__ __forceinline__ int block_excl_scan128(int v, int *tmp, int &total) {
    // exclusive scan over a 128-thread block; tmp: 2 ints of LDS
    const int inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) tmp[w] = inc;
    __syncthreads();
    const int off = w ? tmp[0] : 0;
    total = tmp[0] + tmp[1];
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(kSortThr) void k_tsort(const RepInfo *__restrict__ ri, const Tile *__restrict__ tiles,
                                                    const int32_t *__restrict__ cs_old,
                                                    const int32_t *__restrict__ key,
                                                    const int2 *__restrict__ rmax,
                                                    uint32_t *__restrict__ moved,
                                                    int32_t *__restrict__ tcount, int32_t *__restrict__ locst,
                                                    int32_t *__restrict__ tsrc, int32_t *__restrict__ tslot,
                                                    uint32_t *__restrict__ err, int nt) {
    __shared__ int32_t s_key[kSortCand];
    __shared__ int32_t s_slot[kSortCand];     // old slot of staged candidate j
    __shared__ uint16_t s_cst[kSortRows];     // LDS index of the first staged bead of tile row r (old cell)
    __shared__ int32_t s_cs[3][kSortRowTab];  // cs_old of the staged rows (slot order); extras: piece ends
    __shared__ int32_t s_e[3][2];
    __shared__ int32_t s_pc[4][3];  // per staged column: LDS base, piece counts; [3][0]: total
    __shared__ int32_t s_cnt[kSortRows], s_alow[kSortRows];
    __shared__ uint16_t s_start[kSortRows];  // < kSortCap
    __shared__ int32_t s_mv[kSortMv];  // staged ids of the tile's beads that changed cells
    __shared__ int s_nmv;
    __shared__ int s_tmp[2];

    auto do_tile = [&](const int tix) {
        const Tile T = tiles[tix];
        const RepInfo &R = ri[T.rep];
        const int ncy = R.ncy, ncx = R.ncx;
        const int t = threadIdx.x;
        const int nrows = T.k1 - T.k0;
        const int c0 = R.coff + T.cx * ncy + T.k0;
        int Rx = 1, Ry = 1;
        if (*moved && moved[1 + tix]) {
            // a bead arrives from further than one cell: the replica's
            // largest displacement
            const int2 m = rmax[T.rep];
            Rx = max(1, m.x);
            Ry = max(1, m.y);
        }
        // staged rows k0-1 .. k1 (mod ncy) as row pieces P0 = [a0, b0], P1 =
        // [a1, b1] in increasing row (= slot) order; the three columns ascending
        int a0, b0, a1 = 0, b1 = -1;
        if (nrows + 2 >= ncy) { a0 = 0; b0 = ncy - 1; }
        else if (T.k0 == 0) { a0 = 0; b0 = T.k1; a1 = ncy - 1; b1 = ncy - 1; }
        else if (T.k1 == ncy) { a0 = 0; b0 = 0; a1 = T.k0 - 1; b1 = ncy - 1; }
        else { a0 = T.k0 - 1; b0 = T.k1; }
        const int len0 = b0 - a0 + 1, len1 = b1 - a1 + 1, nr = len0 + len1;
        int col[3];
        {
            const int cm = T.cx == 0 ? ncx - 1 : T.cx - 1, cp = T.cx == ncx - 1 ? 0 : T.cx + 1;
            // ascending order of {cm, cx, cp}
            col[0] = min(cm, min(T.cx, cp));
            col[2] = max(cm, max(T.cx, cp));
            col[1] = cm + T.cx + cp - col[0] - col[2];
        }
        const int qo = col[0] == T.cx ? 0 : (col[1] == T.cx ? 1 : 2);
        bool fast = Rx == 1 && Ry == 1 && ncx >= 3 && nr + 1 <= kSortRowTab;
        auto row_of = [&](int u) { return u < len0 ? a0 + u : a1 + (u - len0); };
        if (fast) {
            constexpr int kPer = (kSortRowTab + kSortThr - 1) / kSortThr;
            int v[3][kPer];
            // every load unconditional (the row index clamped into the staged
            // rows), so that all of them are in flight together
    #pragma unroll
            for (int q = 0; q < 3; q++) {
                const int cbq = R.coff + col[q] * ncy;
    #pragma unroll
                for (int i = 0; i < kPer; i++) v[q][i] = cs_old[cbq + row_of(min(t + i * kSortThr, nr - 1))];
            }
            int ve = 0;
            if (t < 6) {
                const int q = t >> 1;
                const int b = (t & 1) ? (len1 > 0 ? b1 : b0) : b0;
                ve = cs_old[R.coff + col[q] * ncy + b + 1];
            }
    #pragma unroll
            for (int q = 0; q < 3; q++)
    #pragma unroll
                for (int i = 0; i < kPer; i++) {
                    const int u = t + i * kSortThr;
                    if (u < nr) s_cs[q][u] = v[q][i];
                }
            if (t < 6) s_e[t >> 1][t & 1] = ve;
        }
        s_cnt[t] = 0;
        s_cnt[t + kSortThr] = 0;
        s_alow[t] = 0;
        s_alow[t + kSortThr] = 0;
        if (t == 0) s_nmv = 0;
        __syncthreads();
        // per column (LDS, indexed by run-time q): staged count of each piece, LDS base
        if (t == 0) {
            int b = 0;
            for (int q = 0; q < 3; q++) {
                const int m0 = fast ? s_e[q][0] - s_cs[q][0] : 0;
                const int m1 = fast && len1 > 0 ? s_e[q][1] - s_cs[q][len0] : 0;
                s_pc[q][0] = b;
                s_pc[q][1] = m0;
                s_pc[q][2] = m1;
                b += m0 + m1;
            }
            s_pc[3][0] = b;
        }
        __syncthreads();
        const int nst = s_pc[3][0];
        fast = fast && nst <= kSortCand;
        auto lds_row = [&](int q, int u) {  // LDS index of the first bead of staged row u < nr
            return s_pc[q][0] + (u < len0 ? s_cs[q][u] - s_cs[q][0] : s_pc[q][1] + s_cs[q][u] - s_cs[q][len0]);
        };
        auto lds_after = [&](int q, int u) {  // just past the beads of staged row u
            return u + 1 == len0 ? s_pc[q][0] + s_pc[q][1]
                                 : (u + 1 == nr ? s_pc[q][0] + s_pc[q][1] + s_pc[q][2] : lds_row(q, u + 1));
        };
        auto u_of = [&](int row) { return row >= a0 && row <= b0 ? row - a0 : len0 + (row - a1); };
        const int ra = 2 * t, rb = 2 * t + 1;  // this thread's rows (local)
        auto slot_of = [&](int j) {  // old slot of staged candidate j
            const int q = j < s_pc[1][0] ? 0 : (j < s_pc[2][0] ? 1 : 2);
            const int o = j - s_pc[q][0];
            return o < s_pc[q][1] ? s_cs[q][0] + o : s_cs[q][len0] + (o - s_pc[q][1]);
        };
        if (fast) {
            // keys in LDS order; slot of LDS index j from the six pieces
            constexpr int kPer = kSortCand / kSortThr;
            const int nper = (nst + kSortThr - 1) / kSortThr;
            int v[kPer], sl[kPer];
    #pragma unroll
            for (int i = 0; i < kPer; i++)
                if (i < nper) {
                    const int j = t + i * kSortThr;
                    if (j < nst) {
                        sl[i] = slot_of(j);
                        v[i] = key[sl[i]];
                    }
                }
            if (ra < nrows) s_cst[ra] = (uint16_t)lds_row(qo, u_of(T.k0 + ra));
            if (rb < nrows) s_cst[rb] = (uint16_t)lds_row(qo, u_of(T.k0 + rb));
    #pragma unroll
            for (int i = 0; i < kPer; i++)
                if (i < nper) {
                    const int j = t + i * kSortThr;
                    if (j < nst) {
                        s_key[j] = v[i];
                        s_slot[j] = sl[i];
                    }
                }
        }
        __syncthreads();

        const int tb = tix << kSortCapLog;
        auto put = [&](int p, int sl) {  // the tile's p-th bead is old slot sl
            tsrc[tb + p] = sl;
            tslot[sl] = tb + p;
        };
        // rank of a bead that changed cells: the beads of its new cell before it
        // in the cell's window (old-slot order)
        auto mover_rank = [&](int j) {
            const int kc = s_key[j] & 0x7fffffff;
            int wa[2], wb[2];
            const int np = win_pieces(T.k0 + (kc - c0), 1, ncy, wa, wb);
            int rank = 0;
#pragma unroll
            for (int q = 0; q < 3; q++)
#pragma unroll
                for (int pc = 0; pc < 2; pc++) {
                    if (pc >= np) break;
                    const int j0 = lds_row(q, u_of(wa[pc]));
                    const int j1 = min(lds_after(q, u_of(wb[pc])), j);
                    for (int jj = j0; jj < j1; jj++) rank += (s_key[jj] & 0x7fffffff) == kc ? 1 : 0;
                }
            return rank;
        };
        int total;
        if (fast) {
            auto cell_start = [&](int r) { return (int)s_cst[r]; };
            for (int j = t; j < nst; j += kSortThr) {
                const int k = s_key[j];
                const int r = (k & 0x7fffffff) - c0;
                if ((unsigned)r < (unsigned)nrows) {
                    atomicAdd(&s_cnt[r], 1);
                    if (k < 0 && j < cell_start(r)) atomicAdd(&s_alow[r], 1);
                }
            }
            __syncthreads();
            const int ca = ra < nrows ? s_cnt[ra] : 0, cb = rb < nrows ? s_cnt[rb] : 0;
            const int start = block_excl_scan128(ca + cb, s_tmp, total);
            if (total > kSortCap) {
                // reported as an error; the tile places no bead, so every later
                // index stays in bounds
                if (ra < nrows) locst[c0 + ra] = 0;
                if (rb < nrows) locst[c0 + rb] = 0;
                if (t == 0) {
                    tcount[tix] = 0;
                    atomicOr(err, ERR_SORT);
                }
                return;
            }
            s_start[ra] = start;
            s_start[rb] = start + ca;
            if (ra < nrows) locst[c0 + ra] = start;
            if (rb < nrows) locst[c0 + rb] = start + ca;
            __syncthreads();
            for (int j = t; j < nst; j += kSortThr) {
                const int k = s_key[j];
                const int r = (k & 0x7fffffff) - c0;
                if ((unsigned)r >= (unsigned)nrows) continue;
                if (k < 0) {  // changed cells: ranked below, one lane per bead
                    const int m = atomicAdd(&s_nmv, 1);
                    if (m < kSortMv) s_mv[m] = j;
                    else put(s_start[r] + mover_rank(j), s_slot[j]);  // list full: rank in place
                    continue;
                }
                const int o = cell_start(r);
                int d = 0;
                for (int jj = o; jj < j; jj++) d += s_key[jj] < 0 ? 1 : 0;
                put(s_start[r] + s_alow[r] + (j - o) - d, s_slot[j]);
            }
            __syncthreads();
            for (int m = t; m < min(s_nmv, kSortMv); m += kSortThr) {
                const int j = s_mv[m];
                put(s_start[(s_key[j] & 0x7fffffff) - c0] + mover_rank(j), s_slot[j]);
            }
        } else {
            // general path: old cells within (Rx, Ry) of each row's cell, from global memory
            auto visit = [&](int row, auto fn) {
                const int c = c0 + (row - T.k0);
                int ra2[2], rb2[2], ca2[2], cb2[2];
                const int npr = win_pieces(row, Ry, ncy, ra2, rb2);
                const int npc = win_pieces(T.cx, Rx, ncx, ca2, cb2);
    #pragma unroll
                for (int pc = 0; pc < 2; pc++) {
                    if (pc >= npc) break;
                    for (int cc = ca2[pc]; cc <= cb2[pc]; cc++) {
                        const int cbq = R.coff + cc * ncy;
    #pragma unroll
                        for (int pr = 0; pr < 2; pr++) {
                            if (pr >= npr) break;
                            const int j1 = cs_old[cbq + rb2[pr] + 1];
                            for (int j = cs_old[cbq + ra2[pr]]; j < j1; j++)
                                if ((key[j] & 0x7fffffff) == c) fn(j);
                        }
                    }
                }
            };
            int ca = 0, cb = 0;
            if (ra < nrows) visit(T.k0 + ra, [&](int) { ca++; });
            if (rb < nrows) visit(T.k0 + rb, [&](int) { cb++; });
            const int start = block_excl_scan128(ca + cb, s_tmp, total);
            if (total > kSortCap) {
                if (ra < nrows) locst[c0 + ra] = 0;
                if (rb < nrows) locst[c0 + rb] = 0;
                if (t == 0) {
                    tcount[tix] = 0;
                    atomicOr(err, ERR_SORT);
                }
                return;
            }
            int pos = start;
            if (ra < nrows) {
                locst[c0 + ra] = start;
                visit(T.k0 + ra, [&](int j) { put(pos++, j); });
            }
            if (rb < nrows) {
                locst[c0 + rb] = pos;
                visit(T.k0 + rb, [&](int j) { put(pos++, j); });
            }
        }
        if (t == 0) tcount[tix] = total;
    };
    // a grid of at most a few blocks per CU walks the tiles (fewer, longer-lived
    // workgroups)
    const TileWalk W = tile_walk(nt);
    for (int tix = W.first; tix < W.end; tix += W.stride) {
        do_tile(tix);
        __syncthreads();
    }
}

// Exclusive scan of the tile counts -> first new slot of each tile (tiles are
// in cell order, cells in slot order), 1024 * per counts per block (per <= 8
// consecutive counts a thread, loaded together).  With more than one chunk,
// part holds the chunk totals (k_tsum) and each block adds those of the
// chunks before it.
// Block 0 also clears the displacement flag and maxima; when step_delta > 0
// one more block (blockIdx.x == number of chunks) sets the next step's
// per-replica values (the step counter advance of the step sequence) at the
// same time as the scan runs.
constexpr int kTScanChunk = 1024 * 8;  // largest chunk
__global__ __launch_bounds__(1024) void k_tsum(const int32_t *__restrict__ tcount, int nt, int per,
                                               int32_t *__restrict__ part) {
    __shared__ int wsum[16];
    const int b0 = (blockIdx.x * 1024 + threadIdx.x) * per;
    int loc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) loc += k < per && b0 + k < nt ? tcount[b0 + k] : 0;
    const int inc = wave_incl_scan(loc);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; w++) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_tscan(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs, int nrep,
                                                const int32_t *__restrict__ tcount, int nt, int per,
                                                const int32_t *__restrict__ part, int32_t *__restrict__ toff,
                                                uint32_t *__restrict__ moved, int2 *__restrict__ rmax,
                                                int step_delta) {
    __shared__ int wsum[16];
    const int nch = (nt + 1024 * per - 1) / (1024 * max(per, 1));  // scan blocks
    if ((int)blockIdx.x == nch) {  // the step's per-replica values
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (step_delta)
            for (int r = wv; r < nrep; r += 16) step_params_wave(ri[r], rs[r], r, lane, step_delta, -1, nullptr);
        return;
    }
    const int b0 = (blockIdx.x * 1024 + threadIdx.x) * per;
    int c8[8];
    int loc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        c8[k] = k < per && b0 + k < nt ? tcount[b0 + k] : 0;
        loc += c8[k];
    }
    int base = 0;  // counts of the chunks before this one
    if (part)
        for (int q = 0; q < (int)blockIdx.x; q++) base += part[q];
    const int inc = wave_incl_scan(loc);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) woff += wsum[w];
    int run = base + woff + inc - loc;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (k < per && b0 + k < nt) {
            toff[b0 + k] = run;
            run += c8[k];
            moved[1 + b0 + k] = 0u;  // the tile's far-arrival flag, read by k_tsort
        }
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) *moved = 0u;
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) rmax[r] = make_int2(0, 0);
}

struct StateP {
    double2 *xy, *v;
    float2 *vt;
    uint32_t *gt;
    int8_t *mp;
    int32_t *lnk;  // lipid bead: next bead of the chain; other bead: chemistry partner; -1: none
    int32_t *prv;  // lipid bead: previous bead of the chain; -1 otherwise (derived from lnk)
};

// New cell starts of the tile's rows, and the gather of its beads into the new
// order (one block per sort tile; src[new slot] = old slot when src is given,
// for the adjoint's tape); when INTEG each bead is integrated on the
// way (same arithmetic as k_keys, so the new position is the one that was
// keyed): the integrate of cel/engine/oercel.c:2303-2334 fused into the
// scatter of sort_cells, :897-921.
template <bool INTEG>
__global__ __launch_bounds__(kSortThr) void k_place(const RepInfo *__restrict__ ri, const Tile *__restrict__ tiles,
                                                     const int32_t *__restrict__ tcount,
                                                     const int32_t *__restrict__ toff,
                                                     const int32_t *__restrict__ locst,
                                                     const int32_t *__restrict__ tsrc,
                                                     const int32_t *__restrict__ tslot, int32_t *__restrict__ cs_new,
                                                     int32_t *__restrict__ src, StateP in, StateP out,
                                                     const double2 *__restrict__ f, uint32_t *__restrict__ err,
                                                     int nt, const RepStep *__restrict__ rs, ChemList CL,
                                                     uint8_t *__restrict__ cflag) {
    // Each workgroup walks tiles blockIdx.x, + gridDim.x, ...; the next tile's
    // header and bead list are loaded while the current one is placed.
    constexpr int kPl = 2;  // beads per thread per round
    const int t = threadIdx.x;
    const TileWalk W = tile_walk(nt);
    int tix = W.first;
    if (tix >= W.end) return;
    __shared__ int s_cn;  // the tile's precursor-chemistry list entries
    if (t == 0) s_cn = 0;
    __syncthreads();
    Tile T;
    int base, n, sv[kPl];
    auto load_head = [&](int x, Tile &T_, int &b_, int &n_, int *sv_) {
        T_ = tiles[x];
        b_ = toff[x];
        n_ = tcount[x];
        const int tb_ = x << kSortCapLog;
#pragma unroll
        for (int k = 0; k < kPl; k++) sv_[k] = tsrc[tb_ + t + k * kSortThr];  // in bounds; used when < n_
    };
    load_head(tix, T, base, n, sv);
    for (;;) {
        const int nx = tix + W.stride;
        Tile Tn;
        int basen = 0, nn = 0, svn[kPl];
        if (nx < W.end) load_head(nx, Tn, basen, nn, svn);
        const RepInfo &R = ri[T.rep];
        for (int row = T.k0 + t; row < T.k1; row += kSortThr) {
            const int c = R.coff + T.cx * R.ncy + row;
            cs_new[c] = base + locst[c];
            if (T.cx == R.ncx - 1 && row == R.ncy - 1) {
                // the replica's last tile: its beads end at the replica's last slot
                cs_new[c + 1] = R.off + R.N;
                if (base + n != R.off + R.N) atomicOr(err, ERR_SORT);
            }
        }
        const int tb = tix << kSortCapLog;
        for (int p0 = t; p0 < n; p0 += kPl * kSortThr) {
            bool ok[kPl];
#pragma unroll
            for (int k = 0; k < kPl; k++) {
                ok[k] = p0 + k * kSortThr < n;
                if (p0 != t && ok[k]) sv[k] = tsrc[tb + p0 + k * kSortThr];  // rounds after the first
            }
            double2 xq[kPl], vq[kPl], fq[kPl];
            float2 vt[kPl];
            uint32_t gq[kPl];
            int8_t mq[kPl];
            int32_t lq[kPl], pq[kPl];
#pragma unroll
            for (int k = 0; k < kPl; k++) {  // unconditional (slot 0 for a lane without a bead): in flight together
                const int s = ok[k] ? sv[k] : 0;
                xq[k] = in.xy[s];
                vq[k] = in.v[s];
                if (INTEG) fq[k] = f[s];
                else vt[k] = in.vt[s];
                gq[k] = in.gt[s];
                const int8_t m = in.mp[s];
                const int32_t l = in.lnk[s], pv = in.prv[s];
                mq[k] = ok[k] ? m : (int8_t)0;
                lq[k] = ok[k] ? l : -1;
                pq[k] = ok[k] ? pv : -1;
            }
            // (tile, rank) of the partners' old slots; the chemistry partner's
            // gid word at a precursor-chemistry check.  The loads are
            // unconditional (index 0 where there is no partner; lanes without a
            // bead have lq = pq = -1) so that they are in flight together, and
            // so are the tile offsets after them.
            const bool chk = CL.seg && rs[T.rep].chem_now;
            int32_t ts[kPl], tp[kPl], os[kPl], op[kPl];
            uint32_t wp[kPl];
#pragma unroll
            for (int k = 0; k < kPl; k++) {
                const int a = tslot[max(lq[k], 0)], b = tslot[max(pq[k], 0)];
                ts[k] = lq[k] >= 0 ? a : -1;
                tp[k] = mq[k] > 0 && pq[k] >= 0 ? b : -1;
            }
#pragma unroll
            for (int k = 0; k < kPl; k++) {
                wp[k] = 0u;
                if (chk) {
                    const uint32_t w = in.gt[max(lq[k], 0)];
                    wp[k] = mq[k] < 0 && lq[k] >= 0 ? w : 0u;
                }
            }
#pragma unroll
            for (int k = 0; k < kPl; k++) {
                os[k] = toff[max(ts[k], 0) >> kSortCapLog];
                op[k] = toff[max(tp[k], 0) >> kSortCapLog];
            }
#pragma unroll
            for (int k = 0; k < kPl; k++) {
                if (!ok[k]) continue;
                const int d = base + p0 + k * kSortThr;
                double2 q = xq[k], vv = vq[k];
                if (INTEG) {
                    double2 po, vo;
                    integrate1(R, q, vv, fq[k], po, vo, vt[k]);
                    q = po;
                    vv = vo;
                }
                out.xy[d] = q;
                out.v[d] = vv;
                out.vt[d] = vt[k];
                out.gt[d] = gq[k];
                if (src) src[d] = sv[k];
                const int8_t m = mq[k];
                out.mp[d] = m;
                // the partners' new slots: next / chemistry partner, and the
                // chain predecessor of a lipid bead past the head
                const int nl = ts[k] >= 0 ? os[k] + (ts[k] & (kSortCap - 1)) : -1;
                out.lnk[d] = nl;
                out.prv[d] = tp[k] >= 0 ? op[k] + (tp[k] & (kSortCap - 1)) : -1;
                if (chk) cflag[d] = chem_class(gq[k], m, nl);
                if (chk) {
                    const int role = chem_role(R, rs[T.rep], gq[k], m, nl, wp[k]);
                    if (role) {
                        ChemEnt E;
                        E.sr = d | (role << kRoleShift);
                        E.w = gq[k];
                        E.x = q.x;
                        E.y = q.y;
                        *chem_ent(CL, tix, base, atomicAdd(&s_cn, 1)) = E;
                    }
                }
            }
        }
        if (CL.seg) {
            __syncthreads();
            if (t == 0) {
                CL.seg[tix] = ChemSeg{base, s_cn, T.rep, 0};
                s_cn = 0;
            }
            __syncthreads();
        }
        if (nx >= W.end) break;
        tix = nx;
        T = Tn;
        base = basen;
        n = nn;
#pragma unroll
        for (int k = 0; k < kPl; k++) sv[k] = svn[k];
    }
}

// prv from lnk (after a state is loaded or restored), the rule of k_place
__global__ __launch_bounds__(kBeadBlock) void k_build_prv(const RepInfo *__restrict__ ri,
                                                          const int32_t *__restrict__ bblk,
                                                          const int8_t *__restrict__ mp,
                                                          const int32_t *__restrict__ lnk, int32_t *__restrict__ prv) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const int m = mp[s], l = lnk[s];
    if (m <= 0) prv[s] = -1;
    if (m >= 0 && l >= 0) prv[l] = s;
}

// --------------------------------------------------------------------------- pair forces

// entries of a wavefront's pair queue per round (8 bytes of LDS each)
constexpr int kPairQueue = 256;

constexpr float kTheta32 = (float)(3.4641016151377545870548926830117 / 4294967296.0);  // sqrt(12) / 2^32

// Caller-supplied pair noise (oc_forces_theta): theta of the pair (min gid,
// max gid) from a table of keys (lo << 32 | hi), strictly increasing; a pair
// missing from the table gets theta = 0 and raises ERR_INJ.
struct InjTheta {
    const uint64_t *k;
    const float *v;
    int64_t n;
    uint32_t *err;
};
constexpr float kInvTheta32 = (float)(4294967296.0 / 3.4641016151377545870548926830117);  // 2^32 / sqrt(12)

__device__ __forceinline__ int64_t inj_find(const uint64_t *k, int64_t n, uint32_t gi, uint32_t gj) {
    const uint64_t key = gi < gj ? (((uint64_t)gi << 32) | gj) : (((uint64_t)gj << 32) | gi);
    int64_t lo = 0, hi = n;  // first entry >= key
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (k[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && k[lo] == key ? lo : -1;
}

// Force of an interacting pair (0 < r < 1) from its unit vector (ux, uy) =
// (dx, dy) / r: F_ij = s * (ux, uy), s = w (a_ij - gamma w r.v + sigma/sqrt(dt)
// theta_ij), w = 1 - r.  gj: the candidate's gid, a: a_ij; sdtk = sigma /
// sqrt(dt) * sqrt(12) / 2^32 (the draw scale folded in).
// r * rscale is the distance: rscale a power of two, so w = 1 - r rscale in
// one fused multiply-add rounds exactly as the product then the difference.
__device__ __forceinline__ float pair_core(float ux, float uy, float r, float rscale, float vxi, float vyi, uint32_t gi,
                                           float a, float2 vj, uint32_t gj, float gam, float sdtk, uint64_t sseed,
                                           float &w, const InjTheta *inj = nullptr) {
    w = __builtin_fmaf(r, -rscale, 1.0f);
    const float dvx = vxi - vj.x, dvy = vyi - vj.y;
    const float rv = __builtin_fmaf(ux, dvx, uy * dvy);
    float th;
    if (inj) {  // the caller's theta, in the units of the hash word below
        const int64_t q = inj_find(inj->k, inj->n, gi, gj);
        if (q < 0) atomicOr(inj->err, ERR_INJ);
        th = q < 0 ? 0.0f : inj->v[q] * kInvTheta32;
    } else {
        // theta on [-sqrt3, sqrt3) from the top 32 bits of the hash
        const uint64_t h = pair_hash(sseed, gi, gj);
        th = (float)(int32_t)(uint32_t)(h >> 32);
    }
    return w * __builtin_fmaf(sdtk, th, __builtin_fmaf(-gam * w, rv, a));
}

// Staged (fixed-point) coordinates: the low 32 bits of round(x * 2^29) plus
// the periodic image shift (a multiple of round(L * 2^29)), so a difference
// of two staged coordinates is the exact integer difference for |dx| < 4 and
// the two orientations of a pair negate exactly.  The FP32 cutoff test and
// arithmetic run in units of 2^-29 (r2 scaled by 2^58: the same FP32 values,
// scaled by a power of two).
constexpr double kFixScale = 536870912.0;  // 2^29
constexpr float kFixR = 0x1p-29f;           // grid unit
// The cutoff test on staged coordinates, in exact integer arithmetic: the
// squared separation in grid units (|dx|, |dy| < 2^31, within a tile's
// window < 2^31) lies in (0, 2^58), i.e. 0 < r < 1 -- the same for both
// orientations of a pair, and for every kernel variant.
// m1 = -1, passed in a register by the loops that test many candidates, so
// that it is the addend of the first 64-bit multiply-add (two instructions
// for r^2 - 1 instead of three).
__device__ __forceinline__ bool in_cutoff_q(uint2 xi, uint2 xj, int64_t m1 = -1) {
    const int32_t dx = (int32_t)(xi.x - xj.x), dy = (int32_t)(xi.y - xj.y);
    const int64_t r2m1 = (int64_t)dy * dy + ((int64_t)dx * dx + m1);
    return (uint64_t)r2m1 < (1ull << 58) - 1ull;
}
// Low 32 bits of round(v * 2^29) (ties to even), 0 <= v * 2^29 < 2^52: the
// fused multiply-add v * 2^29 + 2^52 rounds once, to the integer grid of
// [2^52, 2^53), so its mantissa field is exactly round(v * 2^29), and its low
// word the result -- one instruction.
__device__ __forceinline__ uint32_t fix_lo(double v) {
    return (uint32_t)__double_as_longlong(__builtin_fma(v, kFixScale, 0x1p52));
}

// Pair term of i and j, added to (ax, ay) in FP64, from their fixed-point
// coordinates xq, qj (FP32 arithmetic, fused multiply-adds; both orientations
// of a pair run the same operations on negated operands, so f_ji = -f_ij
// exactly).  TEST: the cutoff test (false: the caller has made it, the pair
// interacts)
template <bool TEST = true>
__device__ __forceinline__ void pair_one_q(uint2 xq, float vxi, float vyi, uint32_t gi, uint2 qj, float2 vj,
                                           uint32_t gj, const float *ap, float gam, float sdtk, uint64_t sseed,
                                           double &ax, double &ay, float2 &e, bool energy,
                                           const InjTheta *inj = nullptr) {
    if (!TEST || in_cutoff_q(xq, qj)) {
        const float dx = (float)(int32_t)(xq.x - qj.x), dy = (float)(int32_t)(xq.y - qj.y);
        const float r2 = __builtin_fmaf(dx, dx, dy * dy);
        const float ir = __builtin_amdgcn_rsqf(r2);
        const float ux = dx * ir, uy = dy * ir;
        float w;
        const float a = *ap;
        const float fs = pair_core(ux, uy, r2 * ir, kFixR, vxi, vyi, gi, a, vj, gj, gam, sdtk, sseed, w, inj);
        ax += (double)(fs * ux);
        ay += (double)(fs * uy);
        if (energy) {
            e.x += 0.25f * a * w * w;
            e.y += 0.5f * a * w * (1.0f - w);
        }
    }
}

// harmonic bond i -> j (d = minimum image of p_j - p_i), the force of
// cel/engine/oercel.c:1280-1289, on j +fb*d and on i -fb*d with fb = -k_b
// (r - r0) / r.  The length comes from the hardware's FP64 reciprocal square
// root refined by two Newton steps (relative error ~1e-15, against a sqrt and
// a division of ~25 instructions): ir = 1 / r, r = r^2 ir.
__device__ __forceinline__ void bond_vec(const RepInfo &R, double2 pi, double2 pj, double &dx, double &dy, double &r,
                                         double &ir) {
    dx = minimg(pj.x - pi.x, R.L, R.halfL);
    dy = minimg(pj.y - pi.y, R.L, R.halfL);
    const double r2 = dx * dx + dy * dy, h = 0.5 * r2;
    double y = __builtin_amdgcn_rsq(r2);
    y = y * __builtin_fma(-h * y, y, 1.5);
    y = y * __builtin_fma(-h * y, y, 1.5);
    ir = y;
    r = r2 * y;
}
__device__ __forceinline__ double bond_fb(const RepInfo &R, double r, double ir) { return -R.kb * (r - R.r0) * ir; }

// Links and partner data of bead s for its bonded terms, loaded in two
// rounds: bond_links at the start of the bead (the pair loop hides their
// latency), bond_partners after the pair loop.
struct BondPre {
    int m, n1, p1;   // chain position (-1: none), next/partner slot, previous slot
    double2 pn, pp;  // positions of n1 and p1
    uint32_t gn;     // gid word of n1 (chemistry bond)
};
__device__ __forceinline__ void bond_links(int s, const int8_t *__restrict__ mp, const int32_t *__restrict__ lnk,
                                           const int32_t *__restrict__ prv, BondPre &B) {
    B.m = mp[s];
    B.n1 = lnk[s];
    B.p1 = prv[s];
}
__device__ __forceinline__ void bond_partners(const double2 *__restrict__ xy, const uint32_t *__restrict__ gt,
                                              BondPre &B) {
    if (B.n1 >= 0) {
        B.pn = xy[B.n1];
        B.gn = gt[B.n1];
    }
    if (B.m > 0 && B.p1 >= 0) B.pp = xy[B.p1];
}

// Bonded force on bead s in gather form: every bond and angle term that s
// takes part in is evaluated whole, with the terms of chain_forces_linear
// (cel/engine/oercel.c:1267-1308; bond lengths as bond_vec forms them), and s
// adds its share, bonds along the chain first, then angles -- the order in
// which a per-chain pass adds them to each bead, so both beads of a bond
// evaluate it on the same operands and receive exactly opposite forces.  A chemistry bond is
// the bond from the smaller gid to the larger.  mb: largest bond length at s.
// The angle terms (ANG: some replica has k_angle != 0; each replica's own
// k_angle applies) load the beads two links away here.
template <bool ANG>
__device__ __forceinline__ void bonded_gather(const RepInfo &R, double2 ps, uint32_t gs, const BondPre &B,
                                              const double2 *__restrict__ xy, const int32_t *__restrict__ lnk,
                                              const int32_t *__restrict__ prv, double &fx, double &fy, double &mb) {
    const int m = B.m, n1 = B.n1;
    if (m < 0) {
        if (n1 < 0) return;
        const bool lo = (gs & kGidMask) < (B.gn & kGidMask);
        double dx, dy, r, ir;
        bond_vec(R, lo ? ps : B.pn, lo ? B.pn : ps, dx, dy, r, ir);
        mb = fmax(mb, r);
        const double fb = bond_fb(R, r, ir);
        if (lo) {
            fx -= fb * dx;
            fy -= fb * dy;
        } else {
            fx += fb * dx;
            fy += fb * dy;
        }
        return;
    }
    if (!ANG) {
        // bonds to the previous and the next chain bead
        if (m > 0 && B.p1 >= 0) {
            double dx, dy, r, ir;
            bond_vec(R, B.pp, ps, dx, dy, r, ir);
            mb = fmax(mb, r);
            const double fb = bond_fb(R, r, ir);
            fx += fb * dx;
            fy += fb * dy;
        }
        if (n1 >= 0) {
            double dx, dy, r, ir;
            bond_vec(R, ps, B.pn, dx, dy, r, ir);
            mb = fmax(mb, r);
            const double fb = bond_fb(R, r, ir);
            fx -= fb * dx;
            fy -= fb * dy;
        }
        return;
    }
    // chain beads at offsets -2..+2 (index 0..4, s at 2); bond b joins c[b] -> c[b+1]
    const bool ang = R.ka != 0.0;
    int c[5];
    double2 p[5];
    c[2] = 0;
    p[2] = ps;
    c[1] = m > 0 ? B.p1 : -1;
    p[1] = B.pp;
    c[3] = n1;
    p[3] = B.pn;
    c[0] = (ang && c[1] >= 0) ? prv[c[1]] : -1;
    c[4] = (ang && c[3] >= 0) ? lnk[c[3]] : -1;
    if (c[0] >= 0) p[0] = xy[c[0]];
    if (c[4] >= 0) p[4] = xy[c[4]];
    double bx[4], by[4], bl[4], bi[4];
    bool has[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        has[b] = c[b] >= 0 && c[b + 1] >= 0;
        if (has[b]) bond_vec(R, p[b], p[b + 1], bx[b], by[b], bl[b], bi[b]);
    }
    if (has[1]) {  // s is the second bead of bond 1
        mb = fmax(mb, bl[1]);
        const double fb = bond_fb(R, bl[1], bi[1]);
        fx += fb * bx[1];
        fy += fb * by[1];
    }
    if (has[2]) {  // s is the first bead of bond 2
        mb = fmax(mb, bl[2]);
        const double fb = bond_fb(R, bl[2], bi[2]);
        fx -= fb * bx[2];
        fy -= fb * by[2];
    }
    if (!ang) return;
    const double ka = R.ka;
    // angle a uses bonds a, a+1 (beads c[a], c[a+1], c[a+2]); s is its third,
    // second, first bead for a = 0, 1, 2
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (!has[a] || !has[a + 1]) continue;
        const double ax = bx[a], ay = by[a], cxv = bx[a + 1], cyv = by[a + 1];
        const double la = bl[a], lc = bl[a + 1];
        const double ilac = 1.0 / (la * lc);
        const double csn = (ax * cxv + ay * cyv) * ilac;
        const double dUax = -ka * (cxv * ilac - csn * ax / (la * la));
        const double dUay = -ka * (cyv * ilac - csn * ay / (la * la));
        const double dUcx = -ka * (ax * ilac - csn * cxv / (lc * lc));
        const double dUcy = -ka * (ay * ilac - csn * cyv / (lc * lc));
        if (a == 0) {
            fx -= dUcx;
            fy -= dUcy;
        } else if (a == 1) {
            fx += -dUax + dUcx;
            fy += -dUay + dUcy;
        } else {
            fx += dUax;
            fy += dUay;
        }
    }
}

struct PairArgs {
    const RepInfo *ri;
    const RepStep *rs;
    RepStep *rsw;        // running max bond length (maxbond_bits)
    const int8_t *mp;    // bonded terms (gather form) when bonded != 0
    const int32_t *lnk, *prv;
    int bonded;
    const Tile *tiles;
    int ntiles, per_xcd;
    const int32_t *cs;
    const double2 *xy;
    const float2 *vt;
    const uint32_t *gt;
    const double *stir;
    double2 *f;
    double *epot, *evir;  // per-bead pair energy and conservative virial (ENERGY)
    InjTheta inj;          // caller-supplied pair noise (INJ)
    uint32_t *err;
    // next step's cell keys (new_key of the integrated position) when key != 0
    const double2 *v;
    int32_t *key;
    int2 *rmax;
    uint32_t *moved;
    // the state's link words (lnk, writable): a free T's pending formation
    // proposal of this step's check is settled by its bead (chem_settle),
    // with the event
    int32_t *lnkw;
    Event *ev;
    unsigned long long *evn;
    int64_t evcap;
    int record;
};

// A pending proposal of bead s settled before its bonded terms: the link word
// becomes the partner (mutual) or -1; the bond's smaller gid records the
// event once the partner's gid is loaded (chem_formed).
__device__ __forceinline__ bool chem_pending(const PairArgs &A, int slot, BondPre &B) {
    if (B.n1 >= -1) return false;
    B.n1 = chem_settle(slot, B.n1, A.lnk);
    A.lnkw[slot] = B.n1;
    return B.n1 >= 0;
}
__device__ __forceinline__ void chem_formed(const PairArgs &A, const RepStep &S, int rep, uint32_t gi,
                                            const BondPre &B) {
    const uint32_t gn = B.gn & kGidMask;
    if (A.record && gi < gn) push_event(A.ev, A.evn, A.evcap, A.err, S.step, OC_EV_FORM, gi, gn, rep);
}

// Full-shell pair forces of the i-beads of one tile (a column, rows [k0,k1)).
// The candidate beads -- columns cx-1..cx+1, rows k0-M..k1-1+M -- are staged
// in LDS once per tile, row-major: staged row r holds the beads of its three
// columns (column cx-1, cx, cx+1), each at the position of its periodic image
// next to the tile.  A lane's y-window (the window expressions of
// cel/engine/oercel.c:1073-1074) over the three columns is then one
// contiguous LDS range.  Forces are summed in FP64 per bead, no atomics: the
// result is deterministic.
// rows per staged column (tile rows + 2 * margin + 1) of a pair tile
__host__ __device__ constexpr int pair_max_rows(int block) { return 2 * block; }
constexpr int64_t kPairSlotLimit = int64_t(1) << 28;  // slots the staging packs (bits 0-27)

// Occupancy: the kernel issues VALU work at nearly every cycle, and one
// wavefront per SIMD more (7 against 6 at BLOCK 256: at most 72 VGPRs and
// 22.8 KB of LDS a workgroup) hides more of the LDS latency of the candidate
// loop (k_pair 60.3 -> 58.6 us at N = 1M, rho 3); the angle terms and the
// injected noise would spill at 72 VGPRs and the pair queue's LDS allows five,
// so those kernels keep the compiler's choice.
template <int BLOCK, int CAP, int CAPL, bool ENERGY, bool ANG, bool INJ = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(ANG || ENERGY || INJ || CAPL == 2 ? 1 : 7)))
void k_pair(
    PairArgs A) {
    const InjTheta *inj = INJ ? &A.inj : nullptr;  // caller-supplied pair noise (single pass only)
    static_assert(CAP < 65536, "staged indices are 16-bit");
    constexpr int kMaxRows = pair_max_rows(BLOCK);
    static_assert(kMaxRows <= 2 * BLOCK, "two staged rows per thread");
    // Staged candidate j, three 8-byte words at the same stride (one address
    // for all three in the candidate loop): s_st[j] its fixed-point coordinates
    // (its image next to the tile), s_st[CAP + j] its predicted velocity (x, y
    // bits), s_st[2 CAP + j] its gid and the byte offset of its type in a row
    // of the a_ij table.
    __shared__ uint2 s_st[3 * CAP];
    uint2 *const s_xq = s_st;
    // before the staging loads: each candidate's source (slot | column | image)
    uint32_t *s_src = reinterpret_cast<uint32_t *>(s_st + CAP);
    __shared__ uint16_t s_rs[kMaxRows + 1];   // LDS index of the first candidate of staged row r (+ the end)
    __shared__ uint16_t s_ib[CAP];            // LDS index of the tile's q-th i-bead
    __shared__ float s_a[64];
    __shared__ int32_t s_ws[BLOCK / 64][2];

    // XCD-aware tile order: consecutive workgroups go to different XCDs
    // round-robin, so give each XCD a contiguous run of tiles (neighbouring
    // columns share their staged beads through that XCD's L2).
    const int t = (blockIdx.x & 7) * A.per_xcd + (blockIdx.x >> 3);
    if (t >= A.ntiles) return;
    const Tile T = A.tiles[t];
    const RepInfo &R = A.ri[T.rep];
    const RepStep &S = A.rs[T.rep];
    const int ncy = R.ncy, ncx = R.ncx;
    const int M = (int)ceil(R.icsy) + 1;  // staging margin in rows
    const int kb0 = T.k0 - M;             // first staged (unwrapped) row
    const int nr = T.k1 - T.k0 + 2 * M;   // staged rows kb0 .. T.k1 - 1 + M
    const int32_t *cs = A.cs;
    const double L = R.L;
    const int tid = (int)threadIdx.x;
    if (nr + 1 > kMaxRows) {  // the host sizes tiles so that this cannot happen
        if (tid == 0) atomicOr(A.err, ERR_TILE);
        return;
    }
    // the three columns (cell base, x image shift)
    int colb[3];
    double sxc[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int col = T.cx - 1 + c;
        sxc[c] = col < 0 ? -L : (col >= ncx ? L : 0.0);
        col = col < 0 ? col + ncx : (col >= ncx ? col - ncx : col);
        colb[c] = R.coff + col * ncy;
    }
    // the tile's i-beads: slots [slot0, slot0 + ni) of column cx
    const int slot0 = cs[colb[1] + T.k0], ni = cs[colb[1] + T.k1] - slot0;
    if (tid < 64) s_a[tid] = S.a_eff[tid];

    // The prologue costs two rounds of memory latency: the cell starts of the
    // staged rows (thread t: rows t and t + BLOCK, three columns), then the
    // staged beads.  Row r's beads start at the exclusive prefix of the row
    // totals (a block scan of both rows' totals).
    // (loads unconditional, rows past the staged ones clamped to the last, so
    // that all twelve are in flight together)
    int f[2][3], n[2][3], tot[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = kb0 + min(tid + h * BLOCK, nr - 1);
        const int w = k < 0 ? k + ncy : (k < ncy ? k : k - ncy);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            f[h][c] = cs[colb[c] + w];
            n[h][c] = cs[colb[c] + w + 1];
        }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bool in = tid + h * BLOCK < nr;
        tot[h] = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            n[h][c] = in ? n[h][c] - f[h][c] : 0;
            f[h][c] = in ? f[h][c] : 0;
            tot[h] += n[h][c];
        }
    }
    const int lane = tid & 63, wv = tid >> 6;
    const int inc0 = wave_incl_scan(tot[0]), inc1 = wave_incl_scan(tot[1]);
    if (lane == 63) {
        s_ws[wv][0] = inc0;
        s_ws[wv][1] = inc1;
    }
    __syncthreads();
    int off0 = 0, off1 = 0, sum0 = 0, sum1 = 0;
#pragma unroll
    for (int q = 0; q < BLOCK / 64; q++) {
        const int a0 = s_ws[q][0], a1 = s_ws[q][1];
        off0 += q < wv ? a0 : 0;
        off1 += q < wv ? a1 : 0;
        sum0 += a0;
        sum1 += a1;
    }
    const int total = sum0 + sum1;
    const bool staged = total <= CAP;
    const int rs[2] = {off0 + inc0 - tot[0], sum0 + off1 + inc1 - tot[1]};
    if (staged) {
        // row starts; each candidate's source: slot (bits 0-27), column (28-29),
        // y image (30-31: 0 none, 1 -L, 2 +L); the i-bead table
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = tid + h * BLOCK;
            if (r >= nr) continue;
            s_rs[r] = (uint16_t)rs[h];
            const int k = kb0 + r;
            const uint32_t syc = k < 0 ? 1u : (k < ncy ? 0u : 2u);
            const bool irow = k >= T.k0 && k < T.k1;
            int d = rs[h];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const uint32_t code = ((uint32_t)c << 28) | (syc << 30);
                for (int b = 0; b < n[h][c]; b++) {
                    s_src[d + b] = (uint32_t)(f[h][c] + b) | code;
                    if (c == 1 && irow) s_ib[f[h][c] + b - slot0] = (uint16_t)(d + b);
                }
                d += n[h][c];
            }
        }
        if (tid == 0) s_rs[nr] = (uint16_t)total;
    }
    __syncthreads();
    // the x / y image shifts in fixed point: columns 0 and 2, rows below 0 and
    // past ncy - 1 (the staged coordinates, and the global-memory path's)
    const uint32_t fL = fix_lo(L);
    const uint32_t fx0 = sxc[0] < 0.0 ? 0u - fL : (sxc[0] > 0.0 ? fL : 0u);
    const uint32_t fx2 = sxc[2] < 0.0 ? 0u - fL : (sxc[2] > 0.0 ? fL : 0u);
    if (staged) {
        // every load of a thread is issued before its first LDS store
        constexpr int kStgPer = (CAP + BLOCK - 1) / BLOCK;
        double2 sx[kStgPer];
        float2 sv[kStgPer];
        uint32_t sg[kStgPer];
        uint32_t sc[kStgPer];
#pragma unroll
        for (int u = 0; u < kStgPer; u++) {
            if (u * BLOCK >= total) break;
            const int e = tid + u * BLOCK;
            if (e < total) {
                const uint32_t code = s_src[e];
                const int src = (int)(code & 0x0fffffffu), c = (int)((code >> 28) & 3u), syc = (int)(code >> 30);
                sx[u] = A.xy[src];
                sv[u] = A.vt[src];
                sg[u] = A.gt[src];
                sc[u] = (uint32_t)c | ((uint32_t)syc << 2);
            }
        }
        __syncthreads();  // the sources are read: the velocity words take the candidates
#pragma unroll
        for (int u = 0; u < kStgPer; u++) {
            if (u * BLOCK >= total) break;
            const int e = tid + u * BLOCK;
            if (e < total) {
                const uint32_t c = sc[u] & 3u, syc = sc[u] >> 2;
                s_xq[e] = make_uint2(fix_lo(sx[u].x) + (c == 0 ? fx0 : (c == 2 ? fx2 : 0u)),
                                     fix_lo(sx[u].y) + (syc == 1 ? 0u - fL : (syc == 2 ? fL : 0u)));
                s_st[CAP + e] = make_uint2(__float_as_uint(sv[u].x), __float_as_uint(sv[u].y));
                s_st[2 * CAP + e] = make_uint2(sg[u] & kGidMask, (sg[u] >> kGidBits) * 4u);
            }
        }
        __syncthreads();
    }

    const float gam = S.gamma, sdtk = S.sdt * kTheta32;
    const uint64_t sseed = S.sseed;
    // A tile whose candidates overflow the LDS buffer (a dense clump) runs the
    // per-column loop on global memory.  Each lane owns one i-bead at a time.
    double mb = 0.0;  // largest bond length seen by this lane
    // per-bead epilogue: energy, or force (+ stirring, + bonded terms)
    auto finish = [&](int slot, double2 pi, uint32_t gti, const BondPre &B, double ax, double ay, float2 e,
                      double2 vh) {
        if (ENERGY) {
            A.epot[slot] = (double)e.x;
            A.evir[slot] = (double)e.y;
        } else {
            double bx = 0.0;
            if (A.stir) bx = A.stir[slot] - S.stir_mean;
            double fx = bx + ax, fy = ay;
            // a branched lipid's terms come from k_mol (chemistry bonds stay here)
            if (A.bonded && (R.arch == 0 || B.m < 0)) bonded_gather<ANG>(R, pi, gti, B, A.xy, A.lnk, A.prv, fx, fy, mb);
            A.f[slot] = make_double2(fx, fy);
            if (A.key) {
                // the next step's integrate, as k_keys<true> would run it
                double2 q, vo;
                float2 vtn;
                integrate1(R, pi, vh, make_double2(fx, fy), q, vo, vtn);
                new_key(R, T.rep, slot, pi, q, A.key, A.rmax, A.moved);
            }
        }
    };
    // Cross-lane pair queue (CAPL == 2): per wavefront, kPairQueue entries of
    // LDS, each first an (i, j) pair of staged indices, then its force.
    __shared__ uint2 s_q[CAPL == 2 ? (BLOCK / 64) * kPairQueue : 1];
    // candidate j's predicted velocity and gid word: two 8-byte LDS reads at
    // the stride of its coordinates when staged
    auto run = [&](const uint2 *XQ, const float2 *VT, const uint32_t *GT, const bool stg) {
        const double2 *XY = A.xy;
        // candidate j: velocity, gid and the address of a_ij in row arow
        auto cand = [&](int j, const float *arow, float2 &v, uint32_t &g, const float *&ap) {
            if (stg) {
                const uint2 wv = XQ[CAP + j], wg = XQ[2 * CAP + j];
                v = make_float2(__uint_as_float(wv.x), __uint_as_float(wv.y));
                g = wg.x;
                ap = (const float *)((const char *)arow + wg.y);
            } else {
                const uint32_t gt = GT[j];
                v = VT[j];
                g = gt & kGidMask;
                ap = arow + (gt >> kGidBits);
            }
        };
        // block-uniform bead loop: every lane of a wavefront takes part in the
        // wavefront-wide steps, lanes past the tile's last bead (act false)
        // with an empty window
        for (int ib = 0; ib < ni; ib += BLOCK) {
            const int q = ib + tid;
            const bool act = q < ni;
            const int slot = slot0 + (act ? q : 0);
            const int is = stg ? (int)s_ib[act ? q : 0] : slot;  // an i-bead for every lane (ni > 0)
            BondPre B;
            const bool bond = !ENERGY && A.bonded && act;
            if (bond) bond_links(slot, A.mp, A.lnk, A.prv, B);
            const double2 vh = !ENERGY && A.key && act ? A.v[slot] : make_double2(0.0, 0.0);
            const double2 pi = XY[slot];
            // fixed-point coordinates of the bead: staged, or from its position
            // (the global-memory path runs the same arithmetic as the staged one)
            const uint2 xq = stg ? XQ[is] : make_uint2(fix_lo(pi.x), fix_lo(pi.y));
            float2 vi;
            uint32_t gi;
            const float *ai;  // s_a + bead i's type
            cand(is, s_a, vi, gi, ai);
            const float *arow = s_a + 8 * (ai - s_a);  // the a_ij row of i's type
            int khi = (int)((pi.y + 1.0) * R.icsy);
            int klo = (int)((pi.y - 1.0) * R.icsy + ncy) - ncy;
            if (klo < kb0 || khi > T.k1 - 1 + M) {  // cannot happen for rows k0..k1-1 (margin M); guard
                atomicOr(A.err, ERR_TILE);
                klo = max(klo, kb0);
                khi = min(khi, T.k1 - 1 + M);
            }
            double ax = 0.0, ay = 0.0;
            float2 e = make_float2(0.0f, 0.0f);
            if (stg) {
                // the window: staged rows klo..khi of the three columns, one LDS range
                const int st = s_rs[klo - kb0], en = s_rs[khi + 1 - kb0];
                const int nj = act ? en - st : 0;
                int j = st;
                if (CAPL == 2 && !ENERGY && !__any(nj > 64)) {
                    //   CAPL == 2: the marked pairs of all 64 lanes go into one
                    // queue in lane order, the whole wavefront evaluates it 64
                    // pairs at a time (~mean rather than ~max of the pair
                    // count), and each lane sums its own pairs' forces back in
                    // sequence order -- the same FP32 terms added in the same
                    // order as the single pass, so the forces are bit-identical
                    // to it.  Phase 1: bit t of mk marks an interacting candidate t.
                    uint64_t mk = 0;
                    for (int it = 0; it < nj; it++, j++)
                        if (in_cutoff_q(xq, XQ[j])) mk |= 1ull << it;  // the test of pair_one_q
                    // this lane's first entry in the wavefront's queue
                    const int cnt = __popcll(mk);
                    const int incl = wave_incl_scan(cnt);
                    const int off = incl - cnt, qtot = __shfl(incl, 63, 64);
                    uint2 *qb = s_q + (tid >> 6) * kPairQueue;
                    uint64_t rem = mk;
                    int pos = off;
                    // rounds of kPairQueue entries (wavefront-uniform)
                    for (int w0 = 0; w0 < qtot; w0 += kPairQueue) {
                        const int w1 = min(w0 + kPairQueue, qtot);
                        // phase 2: this lane's pairs that fall in the round
                        while (rem != 0 && pos < w1) {
                            const int tk = __builtin_ctzll(rem);
                            rem &= rem - 1;
                            qb[pos - w0] = make_uint2((uint32_t)is | ((uint32_t)(st + tk) << 16), 0u);
                            ++pos;
                        }
                        __syncwarp();
                        // phase 3: the round's pairs, one per lane
                        for (int q3 = (tid & 63); q3 < w1 - w0; q3 += 64) {
                            const uint32_t u = qb[q3].x;
                            const int i1 = (int)(u & 0xffffu), j1 = (int)(u >> 16);
                            const uint2 p1 = XQ[i1], p2 = XQ[j1];
                            float2 v1, v2;
                            uint32_t g1, g2;
                            const float *a1r, *ap;
                            cand(i1, s_a, v1, g1, a1r);
                            cand(j1, s_a + 8 * (a1r - s_a), v2, g2, ap);
                            const float dx = (float)(int32_t)(p1.x - p2.x), dy = (float)(int32_t)(p1.y - p2.y);
                            const float r2 = __builtin_fmaf(dx, dx, dy * dy);
                            const float ir = __builtin_amdgcn_rsqf(r2);
                            const float ux = dx * ir, uy = dy * ir;
                            float w;
                            const float fs = pair_core(ux, uy, r2 * ir, kFixR, v1.x, v1.y, g1, *ap, v2, g2, gam,
                                                       sdtk, sseed, w);
                            qb[q3] = make_uint2(__float_as_uint(fs * ux), __float_as_uint(fs * uy));
                        }
                        __syncwarp();
                        // phase 4: this lane's forces of the round, in sequence order
                        const int a1 = min(off + cnt, w1);
                        for (int p = max(off, w0); p < a1; p++) {
                            const uint2 fq = qb[p - w0];
                            ax += (double)__uint_as_float(fq.x);
                            ay += (double)__uint_as_float(fq.y);
                        }
                        __syncwarp();
                    }
                } else if (CAPL != 1) {
                    // the single pass, the staged words addressed through
                    // one pointer (one induction variable for the loop)
                    for (const uint2 *pj = XQ + st, *pe = XQ + st + nj; pj < pe; pj++) {
                        const uint2 wv = pj[CAP], wg = pj[2 * CAP];
                        const float2 vj = make_float2(__uint_as_float(wv.x), __uint_as_float(wv.y));
                        const float *ap = (const float *)((const char *)arow + wg.y);
                        pair_one_q(xq, vi.x, vi.y, gi, *pj, vj, wg.x, ap, gam, sdtk, sseed, ax, ay, e, ENERGY, inj);
                    }
                } else {
                    //   CAPL == 1: two phases per lane: the cutoff test over
                    // the window marks the interacting candidates among the
                    // first 64 in two 32-bit register masks (later ones are
                    // evaluated in place), then the force and noise arithmetic
                    // runs over the marked ones only.  The candidate counter is
                    // the same in every lane (uniform), so a mark is one OR.
                    uint32_t mw[2] = {0u, 0u};
                    {
                        // the marks without branches: one loop per mask word,
                        // candidates past the 64th evaluated in place
                        const uint2 *pj = XQ + st;
                        const int n0 = min(nj, 32), n1 = min(nj, 64);
                        int64_t m1 = -1;  // in a register (an empty asm hides the constant)
                        asm volatile("" : "+v"(m1));
                        int it = 0;
                        for (; it + 1 < n0; it += 2, pj += 2)
                            mw[0] |= ((uint32_t)in_cutoff_q(xq, pj[0], m1) | (uint32_t)in_cutoff_q(xq, pj[1], m1) << 1) << it;
                        for (; it < n0; it++, pj++) mw[0] |= (uint32_t)in_cutoff_q(xq, *pj, m1) << it;
                        for (; it + 1 < n1; it += 2, pj += 2)
                            mw[1] |= ((uint32_t)in_cutoff_q(xq, pj[0], m1) | (uint32_t)in_cutoff_q(xq, pj[1], m1) << 1)
                                     << (it - 32);
                        for (; it < n1; it++, pj++) mw[1] |= (uint32_t)in_cutoff_q(xq, *pj, m1) << (it - 32);
                        for (; it < nj; it++, pj++) {
                            if (!in_cutoff_q(xq, *pj)) continue;
                            const uint2 wv = pj[CAP], wg = pj[2 * CAP];
                            const float2 vj = make_float2(__uint_as_float(wv.x), __uint_as_float(wv.y));
                            const float *ap = (const float *)((const char *)arow + wg.y);
                            pair_one_q<false>(xq, vi.x, vi.y, gi, *pj, vj, wg.x, ap, gam, sdtk, sseed, ax, ay, e,
                                              ENERGY, inj);
                        }
                    }
                    // pass 2: the marked candidates in window order, the
                    // mask as two 32-bit words (most windows fit the first)
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        uint32_t m = mw[h];
                        const uint2 *pb = XQ + st + 32 * h;
                        while (m) {
                            const uint2 *pj = pb + __builtin_ctz(m);
                            m &= m - 1;
                            const uint2 wv = pj[CAP], wg = pj[2 * CAP];
                            const float2 vj = make_float2(__uint_as_float(wv.x), __uint_as_float(wv.y));
                            const float *ap = (const float *)((const char *)arow + wg.y);
                            pair_one_q<false>(xq, vi.x, vi.y, gi, *pj, vj, wg.x, ap, gam, sdtk, sseed, ax, ay, e,
                                              ENERGY, inj);
                        }
                    }
                }
            } else if (act) {
                // global memory: per column and periodic piece (image shifts),
                // candidates in fixed point as the staging forms them, so a
                // pair's two orientations cancel exactly whichever path each
                // side of it takes
                for (int c = 0; c < 3; c++) {
                    const uint32_t sxq = c == 0 ? fx0 : (c == 2 ? fx2 : 0u);
                    // window split at the periodic seams: rows < 0 (image -L),
                    // 0..ncy-1, >= ncy (image +L)
                    for (int pq = 0; pq < 3; pq++) {
                        int pa, pb, wrap;
                        uint32_t syq;
                        if (pq == 0) { pa = klo; pb = min(khi, -1); syq = 0u - fL; wrap = ncy; }
                        else if (pq == 1) { pa = max(klo, 0); pb = min(khi, ncy - 1); syq = 0u; wrap = 0; }
                        else { pa = max(klo, ncy); pb = khi; syq = fL; wrap = -ncy; }
                        if (pa > pb) continue;
                        const int ja = cs[colb[c] + pa + wrap], jb = cs[colb[c] + pb + wrap + 1];
                        for (int jj = ja; jj < jb; jj++)
                        {
                            float2 vj;
                            uint32_t gj;
                            const float *ap;
                            cand(jj, arow, vj, gj, ap);
                            const double2 pj = XY[jj];
                            const uint2 qj = make_uint2(fix_lo(pj.x) + sxq, fix_lo(pj.y) + syq);
                            pair_one_q(xq, vi.x, vi.y, gi, qj, vj, gj, ap, gam, sdtk, sseed, ax, ay, e, ENERGY, inj);
                        }
                    }
                }
            }
            if (act) {
                const bool formed = bond && chem_pending(A, slot, B);
                if (bond) bond_partners(A.xy, A.gt, B);
                if (formed) chem_formed(A, S, T.rep, gi & kGidMask, B);
                finish(slot, pi, gi, B, ax, ay, e, vh);
            }
        }
    };
    if (staged) run(s_st, nullptr, nullptr, true);
    else run(nullptr, A.vt, A.gt, false);
    if (ENERGY || !A.bonded) return;
    // running max bond length of the replica: block max, then an atomic only
    // when it exceeds the stored value (positive doubles order as their bits)
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

// --------------------------------------------------------------------------- bonded forces

// harmonic bond i -> j, the arithmetic of cel/engine/oercel.c:1280-1289
template <bool WF>
__device__ __forceinline__ void bond1(const RepInfo &R, int i, int j, const double2 *xy, double2 *f, double &mb,
                                      double &eb, double &bx, double &by, double &bl) {
    const double2 pi = xy[i], pj = xy[j];
    const double dx = minimg(pj.x - pi.x, R.L, R.halfL), dy = minimg(pj.y - pi.y, R.L, R.halfL);
    const double r = sqrt(dx * dx + dy * dy);
    bx = dx; by = dy; bl = r;
    if (r > mb) mb = r;
    const double fb = -R.kb * (r - R.r0) / r;
    if (WF) {
        f[j].x += fb * dx;
        f[j].y += fb * dy;
        f[i].x -= fb * dx;
        f[i].y -= fb * dy;
    }
    eb += 0.5 * R.kb * (r - R.r0) * (r - R.r0);
}

// One thread per bead: a lipid head walks its chain (bonds, then angle terms,
// chain_forces_linear, cel/engine/oercel.c:1267-1308); the smaller-gid bead of
// a chemistry bond applies that bond.  Every bead belongs to at most one of
// them, so the writes never collide.
template <bool ENERGY>
__global__ __launch_bounds__(kBeadBlock) void k_bonded(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                       const int32_t *__restrict__ bblk,
                                                       const double2 *__restrict__ xy, const uint32_t *__restrict__ gt,
                                                       const int8_t *__restrict__ mp, const int32_t *__restrict__ lnk,
                                                       double2 *__restrict__ f, double *__restrict__ ebond,
                                                       uint32_t *__restrict__ err) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    double mb = 0.0, eb = 0.0;
    bool owner = false;
    if (loc < R.N) {
        const int s = R.off + loc;
        const int m = mp[s];
        if (m == 0 && R.arch == 0) {
            owner = true;
            int sl[OC_MAXMLEN];
            double bx[OC_MAXMLEN], by[OC_MAXMLEN], bl[OC_MAXMLEN];
            int n = 0;
            for (int d = s; d >= 0 && n < R.mlen; d = lnk[d]) sl[n++] = d;
            if (n != R.mlen) atomicOr(err, ERR_CHAIN);
            for (int b = 0; b < n - 1; b++) bond1<!ENERGY>(R, sl[b], sl[b + 1], xy, f, mb, eb, bx[b], by[b], bl[b]);
            const double ka = R.ka;
            if (ka != 0.0) {
                for (int b = 0; b < n - 2; b++) {
                    const int i = sl[b], j = sl[b + 1], k = sl[b + 2];
                    const double ax = bx[b], ay = by[b], cxv = bx[b + 1], cyv = by[b + 1];
                    const double la = bl[b], lc = bl[b + 1];
                    const double ilac = 1.0 / (la * lc);
                    const double csn = (ax * cxv + ay * cyv) * ilac;
                    const double dUax = -ka * (cxv * ilac - csn * ax / (la * la));
                    const double dUay = -ka * (cyv * ilac - csn * ay / (la * la));
                    const double dUcx = -ka * (ax * ilac - csn * cxv / (lc * lc));
                    const double dUcy = -ka * (ay * ilac - csn * cyv / (lc * lc));
                    if (!ENERGY) {
                        f[i].x += dUax;
                        f[i].y += dUay;
                        f[j].x += -dUax + dUcx;
                        f[j].y += -dUay + dUcy;
                        f[k].x -= dUcx;
                        f[k].y -= dUcy;
                    }
                    eb += ka * (1.0 - csn);
                }
            }
        } else if (m < 0) {
            const int p = lnk[s];
            if (p >= 0 && (gt[s] & kGidMask) < (gt[p] & kGidMask)) {
                owner = true;
                double bx, by, bl;
                bond1<!ENERGY>(R, s, p, xy, f, mb, eb, bx, by, bl);
            }
        }
        if (ENERGY) ebond[s] = owner ? eb : 0.0;
    }
    if (!ENERGY) {
        uint64_t bits = (uint64_t)__double_as_longlong(mb);
        bits = wave_max_u64(bits);
        unsigned long long *mbw = (unsigned long long *)&rs[r].maxbond_bits;
        if ((threadIdx.x & 63) == 0 && bits > __hip_atomic_load(mbw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(mbw, (unsigned long long)bits);
    }
}

// Bonded terms of branched lipids (pi, y, bola): one thread per bead; the head
// (mpos 0) of a lipid of a replica with a template collects its molecule's
// slots along the chain (template order) and evaluates the template's bonds,
// then its angle terms, each added to the beads it acts on in template order
// -- the arithmetic of mol_forces (cel/engine/oercel.c:1218-1257).  The
// molecule's forces are read once, updated in that order and written back:
// every bead belongs to one molecule, so the writes never collide, and each
// bead's sum is its pair force plus its terms in the reference's order.
// ENERGY: the molecule's bond and angle energy at its head (ebond), no forces.
template <bool ENERGY>
__global__ __launch_bounds__(kBeadBlock) void k_mol(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                    const int32_t *__restrict__ bblk,
                                                    const double2 *__restrict__ xy, const int8_t *__restrict__ mp,
                                                    const int32_t *__restrict__ lnk, double2 *__restrict__ f,
                                                    double *__restrict__ ebond, uint32_t *__restrict__ err) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    double mb = 0.0;
    if (R.arch != 0 && loc < R.N) {
        const int s = R.off + loc;
        if (mp[s] == 0) {
            int sl[OC_MAXMLEN];
            int n = 0;
            for (int d = s; d >= 0 && n < R.mlen; d = lnk[d]) sl[n++] = d;
            double eb = 0.0;
            if (n != R.mlen) {
                atomicOr(err, ERR_CHAIN);
            } else {
                double2 q[OC_MAXMLEN], fl[OC_MAXMLEN];
                for (int b = 0; b < n; b++) {
                    q[b] = xy[sl[b]];
                    if (!ENERGY) fl[b] = f[sl[b]];
                }
                const double L = R.L, hL = R.halfL, kb = R.kb;
                for (int t = 0; t < R.tnb; t++) {
                    const int i = R.bi[t], j = R.bj[t];
                    const double dx = minimg(q[j].x - q[i].x, L, hL), dy = minimg(q[j].y - q[i].y, L, hL);
                    const double rr = sqrt(dx * dx + dy * dy);
                    if (rr > mb) mb = rr;
                    const double fb = -kb * (rr - R.br0[t]) / rr;  // force on j along (dx, dy)
                    if (!ENERGY) {
                        fl[j].x += fb * dx;
                        fl[j].y += fb * dy;
                        fl[i].x -= fb * dx;
                        fl[i].y -= fb * dy;
                    }
                    eb += 0.5 * kb * (rr - R.br0[t]) * (rr - R.br0[t]);
                }
                for (int t = 0; t < R.tna; t++) {
                    const int i = R.ai[t], j = R.aj[t], k = R.ak[t];
                    const double ax = minimg(q[j].x - q[i].x, L, hL), ay = minimg(q[j].y - q[i].y, L, hL);
                    const double cxv = minimg(q[k].x - q[j].x, L, hL), cyv = minimg(q[k].y - q[j].y, L, hL);
                    const double la = sqrt(ax * ax + ay * ay), lc = sqrt(cxv * cxv + cyv * cyv);
                    const double ilac = 1.0 / (la * lc);
                    const double cs = (ax * cxv + ay * cyv) * ilac;
                    // g = dU/dcos: kind 0, U = K (1 - cos); kind 1, U = K (cos - c0)^2
                    const double g = R.akind[t] == 0 ? -R.aK[t] : 2.0 * R.aK[t] * (cs - R.ac0[t]);
                    if (!ENERGY) {
                        const double dUax = g * (cxv * ilac - cs * ax / (la * la));
                        const double dUay = g * (cyv * ilac - cs * ay / (la * la));
                        const double dUcx = g * (ax * ilac - cs * cxv / (lc * lc));
                        const double dUcy = g * (ay * ilac - cs * cyv / (lc * lc));
                        fl[i].x += dUax;
                        fl[i].y += dUay;
                        fl[j].x += -dUax + dUcx;
                        fl[j].y += -dUay + dUcy;
                        fl[k].x -= dUcx;
                        fl[k].y -= dUcy;
                    }
                    eb += R.akind[t] == 0 ? R.aK[t] * (1.0 - cs) : R.aK[t] * (cs - R.ac0[t]) * (cs - R.ac0[t]);
                }
                if (!ENERGY)
                    for (int b = 0; b < n; b++) f[sl[b]] = fl[b];
            }
            if (ENERGY) ebond[s] = eb;
        }
    }
    if (!ENERGY) {
        uint64_t bits = (uint64_t)__double_as_longlong(mb);
        bits = wave_max_u64(bits);
        unsigned long long *mbw = (unsigned long long *)&rs[r].maxbond_bits;
        if ((threadIdx.x & 63) == 0 && bits > __hip_atomic_load(mbw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(mbw, (unsigned long long)bits);
    }
}

// --------------------------------------------------------------------------- precursor chemistry

// Visit the candidate beads of the cell window of bead i (cutoff <= 1) until
// fn(j) returns false; fn loads what it needs (type first, position last).
// Within one column the window's rows are contiguous cells, hence one
// contiguous slot range per periodic piece (at most two per column); the slot
// loop is unrolled so independent candidate loads overlap.
template <typename Fn>
__device__ __forceinline__ void visit_cands(const RepInfo &R, const int32_t *cs, int i, double2 p, Fn fn) {
    const int cx = cell_x(R, p.x);
    const int khi = (int)((p.y + 1.0) * R.icsy);
    const int klo = (int)((p.y - 1.0) * R.icsy + R.ncy) - R.ncy;
    for (int dc = -1; dc <= 1; dc++) {
        int col = cx + dc;
        col = col < 0 ? col + R.ncx : (col >= R.ncx ? col - R.ncx : col);
        const int cb = R.coff + col * R.ncy;
        // the window spans fewer rows than the column (L >= 8), so at most
        // one seam: rows wrapped from the top or from the bottom
        int ra[2], rb[2], np = 0;
        if (klo < 0) { ra[np] = klo + R.ncy; rb[np] = R.ncy - 1; np++; }
        ra[np] = max(klo, 0); rb[np] = min(khi, R.ncy - 1); np++;
        if (khi >= R.ncy && np < 2) { ra[np] = 0; rb[np] = khi - R.ncy; np++; }
        for (int q = 0; q < np; q++) {
            const int j0 = cs[cb + ra[q]], j1 = cs[cb + rb[q] + 1];
#pragma unroll 4
            for (int j = j0; j < j1; j++) {
                if (j == i) continue;
                if (!fn(j)) return;
            }
        }
    }
}

// FP64 minimum-image squared distance (the oracle's arithmetic)
__device__ __forceinline__ double dist2(const RepInfo &R, double2 p, double2 q) {
    const double dx = minimg(q.x - p.x, R.L, R.halfL), dy = minimg(q.y - p.y, R.L, R.halfL);
    return dx * dx + dy * dy;
}

constexpr int kChemGroup = 4;  // lanes per entry
constexpr int kChemGroupWide = 2;  // lanes per entry in a segment of more than 64 / kChemGroup entries
constexpr int kChemCls = 8;    // class bytes a lane loads per round
constexpr int kChemPos = 2;    // positions of matching candidates a lane loads per round
constexpr int kChemGrid = 1 << 16;  // wavefronts walking the segments of the check's list

// The check's list for a state that was not just placed (a chemistry check
// on a loaded state): one segment per bead block, and every bead's class.
__global__ __launch_bounds__(kBeadBlock) void k_chem_list(const RepInfo *__restrict__ ri,
                                                          const RepStep *__restrict__ rs,
                                                          const int32_t *__restrict__ bblk,
                                                          const double2 *__restrict__ xy,
                                                          const uint32_t *__restrict__ gt,
                                                          const int8_t *__restrict__ mp,
                                                          const int32_t *__restrict__ lnk, ChemList CL,
                                                          uint8_t *__restrict__ cls, ChemInj inj) {
    __shared__ int s_n;
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const RepStep &S = rs[r];
    const int base = R.off + (blockIdx.x - R.bblk0) * kBeadBlock;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (S.chem_now && loc < R.N) {
        const int s = R.off + loc;
        const uint32_t w = gt[s];
        const int m = mp[s], l = lnk[s];
        cls[s] = chem_class(w, m, l);
        const int role = chem_role(R, S, w, m, l, m < 0 && l >= 0 ? gt[l] : 0u, inj.uconv ? &inj : nullptr);
        if (role) {
            ChemEnt E;
            E.sr = s | (role << kRoleShift);
            E.w = w;
            const double2 p = xy[s];
            E.x = p.x;
            E.y = p.y;
            *chem_ent(CL, blockIdx.x, base, atomicAdd(&s_n, 1)) = E;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) CL.seg[blockIdx.x] = ChemSeg{base, s_n, r, 0};
}

// Search of a listed bead s (gid g) at p by the GRP lanes of its group
// (gl: the lane in the group; every lane returns the group's result).  conv:
// any lipid tail within r_conv.  Otherwise (a free T): the free T within
// r_form with the smallest symmetric draw pair_u01(kform) below p_form (ties:
// the smaller gid), as its slot (-1: none) and gid.  The window's candidates
// are dealt round robin; a lane reads each candidate's class byte, and the
// position (and gid) only of a candidate of the class searched for.
template <int GRP>
__device__ __forceinline__ void chem_search(const RepInfo &R, const RepStep &S, const int32_t *__restrict__ cs,
                                            const double2 *__restrict__ xy, const uint32_t *__restrict__ gt,
                                            const uint8_t *__restrict__ cls, const ChemInj &inj, int s, uint32_t g,
                                            double2 p, bool conv, int gl, bool &found, int32_t &bsel,
                                            uint32_t &bg) {
    const Win W = chem_window(R, p, conv ? R.rwc : R.rwf,
                              [&](int col, int row) { return cs[R.coff + col * R.ncy + row]; });
    const uint8_t want = conv ? CC_LIPT : CC_FREET;
    found = false;
    uint64_t bkey = ~0ull;  // (bits of u) of the best accepted draw
    bg = 0xffffffffu;
    bsel = -1;
    // candidate kk of the window (its six pieces in order)
    auto win_at = [&](int kk) {
        int j = -1;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (j < 0 && kk < W.cnt[q]) j = W.first[q] + kk;
            kk -= W.cnt[q];
        }
        return j;
    };
    // A round: the class bytes of up to kChemCls candidates of this lane
    // (loaded together), then the positions (and gids) of the ones of the
    // searched class, kChemPos at a time -- one load round each for a window
    // of up to GRP * kChemCls candidates.
    for (int k0 = gl; k0 < W.total; k0 += GRP * kChemCls) {
        uint8_t fc[kChemCls];
        int jc[kChemCls];
#pragma unroll
        for (int i = 0; i < kChemCls; i++) {
            const int kk = k0 + i * GRP;
            jc[i] = kk < W.total ? win_at(kk) : -1;
            fc[i] = cls[max(jc[i], 0)];  // unconditional: the loads in flight together
        }
#pragma unroll
        for (int i = 0; i < kChemCls; i++) fc[i] = jc[i] >= 0 && jc[i] != s ? fc[i] : 0;
        uint32_t mk = 0;
#pragma unroll
        for (int i = 0; i < kChemCls; i++) mk |= (fc[i] & want) ? 1u << i : 0u;
        while (mk) {
            int jm[kChemPos];
            double2 pc[kChemPos];
            uint32_t gc[kChemPos];
#pragma unroll
            for (int m = 0; m < kChemPos; m++) {
                jm[m] = -1;
                if (mk) {
                    jm[m] = win_at(k0 + __builtin_ctz(mk) * GRP);
                    mk &= mk - 1;
                }
            }
#pragma unroll
            for (int m = 0; m < kChemPos; m++) {  // every load of the round issued before the first is used
                pc[m] = xy[max(jm[m], 0)];
                gc[m] = gt[max(jm[m], 0)] & kGidMask;
            }
#pragma unroll
            for (int m = 0; m < kChemPos; m++) {
                if (jm[m] < 0) continue;
                if (conv) {
                    if (dist2(R, p, pc[m]) < R.rc2) found = true;
                } else if (dist2(R, p, pc[m]) < R.rf2) {
                    // draws in [0,1) order as their bits
                    const uint32_t gg = gc[m];
                    double u;
                    if (inj.fk) {  // the caller's formation uniform of the pair
                        const int64_t q = inj_find(inj.fk, inj.nf, g, gg);
                        if (q < 0) atomicOr(inj.err, ERR_INJ);
                        u = q < 0 ? 2.0 : inj.fv[q];
                    } else {
                        u = pair_u01(S.kform, g, gg);
                    }
                    const uint64_t ku = (uint64_t)__double_as_longlong(u);
                    if (u < R.p_form && (ku < bkey || (ku == bkey && gg < bg))) {
                        bkey = ku;
                        bg = gg;
                        bsel = jm[m];
                    }
                }
            }
        }
    }
    for (int o = GRP / 2; o > 0; o >>= 1) {
        found |= __shfl_xor((int)found, o, GRP) != 0;
        const uint64_t ok = __shfl_xor(bkey, o, GRP);
        const uint32_t og = __shfl_xor(bg, o, GRP);
        const int32_t oj = __shfl_xor(bsel, o, GRP);
        if (ok < bkey || (ok == bkey && og < bg)) {
            bkey = ok;
            bg = og;
            bsel = oj;
        }
    }
}

// The check (own definition), decided from the state at its start for the
// listed beads and applied at once, one wavefront per segment, kChemGroup
// lanes per entry (kChemGroupWide in a large segment; the segment header and
// its inline entries are loaded together):
//   - a break (the bond's smaller gid drew it): both link words cleared;
//   - a searching P with a lipid tail within r_conv becomes T;
//   - a searching free T with a proposal writes it into its link word
//     (chem_prop_code); the force pass that follows (or k_chem_resolve)
//     bonds the mutual ones.
// The searches read only what the check does not change -- the class bytes,
// positions and gids -- so the writes do not race with them.  Every write has
// one writer; events are sorted when read.
// The entries of segment k, GRP lanes each (E: entry lane / GRP of the
// first round when preloaded).
template <int GRP>
__device__ __forceinline__ void chem_segment(const RepInfo &R, const RepStep &S, const ChemSeg &G, int k, ChemList CL,
                                             bool preloaded, ChemEnt E, const int32_t *__restrict__ cs,
                                             const double2 *__restrict__ xy, uint32_t *__restrict__ gt,
                                             int32_t *__restrict__ lnk, const uint8_t *__restrict__ cls,
                                             Event *__restrict__ ev, unsigned long long *__restrict__ evn,
                                             int64_t evcap, int record, uint32_t *__restrict__ err,
                                             const ChemInj &inj) {
    constexpr int per = 64 / GRP;
    const int lane = (int)threadIdx.x & 63, gl = lane & (GRP - 1), ge = lane / GRP;
    for (int e = ge; e < G.cnt; e += per) {
        if (!preloaded || e != ge) E = *chem_ent(CL, k, G.base, e);
        const int s = E.sr & ((1 << kRoleShift) - 1), role = (int)((uint32_t)E.sr >> kRoleShift);
        const uint32_t g = E.w & kGidMask, ty = E.w >> kGidBits;
        if (!(role & 2)) {  // a break (of a bead in no search)
            if (gl == 0) {
                const int pq = lnk[s];
                const uint32_t gp = gt[pq] & kGidMask;
                lnk[s] = -1;
                lnk[pq] = -1;
                if (record) push_event(ev, evn, evcap, err, S.step, OC_EV_BREAK, min(g, gp), max(g, gp), G.rep);
            }
            continue;
        }
        bool found;
        int32_t b;
        uint32_t bg;
        chem_search<GRP>(R, S, cs, xy, gt, cls, inj, s, g, make_double2(E.x, E.y), ty == OC_N, gl, found, b, bg);
        if (gl != 0) continue;
        if (ty == OC_N) {
            if (found) {
                gt[s] = g | ((uint32_t)OC_T << kGidBits);
                if (record) push_event(ev, evn, evcap, err, S.step, OC_EV_CONV, g, g, G.rep);
            }
        } else if (b >= 0) {
            lnk[s] = chem_prop_code(b);
        }
    }
}

__global__ __launch_bounds__(64) void k_chem_check(const RepInfo *__restrict__ ri, const RepStep *__restrict__ rs,
                                                   const int32_t *__restrict__ cs, const double2 *__restrict__ xy,
                                                   uint32_t *__restrict__ gt, int32_t *__restrict__ lnk,
                                                   const uint8_t *__restrict__ cls, ChemList CL, int nseg,
                                                   Event *__restrict__ ev, unsigned long long *__restrict__ evn,
                                                   int64_t evcap, int record, uint32_t *__restrict__ err,
                                                   ChemInj inj) {
    constexpr int per = 64 / kChemGroup;
    static_assert(per <= kChemInl, "a wavefront's first entries are inline");
    // segments walked in the placement's XCD bands (their beads are in this
    // XCD's L2)
    const TileWalk Wk = tile_walk(nseg);
    for (int k = Wk.first; k < Wk.end; k += Wk.stride) {
        const ChemSeg G = CL.seg[k];
        // read with the header (used when the segment has at most per entries)
        const ChemEnt E = CL.inl[(size_t)k * kChemInl + (int)(threadIdx.x & 63) / kChemGroup];
        if (G.cnt <= 0) continue;
        const RepInfo &R = ri[G.rep];
        const RepStep &S = rs[G.rep];
        // a segment of more entries than a round of kChemGroup lanes each
        // takes is searched kChemGroupWide lanes an entry (fewer rounds; the
        // few such segments would otherwise end last)
        if (G.cnt <= per)
            chem_segment<kChemGroup>(R, S, G, k, CL, true, E, cs, xy, gt, lnk, cls, ev, evn, evcap, record, err, inj);
        else
            chem_segment<kChemGroupWide>(R, S, G, k, CL, false, E, cs, xy, gt, lnk, cls, ev, evn, evcap, record, err,
                                         inj);
    }
}

__device__ __forceinline__ bool R_react(const RepInfo &R) { return R.react != 0; }


// The decay pass and the food supply of the reference's chemistry check,
// cel/engine/oercel.c:1656-1753, as per-bead passes over every bead of the
// replicas whose check it is (after the catalyst reactions).
// Decay: a lipid whose first bead (mpos 0) draws hash_u01(kld, gid) < p_ld
// becomes water -- that bead's thread walks the chain and clears type, mpos
// and link of each bead (every bead belongs to one lipid: no collisions) --
// a catalyst with hash_u01(kkd, gid) < p_kd and a nutrient N with
// hash_u01(knd, gid) < p_nd become W.  All decisions read the state before
// the pass.
__global__ __launch_bounds__(kBeadBlock) void k_rdecay(const RepInfo *__restrict__ ri, const RepStep *__restrict__ rs,
                                                       const int32_t *__restrict__ bblk, const double2 *__restrict__ xy,
                                                       uint32_t *__restrict__ gt, int8_t *__restrict__ mp,
                                                       int32_t *__restrict__ lnk, Event *__restrict__ ev,
                                                       unsigned long long *__restrict__ evn, int64_t evcap, int record,
                                                       uint32_t *__restrict__ err) {
    const int r = bblk[blockIdx.x];
    const RepStep &S = rs[r];
    // with the catalyst reactions, a check after a step that moved a bead by
    // the chemistry margin is reported (ERR_DISP, k_react_prep) and changes
    // nothing: no reaction, decay or feed
    if (!S.rchem_now || (S.disp && R_react(ri[r]))) return;
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const uint32_t w = gt[s], g = w & kGidMask, ty = w >> kGidBits;
    const int m = mp[s];
    if (m == 0 && R.p_ld > 0 && hash_u01(S.kld, g) < R.p_ld) {
        const double2 p = xy[s];
        int c = s, n = 0;
        while (c >= 0 && n < R.mlen) {
            const int nx = lnk[c];
            gt[c] = (gt[c] & kGidMask) | ((uint32_t)OC_W << kGidBits);
            mp[c] = -1;
            lnk[c] = -1;
            c = nx;
            n++;
        }
        if (record) push_event(ev, evn, evcap, err, S.step, OC_EV_DLIP, g, g, r, p.x, p.y);
    } else if ((ty == OC_KA || ty == OC_KB) && R.p_kd > 0 && hash_u01(S.kkd, g) < R.p_kd) {
        gt[s] = g | ((uint32_t)OC_W << kGidBits);
        if (record) {
            const double2 p = xy[s];
            push_event(ev, evn, evcap, err, S.step, OC_EV_DK, g, ty == OC_KB ? 1u : 0u, r, p.x, p.y);
        }
    } else if (ty == OC_N && R.p_nd > 0 && hash_u01(S.knd, g) < R.p_nd) {
        gt[s] = g | ((uint32_t)OC_W << kGidBits);
        if (record) {
            const double2 p = xy[s];
            push_event(ev, evn, evcap, err, S.step, OC_EV_DN, g, g, r, p.x, p.y);
        }
    }
}

// inside the fixed vent region?  cel/engine/oercel.c:1702-1709
__device__ __forceinline__ bool in_vent(const RepInfo &R, double x, double y) {
    if (R.vent_shape == OC_VENT_DISK) {
        const double dx = minimg(x - R.vent_x, R.L, R.halfL), dy = minimg(y - R.vent_y, R.L, R.halfL);
        return dx * dx + dy * dy < R.vent_r2;
    }
    if (R.vent_shape == OC_VENT_VSTRIP) return fabs(minimg(x - R.vent_x, R.L, R.halfL)) < R.vent_hw;
    return fabs(minimg(y - R.vent_y, R.L, R.halfL)) < R.vent_hw;
}

// feed candidate: a W (inside the vent region for feed = vent) whose draw
// hash_u01(kfeed, gid) < p_feed
__device__ __forceinline__ bool feed_candidate(const RepInfo &R, const RepStep &S, uint32_t w, double2 p) {
    return (w >> kGidBits) == OC_W && hash_u01(S.kfeed, w & kGidMask) < R.p_feed &&
           (R.feed_mode != OC_FEED_VENT || in_vent(R, p.x, p.y));
}
// feed-cap key of a candidate: the high half of hash_u64(kcap, gid), the gid
// below (unique), cel/engine_gpu/oercel_gpu.cu:1963
__device__ __forceinline__ uint64_t feed_key(const RepStep &S, uint32_t g) {
    return (mix13(S.kcap ^ ((uint64_t)g * 0xd1b54a32d192ed03ULL)) & 0xFFFFFFFF00000000ULL) | g;
}

// Food supply after the decay pass (so a bead that decayed to W may be fed in
// the same check): every candidate becomes N -- with a feed cap, the ones
// whose key is at most the threshold k_rfeed_thr left (none when it keeps
// none).
__global__ __launch_bounds__(kBeadBlock) void k_rfeed(const RepInfo *__restrict__ ri, const RepStep *__restrict__ rs,
                                                      const int32_t *__restrict__ bblk, const double2 *__restrict__ xy,
                                                      uint32_t *__restrict__ gt, Event *__restrict__ ev,
                                                      unsigned long long *__restrict__ evn, int64_t evcap, int record,
                                                      uint32_t *__restrict__ err) {
    const int r = bblk[blockIdx.x];
    const RepStep &S = rs[r];
    if (!S.rchem_now || (S.disp && R_react(ri[r]))) return;  // see k_rdecay
    const RepInfo &R = ri[r];
    if (R.feed_mode == OC_FEED_NONE || !(R.p_feed > 0)) return;
    if (R.fcap_on && S.feed_add == 0) return;
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    const uint32_t w = gt[s], g = w & kGidMask;
    if ((w >> kGidBits) != OC_W) return;
    const double2 p = xy[s];
    if (!feed_candidate(R, S, w, p)) return;
    if (R.fcap_on && feed_key(S, g) > S.feed_thr) return;
    gt[s] = g | ((uint32_t)OC_N << kGidBits);
    if (record) push_event(ev, evn, evcap, err, S.step, OC_EV_FEED, g, g, r, p.x, p.y);
}

// A listed free T's pending proposal settled (the rule the force pass applies
// to each bead, chem_settle): the bond forms when the proposal is mutual.
// For a check on its own (oc_chem), after k_chem_check.
__global__ __launch_bounds__(64) void k_chem_resolve(const RepStep *__restrict__ rs, const uint32_t *__restrict__ gt,
                                                     int32_t *__restrict__ lnk, ChemList CL, int nseg,
                                                     Event *__restrict__ ev, unsigned long long *__restrict__ evn,
                                                     int64_t evcap, int record, uint32_t *__restrict__ err) {
    const TileWalk Wk = tile_walk(nseg);
    for (int k = Wk.first; k < Wk.end; k += Wk.stride) {
        const ChemSeg G = CL.seg[k];
        if (G.cnt <= 0) continue;
        const RepStep &S = rs[G.rep];
        for (int e = threadIdx.x; e < G.cnt; e += 64) {
            const ChemEnt &E = *chem_ent(CL, k, G.base, e);
            const int s = E.sr & ((1 << kRoleShift) - 1);
            const int l = lnk[s];
            if (l >= -1) continue;
            const uint32_t g = E.w & kGidMask;
            const int n = chem_settle(s, l, lnk);
            lnk[s] = n;
            if (n >= 0 && record) {
                const uint32_t gb = gt[n] & kGidMask;
                if (g < gb) push_event(ev, evn, evcap, err, S.step, OC_EV_FORM, g, gb, G.rep);
            }
        }
    }
}

// --------------------------------------------------------------------------- reductions

constexpr int kNStat = 16;  // k2, px, py, epair, ebond, count[7], lipids, bonds, stir, virial

// Fixed-partition FP64 sums: chunk b of a replica covers beads
// [b*kRedChunk, (b+1)*kRedChunk); a thread sums a strided subset in order,
// then a fixed tree.  The result depends only on the slot order.
__global__ __launch_bounds__(256) void k_stats_partial(const RepInfo *__restrict__ ri, const int32_t *__restrict__ rblk,
                                                       const double2 *__restrict__ v, const double2 *__restrict__ f,
                                                       const uint32_t *__restrict__ gt, const int8_t *__restrict__ mp,
                                                       const int32_t *__restrict__ lnk,
                                                       const double *__restrict__ epot,
                                                       const double *__restrict__ ebond,
                                                       const double *__restrict__ evir, double *__restrict__ part) {
    __shared__ double red[kNStat][256];
    const int r = rblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int b = blockIdx.x - R.rblk0;
    double acc[kNStat];
    for (int k = 0; k < kNStat; k++) acc[k] = 0.0;
    for (int q = 0; q < kRedChunk / 256; q++) {
        const int loc = b * kRedChunk + q * 256 + threadIdx.x;
        if (loc >= R.N) break;
        const int s = R.off + loc;
        const double2 vv = v[s], ff = f[s];
        const double vx = vv.x + R.hdt * ff.x, vy = vv.y + R.hdt * ff.y;
        acc[0] += vx * vx + vy * vy;
        acc[1] += vx;
        acc[2] += vy;
        acc[3] += epot[s];
        acc[4] += ebond[s];
        const uint32_t ty = gt[s] >> kGidBits;
        for (int k = 0; k < 7; k++) acc[5 + k] += ty == (uint32_t)k ? 1.0 : 0.0;
        acc[12] += mp[s] == 0 ? 1.0 : 0.0;
        const int l = lnk[s];
        acc[13] += (mp[s] < 0 && l >= 0 && (gt[s] & kGidMask) < (gt[l] & kGidMask)) ? 1.0 : 0.0;
        acc[15] += evir[s];
    }
    for (int k = 0; k < kNStat; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < kNStat; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kNStat) part[(size_t)blockIdx.x * kNStat + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_stats_final(const RepInfo *__restrict__ ri, const double *__restrict__ part,
                                                     double *__restrict__ out) {
    __shared__ double red[256];
    const RepInfo &R = ri[blockIdx.x];
    const int nb = (R.N + kRedChunk - 1) / kRedChunk;
    for (int k = 0; k < kNStat; k++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) s += part[(size_t)(R.rblk0 + b) * kNStat + k];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * kNStat + k] = red[0];
        __syncthreads();
    }
}

// stirring body force f_x = A sin(2 pi n y / L) and its per-replica sum
// (force_init, cel/engine/oercel.c:1161-1196)
__global__ __launch_bounds__(256) void k_stir_partial(const RepInfo *__restrict__ ri, const int32_t *__restrict__ rblk,
                                                      const double2 *__restrict__ xy, double *__restrict__ stir,
                                                      double *__restrict__ part) {
    __shared__ double red[256];
    const int r = rblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int b = blockIdx.x - R.rblk0;
    double acc = 0.0;
    for (int q = 0; q < kRedChunk / 256; q++) {
        const int loc = b * kRedChunk + q * 256 + threadIdx.x;
        if (loc >= R.N) break;
        const int s = R.off + loc;
        const double fs = R.has_stir ? R.stir_A * sin(2.0 * M_PI * (R.stir_kn * xy[s].y)) : 0.0;
        stir[s] = fs;
        acc += fs;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_stir_final(const RepInfo *__restrict__ ri, RepStep *__restrict__ rs,
                                                    const double *__restrict__ part) {
    __shared__ double red[256];
    const RepInfo &R = ri[blockIdx.x];
    const int nb = (R.N + kRedChunk - 1) / kRedChunk;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[R.rblk0 + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) rs[blockIdx.x].stir_mean = red[0] / R.N;
}

// vh = v - dt/2 f after the first forces of a fresh start (cel/engine/oercel.c:2207-2211)
__global__ __launch_bounds__(kBeadBlock) void k_half_kick(const RepInfo *__restrict__ ri,
                                                          const int32_t *__restrict__ bblk, double2 *__restrict__ v,
                                                          const double2 *__restrict__ f) {
    const int r = bblk[blockIdx.x];
    const RepInfo &R = ri[r];
    const int loc = (blockIdx.x - R.bblk0) * kBeadBlock + threadIdx.x;
    if (loc >= R.N) return;
    const int s = R.off + loc;
    double2 vv = v[s];
    const double2 ff = f[s];
    vv.x -= 0.5 * R.dt * ff.x;
    vv.y -= 0.5 * R.dt * ff.y;
    v[s] = vv;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace oc
