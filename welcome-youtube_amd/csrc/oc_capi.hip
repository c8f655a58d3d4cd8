// oc_capi.hip -- host side of the C ABI declared in include/oercel_mi.h:
// engine handle (device buffers, block tables, pair tiles), state load/store in
// gid order, the step launch sequence, hipGraph capture, stats and events.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "oc_adjoint.hpp"
#include "oc_mol.hpp"
#include "oc_pairfx.hpp"
#include "oc_react.hpp"

using namespace oc;

#define OC_API extern "C" __attribute__((visibility("default")))

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(OC_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct oc_engine {
    int nrep = 0;
    std::vector<oc_params> prm;
    std::vector<RepInfo> ri;
    int64_t Ntot = 0, Ctot = 0;
    std::vector<int32_t> bblk, rblk;
    std::vector<int2> cols;  // (replica, column) of every cell column
    std::vector<Tile> tiles;   // pair-force tiles
    std::vector<Tile> stiles;  // sort tiles (rows of new cells, one thread each)
    int pair_block = 256, tile_beads = 0;
    int pair_variant = 1;  // k_pair: 0 single pass, 1 per-lane two phases (default), 2 cross-lane pair queue
    int per_xcd = 0;
    // device
    RepInfo *d_ri = nullptr;
    RepStep *d_rs = nullptr;
    int32_t *d_bblk = nullptr, *d_rblk = nullptr;
    std::vector<int2> trng;     // per replica: first pair tile, tile count
    int2 *d_trng = nullptr;
    int2 *d_brng = nullptr;     // per replica: first bead-block partial of the reverse step, count
    Tile *d_tiles = nullptr, *d_stiles = nullptr;
    StateP st[2];
    int cur = 0;
    double2 *d_f = nullptr;
    int32_t *d_key = nullptr, *d_cs[2] = {nullptr, nullptr};
    // sort: per tile bead count, first new slot, per-tile bead list; per bead
    // its (tile, rank); per cell its rank offset within the tile
    int32_t *d_tcount = nullptr, *d_toff = nullptr, *d_tsrc = nullptr, *d_tslot = nullptr, *d_locst = nullptr;
    uint32_t *d_moved = nullptr;  // a bead moved by more than one cell this step
    int32_t *rec_src = nullptr;   // when set, the sort writes src (new slot -> old slot) there
    int32_t *d_tpart = nullptr;  // chunk totals of the tile-count scan
    int2 *d_rmax = nullptr;      // per replica: largest cell displacement of a bead that moved > 1 cell
    bool keys_fresh = false;     // d_key holds the keys of the current state (written by the force kernel)
    double *d_stir = nullptr, *d_epot = nullptr, *d_evir = nullptr, *d_ebond = nullptr, *d_part = nullptr,
           *d_statout = nullptr;
    uint8_t *d_ccls = nullptr;  // chemistry class byte of every bead at a check
    // precursor chemistry: the check's list of beads with a role, in segments
    // (one per placement tile, or per bead block of k_chem_list)
    ChemEnt *d_cinl = nullptr, *d_cspill = nullptr;  // the check's list: inline entries, spill by slot
    ChemSeg *d_cseg = nullptr;
    int cseg_n = 0;  // segments of the last list
    Event *d_ev = nullptr;
    unsigned long long *d_evn = nullptr;
    int64_t evcap = 0;
    int record = 1;
    uint32_t *d_err = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // with the reference chemistry: one graph of n_chem steps from a check
    // (chemistry kernels only in its last step), per starting buffer
    hipGraphExec_t pexec[2] = {nullptr, nullptr};
    hipGraph_t pgraph[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    // oc_launch_pair: the last count's launches as one graph, per state buffer
    hipGraphExec_t texec[2] = {nullptr, nullptr};
    hipGraph_t tgraph[2] = {nullptr, nullptr};
    int32_t treps[2] = {0, 0};
    bool have_state = false, any_stir = false, any_chem = false, any_angle = false;
    bool any_tmpl = false;  // some replica has branched lipids (k_mol; the force kernel leaves no keys)
    bool any_rchem = false; // some replica runs the reference chemistry: the step integrates, runs it, then sorts
    bool any_react = false, any_fcap = false;  // catalyst reactions; feed cap
    bool nokeys = false;    // the force kernel leaves no cell keys (branched lipids or the reference chemistry)
    int32_t *d_cat = nullptr, *d_catn = nullptr;  // active catalysts per replica (kCatCap each), counts
    int2 *d_ccell = nullptr;                      // their cells
    uint8_t *d_cdone = nullptr;                   // reacted in k_react_free
    GBead *d_gs = nullptr;                        // per-wavefront gather lists of the reactions
    int gcap = 0;                                 // beads per gather list
    int32_t *d_npart = nullptr;                   // feed cap: N count per bead block
    uint64_t *d_fkey = nullptr;                   // feed-cap candidate keys (per replica at its first slot)
    int2 *d_fcnt = nullptr;                       // feed cap: (number of N, candidates) per replica
    std::vector<std::vector<oc_event>> evq;
    std::vector<void *> allocs;
    bool events_dropped = false;
    struct Adj *adj = nullptr;
};

template <class T>
static int dalloc(oc_engine *h, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(OC_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

static int grid_of(int64_t n, int b) { return (int)((n + b - 1) / b); }

// --------------------------------------------------------------------------- replica constants

static int make_repinfo(const oc_params &p, int32_t N, RepInfo &R) {
    memset(&R, 0, sizeof R);
    if (!(p.L >= 8.0)) return fail(OC_EINVAL, "L must be >= 8");
    if (p.nsub < 1 || p.nsub > 8) return fail(OC_EINVAL, "nsub must be 1..8");
    if (p.dt <= 0) return fail(OC_EINVAL, "dt must be > 0");
    if (N < 2 || N >= (1 << kGidBits)) return fail(OC_EINVAL, "N must be in [2, 2^29)");
    MolTmpl M;
    const char *why = "";
    if (!build_template(p, M, &why)) return fail(OC_EINVAL, why);
    if (p.chem && (p.r_conv > 1.0 || p.r_form > 1.0)) return fail(OC_EINVAL, "r_conv and r_form must be <= 1");
    R.N = N;
    R.ncx = (int)floor(p.L);
    R.ncy = (int)floor(p.L) * p.nsub;
    R.ncell = R.ncx * R.ncy;
    R.mlen = M.mlen;
    R.nH = M.nH;
    R.arch = M.arch;
    R.tnb = M.nb;
    R.tna = M.na;
    for (int q = 0; q < M.nb; q++) {
        R.bi[q] = M.bi[q];
        R.bj[q] = M.bj[q];
        R.br0[q] = M.br0[q];
    }
    for (int q = 0; q < M.na; q++) {
        R.ai[q] = M.ai[q];
        R.aj[q] = M.aj[q];
        R.ak[q] = M.ak[q];
        R.akind[q] = M.akind[q];
        R.aK[q] = M.aK[q];
        R.ac0[q] = M.ac0[q];
    }
    R.chem = p.chem;
    R.n_chem = p.n_chem < 1 ? 1 : p.n_chem;
    R.has_stir = p.stir_A != 0.0;
    R.L = p.L;
    R.halfL = 0.5 * p.L;
    R.icsx = R.ncx / p.L;
    R.icsy = R.ncy / p.L;
    R.csx = p.L / R.ncx;
    R.csy = p.L / R.ncy;
    R.dt = p.dt;
    R.dt2 = p.dt * p.dt;
    R.cvt = (0.5 + p.lambda) * p.dt;
    R.hdt = 0.5 * p.dt;
    R.gamma = p.gamma;
    R.kT = p.kT;
    R.kb = p.k_bond;
    R.r0 = p.r0;
    R.ka = p.k_angle;
    R.stir_A = p.stir_A;
    R.stir_kn = (double)p.stir_n / p.L;
    R.kT_amp = p.kT_amp;
    R.kT_period = p.kT_period;
    R.t_anchor = p.t_anchor;
    R.step_anchor = p.step_anchor;
    R.warmup = p.warmup;
    R.warm_end = p.warm_end;
    R.seed = p.seed;
    R.p_conv = p.p_conv;
    R.rc2 = p.r_conv * p.r_conv;
    R.p_form = p.p_form;
    R.rf2 = p.r_form * p.r_form;
    // the rows a search visits: within r + 1e-6 of the bead (the margin covers
    // the FP64 rounding of the row bounds; beads further off fail the test)
    R.rwc = p.r_conv + 1e-6;
    R.rwf = p.r_form + 1e-6;
    R.rchem = p.rchem;
    R.feed_mode = p.feed_mode;
    R.vent_shape = p.vent_shape;
    R.p_ld = p.p_ld;
    R.p_nd = p.p_nd;
    R.p_feed = p.p_feed;
    R.vent_x = p.vent_x;
    R.vent_y = p.vent_y;
    R.vent_r2 = p.vent_r * p.vent_r;
    R.vent_hw = 0.5 * p.vent_w;
    if (p.rchem && (p.feed_mode < 0 || p.feed_mode > 2 || p.vent_shape < 0 || p.vent_shape > 2))
        return fail(OC_EINVAL, "bad feed mode or vent shape");
    R.p_break = p.p_break;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) R.a[8 * i + j] = (i < OC_NTYPES && j < OC_NTYPES) ? p.a[i][j] : 0.0;
    // catalyst chemistry and feed cap (cel/engine_gpu/oercel_gpu.cu:1773-1797)
    R.p_syn[0] = p.p_syn_a;
    R.p_syn[1] = p.p_syn_b;
    R.p_rep[0] = p.p_rep_a;
    R.p_rep[1] = p.p_rep_b;
    R.mu = p.mu;
    R.p_kd = p.p_kd;
    R.r2cat = p.r_cat * p.r_cat;
    R.syn_tries = std::max(p.syn_tries, 1);
    R.syn_relax = std::max(p.syn_relax, 0);
    R.nsynN = p.syn_nN > 0 ? p.syn_nN : M.mlen;
    R.react = p.rchem && (p.p_syn_a > 0 || p.p_syn_b > 0 || p.p_rep_a > 0 || p.p_rep_b > 0);
    for (int b = 0; b < M.mlen; b++) {
        R.tpx[b] = M.px[b];
        R.tpy[b] = M.py[b];
    }
    if (R.react) {
        if (p.syn_nN < 0 || p.syn_nN > M.mlen) return fail(OC_EINVAL, "syn_nN must be 0..lipid length");
        if (!(p.r_cat > 0)) return fail(OC_EINVAL, "r_cat must be > 0");
        const double margin = 0.5, reach = 1.0 + 0.05 * R.syn_relax;
        R.rwx = (int)ceil((p.r_cat + margin) * R.icsx);
        R.rwy = (int)ceil((p.r_cat + margin) * R.icsy);
        R.owx = (int)ceil((1.0 + margin) * R.icsx);
        R.owy = (int)ceil((1.0 + margin) * R.icsy);
        R.lwx = (int)ceil((reach + margin) * R.icsx);
        R.lwy = (int)ceil((reach + margin) * R.icsy);
        if (2 * std::max({R.rwx, R.owx, R.lwx}) + 1 > R.ncx || 2 * std::max({R.rwy, R.owy, R.lwy}) + 1 > R.ncy)
            return fail(OC_EINVAL, "box too small for the catalyst chemistry's search windows (r_cat, syn_relax)");
        // cells a reaction writes: its r_cat window; cells it reads: that
        // window, and the overlap and relaxation windows around the new
        // lipid's beads, which lie within r_cat + E (E: the template's extent
        // from its centroid) of the catalyst
        double E = 0.0;
        for (int b = 0; b < M.mlen; b++) E = std::max(E, std::hypot(M.px[b], M.py[b]));
        const double d = p.r_cat + E + 1e-9;
        const int rdx = std::max(R.rwx, (int)ceil(d * R.icsx) + 1 + std::max(R.owx, R.lwx));
        const int rdy = std::max(R.rwy, (int)ceil(d * R.icsy) + 1 + std::max(R.owy, R.lwy));
        R.cfx = R.rwx + rdx;
        R.cfy = R.rwy + rdy;
        R.rdx = rdx;
        R.rdy = rdy;
        const double rg = d + std::max(1.0, reach) + 1e-6, ro = d + 1.0 + 1e-6;
        R.rg2 = rg * rg;
        R.ro2 = ro * ro;
    }
    R.fcap_on = p.rchem && p.feed_cap > 0 && p.feed_mode != OC_FEED_NONE && p.p_feed > 0;
    if (p.rchem && (p.feed_cap < 0 || p.feed_cap >= 1)) return fail(OC_EINVAL, "feed_cap must be in [0, 1)");
    R.fcap_n = (int32_t)floor(p.feed_cap * (double)N + 1e-9);
    return 0;
}

// pair tiles: rows per tile from the expected beads per row
static int build_tiles(oc_engine *h) {
    h->tiles.clear();
    const int target = h->tile_beads > 0 ? h->tile_beads : (h->pair_block * 9) / 10;
    h->trng.assign(h->nrep, make_int2(0, 0));
    for (int r = 0; r < h->nrep; r++) {
        const RepInfo &R = h->ri[r];
        h->trng[r].x = (int)h->tiles.size();
        const double per_row = (double)R.N / ((double)R.ncx * R.ncy);
        const int M = (int)ceil(R.icsy) + 1;
        int rows = (int)lround(target / std::max(per_row, 1e-9));
        rows = std::max(1, std::min(rows, R.ncy));
        rows = std::min(rows, pair_max_rows(h->pair_block) - 2 * M - 1);
        if (rows < 1) return fail(OC_EINVAL, "nsub too large for the pair tile buffer");
        const int nt = (R.ncy + rows - 1) / rows;
        for (int cx = 0; cx < R.ncx; cx++)
            for (int t = 0; t < nt; t++) {
                Tile T;
                T.rep = r;
                T.cx = cx;
                T.k0 = (int)(((int64_t)t * R.ncy) / nt);
                T.k1 = (int)(((int64_t)(t + 1) * R.ncy) / nt);
                h->tiles.push_back(T);
            }
        h->trng[r].y = (int)h->tiles.size() - h->trng[r].x;
    }
    h->per_xcd = (int)((h->tiles.size() + 7) / 8);
    if (h->d_tiles) {
        (void)hipFree(h->d_tiles);
        h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void *)h->d_tiles), h->allocs.end());
        h->d_tiles = nullptr;
    }
    if (int rc = dalloc(h, &h->d_tiles, h->tiles.size())) return rc;
    HIPCHK(hipMemcpy(h->d_tiles, h->tiles.data(), h->tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (!h->d_trng)
        if (int rc = dalloc(h, &h->d_trng, h->nrep)) return rc;
    HIPCHK(hipMemcpy(h->d_trng, h->trng.data(), h->nrep * sizeof(int2), hipMemcpyHostToDevice));
    return 0;
}

static void drop_graphs(oc_engine *h) {
    for (int k = 0; k < 2; k++) {
        if (h->gexec[k]) (void)hipGraphExecDestroy(h->gexec[k]);
        if (h->graph[k]) (void)hipGraphDestroy(h->graph[k]);
        h->gexec[k] = nullptr;
        h->graph[k] = nullptr;
        if (h->pexec[k]) (void)hipGraphExecDestroy(h->pexec[k]);
        if (h->pgraph[k]) (void)hipGraphDestroy(h->pgraph[k]);
        h->pexec[k] = nullptr;
        h->pgraph[k] = nullptr;
        if (h->texec[k]) (void)hipGraphExecDestroy(h->texec[k]);
        if (h->tgraph[k]) (void)hipGraphDestroy(h->tgraph[k]);
        h->texec[k] = nullptr;
        h->tgraph[k] = nullptr;
        h->treps[k] = 0;
    }
}

// --------------------------------------------------------------------------- launches

// pair forces (+ bonded terms when bonded, in the same pass), or the pair
// energies when energy
static void launch_pair(oc_engine *h, bool energy, bool bonded = false, const InjTheta *inj = nullptr) {
    const int c = h->cur;
    PairArgs A;
    A.ri = h->d_ri;
    A.rs = h->d_rs;
    A.rsw = h->d_rs;
    A.mp = h->st[c].mp;
    A.lnk = h->st[c].lnk;
    A.prv = h->st[c].prv;
    A.bonded = bonded && !energy;
    A.tiles = h->d_tiles;
    A.ntiles = (int)h->tiles.size();
    A.per_xcd = h->per_xcd;
    A.cs = h->d_cs[c];
    A.xy = h->st[c].xy;
    A.vt = h->st[c].vt;
    A.gt = h->st[c].gt;
    A.stir = h->any_stir ? h->d_stir : nullptr;
    A.f = h->d_f;
    A.epot = h->d_epot;
    A.evir = h->d_evir;
    A.err = h->d_err;
    // forces also leave the next step's cell keys
    A.v = h->st[c].v;
    A.key = energy || h->nokeys ? nullptr : h->d_key;
    A.rmax = h->d_rmax;
    A.moved = h->d_moved;
    if (!energy) h->keys_fresh = !h->nokeys;
    A.lnkw = h->st[c].lnk;
    A.ev = h->d_ev;
    A.evn = h->d_evn;
    A.evcap = h->evcap;
    A.record = h->record;
    const dim3 grid(8 * h->per_xcd);
    // CAP: staged candidates (3.5x the block; 812 at 256 threads, so that seven
    // workgroups fit a CU's LDS: tiles of ~230 beads stage ~710 at rho 3, and
    // a tile with more runs on global memory); the variant picks the single
    // pass, the per-lane two phases or the cross-lane pair queue
    const int V = h->pair_variant;
#define OC_PAIR(B, C, CL)                                                                        \
    do {                                                                                         \
        if (energy) k_pair<B, C, CL, true, false><<<grid, B, 0, h->stream>>>(A);                  \
        else if (h->any_angle) k_pair<B, C, CL, false, true><<<grid, B, 0, h->stream>>>(A);       \
        else k_pair<B, C, CL, false, false><<<grid, B, 0, h->stream>>>(A);                        \
    } while (0)
#define OC_PAIRFX(B, C)                                                                          \
    do {                                                                                         \
        if (h->any_angle) k_pair_fx<B, C, kPairList, true><<<grid, B, 0, h->stream>>>(A);         \
        else k_pair_fx<B, C, kPairList, false><<<grid, B, 0, h->stream>>>(A);                     \
    } while (0)
    if (inj) {  // caller-supplied pair noise: the single pass
        A.inj = *inj;
#define OC_PAIRINJ(B, C)                                                                         \
    do {                                                                                         \
        if (h->any_angle) k_pair<B, C, 0, false, true, true><<<grid, B, 0, h->stream>>>(A);       \
        else k_pair<B, C, 0, false, false, true><<<grid, B, 0, h->stream>>>(A);                   \
    } while (0)
        switch (h->pair_block) {
        case 64: OC_PAIRINJ(64, 224); break;
        case 128: OC_PAIRINJ(128, 448); break;
        default: OC_PAIRINJ(256, 812); break;
        }
#undef OC_PAIRINJ
        return;
    }
    if (V == 3 && !energy) {
        switch (h->pair_block) {
        case 64: OC_PAIRFX(64, 224); break;
        case 128: OC_PAIRFX(128, 448); break;
        default: OC_PAIRFX(256, 896); break;
        }
        return;
    }
#undef OC_PAIRFX
    switch (h->pair_block) {
    case 64:
        if (V == 2) OC_PAIR(64, 224, 2); else if (V == 1) OC_PAIR(64, 224, 1); else OC_PAIR(64, 224, 0);
        break;
    case 128:
        if (V == 2) OC_PAIR(128, 448, 2); else if (V == 1) OC_PAIR(128, 448, 1); else OC_PAIR(128, 448, 0);
        break;
    default:
        if (V == 2) OC_PAIR(256, 812, 2); else if (V == 1) OC_PAIR(256, 812, 1); else OC_PAIR(256, 812, 0);
        break;
    }
#undef OC_PAIR
}

static void launch_stir(oc_engine *h) {
    if (!h->any_stir) return;
    const int c = h->cur;
    const int nrb = (int)h->rblk.size();
    k_stir_partial<<<nrb, 256, 0, h->stream>>>(h->d_ri, h->d_rblk, h->st[c].xy, h->d_stir, h->d_part);
    k_stir_final<<<h->nrep, 256, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_part);
}

static void launch_forces(oc_engine *h, bool with_bonded, const InjTheta *inj = nullptr) {
    launch_stir(h);
    launch_pair(h, false, with_bonded, inj);
    if (with_bonded && h->any_tmpl) {
        const int c = h->cur;
        k_mol<false><<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, h->st[c].xy,
                                                                       h->st[c].mp, h->st[c].lnk, h->d_f, nullptr,
                                                                       h->d_err);
    }
}

static void launch_build_prv(oc_engine *h) {
    const int c = h->cur;
    k_build_prv<<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, h->st[c].mp, h->st[c].lnk,
                                                                  h->st[c].prv);
}

// The reference chemistry of the replicas whose check it is: catalyst
// reactions, decay, food supply (with its cap), in the reference's order.
// list: list the active catalysts here (a step's k_integ lists them).
static void launch_rchem(oc_engine *h, bool list) {
    if (!h->any_rchem) return;
    const int c = h->cur, nb = (int)h->bblk.size();
    StateP &S = h->st[c];
    if (h->any_react) {
        if (list) k_react_list<<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, S.gt, h->d_cat, h->d_catn);
        ReactBufs B;
        B.xy = S.xy;
        B.v = S.v;
        B.vt = S.vt;
        B.gt = S.gt;
        B.mp = S.mp;
        B.lnk = S.lnk;
        B.prv = S.prv;
        B.cs = h->d_cs[c];
        B.key = h->d_key;
        B.rmax = h->d_rmax;
        B.moved = h->d_moved;
        B.cat = h->d_cat;
        B.catn = h->d_catn;
        B.ccell = h->d_ccell;
        B.cdone = h->d_cdone;
        B.gs = h->d_gs;
        B.gcap = h->gcap;
        B.ev = h->d_ev;
        B.evn = h->d_evn;
        B.evcap = h->evcap;
        B.record = h->record;
        B.err = h->d_err;
        k_react_prep<<<h->nrep, kPrepThr, 0, h->stream>>>(h->d_ri, h->d_rs, B);
        k_react_free<<<kFreeBlocks, 256, 0, h->stream>>>(h->d_ri, h->d_rs, B, h->nrep);
        k_react_rest<<<h->nrep, kRestThr, 0, h->stream>>>(h->d_ri, h->d_rs, B, kFreeBlocks * 4);
    }
    k_rdecay<<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, S.xy, S.gt, S.mp, S.lnk, h->d_ev, h->d_evn,
                                               h->evcap, h->record, h->d_err);
    if (h->any_fcap) {
        k_rfeed_count<<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, S.xy, S.gt, h->d_fkey, h->d_fcnt,
                                                        h->d_npart);
        k_rfeed_thr<<<h->nrep, 256, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_fkey, h->d_fcnt, h->d_npart);
    }
    k_rfeed<<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, S.xy, S.gt, h->d_ev, h->d_evn, h->evcap,
                                              h->record, h->d_err);
}

// The check's list of the engine's buffers.
static ChemList chem_list(oc_engine *h) { return ChemList{h->d_cseg, h->d_cinl, h->d_cspill}; }

// the precursor chemistry (after the sort, on the fresh cell list): one pass
// decides and applies the check; a formation proposal waits in the proposer's
// link word for the force pass of the step (settled: settle = false), or is
// settled by k_chem_resolve (a check on its own).
static void launch_pchem(oc_engine *h, bool listed, const ChemInj *inj = nullptr, bool settle = true) {
    if (!h->any_chem) return;
    const int c = h->cur;
    const ChemInj ci = inj ? *inj : ChemInj{nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    const ChemList CL = chem_list(h);
    if (!listed) {
        // a state that was not just placed: list the roles (and classes) of every bead
        h->cseg_n = (int)h->bblk.size();
        k_chem_list<<<h->cseg_n, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, h->st[c].xy, h->st[c].gt,
                                                             h->st[c].mp, h->st[c].lnk, CL, h->d_ccls, ci);
    }
    // one wavefront per segment (XCD bands: a multiple of 8 workgroups)
    const int ng = h->cseg_n >= 8 ? (std::min(h->cseg_n, kChemGrid) + 7) & ~7 : h->cseg_n;
    k_chem_check<<<ng, 64, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_cs[c], h->st[c].xy, h->st[c].gt, h->st[c].lnk,
                                           h->d_ccls, CL, h->cseg_n, h->d_ev, h->d_evn, h->evcap, h->record,
                                           h->d_err, ci);
    if (settle)
        k_chem_resolve<<<ng, 64, 0, h->stream>>>(h->d_rs, h->st[c].gt, h->st[c].lnk, CL, h->cseg_n, h->d_ev,
                                                 h->d_evn, h->evcap, h->record, h->d_err);
}

// one chemistry check on the current (sorted) state; the reference
// chemistry's syntheses move beads, so its state is sorted again
static void launch_sort(oc_engine *h, bool integ, int step_delta, bool keys_ready);
static void launch_keys(oc_engine *h, bool integ);
static void launch_chem(oc_engine *h) {
    if (h->any_rchem) {
        launch_keys(h, false);
        launch_rchem(h, true);
        launch_sort(h, false, 0, true);
    }
    launch_pchem(h, h->any_rchem);  // the placement lists the roles
}

// sort (with integration when integ): keys, tile sort, tile offsets,
// placement; flips cur.  step_delta advances every replica's step counter and
// per-step values (in k_tscan; no later kernel of the sort reads them).
static void launch_keys(oc_engine *h, bool integ) {
    const int c = h->cur, nb = (int)h->bblk.size();
    if (integ)
        k_keys<true><<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, h->st[c].xy, h->st[c].v, h->d_f, h->d_key,
                                                       h->d_rmax, h->d_moved);
    else
        k_keys<false><<<nb, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, h->st[c].xy, h->st[c].v, h->d_f,
                                                        h->d_key, h->d_rmax, h->d_moved);
}

static void launch_sort(oc_engine *h, bool integ, int step_delta, bool keys_ready = false) {
    const int c = h->cur, n = c ^ 1;
    const int nt = (int)h->stiles.size();
    // the force kernel left the keys of this state's integrate unless the
    // state changed since (keys_ready: the caller's kernels wrote them)
    if (!keys_ready && !(integ && h->keys_fresh)) launch_keys(h, integ);
    // a multiple of 8 workgroups (one band of tiles per XCD) when there are 8 tiles or more
    const int ng = nt >= 8 ? std::min(nt, kSortGrid) & ~7 : nt;
    k_tsort<<<ng, kSortThr, 0, h->stream>>>(h->d_ri, h->d_stiles, h->d_cs[c], h->d_key, h->d_rmax, h->d_moved,
                                             h->d_tcount, h->d_locst, h->d_tsrc, h->d_tslot, h->d_err, nt);
    const int nch = (nt + kTScanChunk - 1) / kTScanChunk, per = nch > 1 ? 8 : (nt + 1023) / 1024;
    if (nch > 1) k_tsum<<<nch, 1024, 0, h->stream>>>(h->d_tcount, nt, per, h->d_tpart);
    k_tscan<<<nch + (step_delta ? 1 : 0), 1024, 0, h->stream>>>(h->d_ri, h->d_rs, h->nrep, h->d_tcount, nt, per,
                                         nch > 1 ? h->d_tpart : nullptr, h->d_toff, h->d_moved, h->d_rmax, step_delta);
    // the placement lists the precursor chemistry's roles of the replicas whose
    // check it is, and clears every bead's decision flag
    const ChemList CL = h->any_chem ? chem_list(h) : ChemList{nullptr, nullptr, nullptr};
    uint8_t *cf = h->any_chem ? h->d_ccls : nullptr;
    if (h->any_chem) h->cseg_n = nt;
    int32_t *src = h->rec_src;
    if (integ)
        k_place<true><<<ng, kSortThr, 0, h->stream>>>(h->d_ri, h->d_stiles, h->d_tcount, h->d_toff, h->d_locst,
                                                       h->d_tsrc, h->d_tslot, h->d_cs[n], src, h->st[c],
                                                       h->st[n], h->d_f, h->d_err, nt, h->d_rs, CL, cf);
    else
        k_place<false><<<ng, kSortThr, 0, h->stream>>>(h->d_ri, h->d_stiles, h->d_tcount, h->d_toff, h->d_locst,
                                                        h->d_tsrc, h->d_tslot, h->d_cs[n], src, h->st[c],
                                                        h->st[n], h->d_f, h->d_err, nt, h->d_rs, CL, cf);
    h->cur = n;
    h->keys_fresh = false;
}

static void launch_step_params(oc_engine *h, int64_t delta, int64_t set_step) {
    k_step_params<<<h->nrep, 64, 0, h->stream>>>(h->d_ri, h->d_rs, delta, set_step, nullptr);
}


// one full step of every replica: cel/engine/oercel.c:2300-2362.  With the
// reference chemistry in the reference's order: integrate, its chemistry on
// the previous cell index (catalyst reactions move beads), sort, then the
// precursor chemistry and the forces; without it the integrate is fused into
// the sort and every chemistry runs after it.
// chem = false leaves out the chemistry kernels (a step that is no check of
// any replica: they would find nothing to do).
static void enqueue_step(oc_engine *h, bool chem = true) {
    if (h->any_rchem) {
        const int c = h->cur;
        launch_step_params(h, 1, -1);
        k_integ<<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, h->st[c].xy,
                                                                    h->st[c].v, h->st[c].vt, h->d_f, h->d_key,
                                                                    h->d_rmax, h->d_moved, h->st[c].gt, h->d_cat,
                                                                    h->d_catn);
        if (chem) launch_rchem(h, false);
        launch_sort(h, false, 0, true);
        if (chem) launch_pchem(h, true, nullptr, false);
    } else {
        launch_sort(h, true, 1);
        launch_pchem(h, true, nullptr, false);
    }
    // the force pass settles the check's formation proposals
    launch_forces(h, true);
}

static int check_err(oc_engine *h) {
    uint32_t e = 0;
    HIPCHK(hipMemcpyAsync(&e, h->d_err, sizeof e, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (e & ERR_EVOVF) {
        // reported once, by the next oc_get_events; the bit is cleared so that
        // later drains report only new overflows
        h->events_dropped = true;
        const uint32_t cleared = e & ~ERR_EVOVF;
        HIPCHK(hipMemcpy(h->d_err, &cleared, sizeof cleared, hipMemcpyHostToDevice));
    }
    if (e & ~ERR_EVOVF) {
        std::string m = "device error:";
        if (e & ERR_SORT) m += " sort tile overflow or bead count mismatch (sort)";
        if (e & ERR_CHAIN) m += " broken lipid chain";
        if (e & ERR_TILE) m += " pair tile window outside its staged rows";
        if (e & ERR_CATCAP)
            m += " catalyst chemistry over capacity (more than 4096 active catalysts in one check of a replica, or "
                 "more beads within a reaction's reach than its gather list holds)";
        if (e & ERR_INJ)
            m += " a pair within the cutoff (or a formation candidate) is missing from the caller's table";
        if (e & ERR_DISP)
            m += " a bead moved 0.5 or more in the step before a chemistry check (the cell index of the last sort "
                 "no longer covers the reaction windows; reduce dt)";
        return fail(OC_EDEVICE, m);
    }
    return 0;
}

// --------------------------------------------------------------------------- host-side soup (init)

namespace {
struct HRng {
    uint64_t s[4];
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    // xoshiro256** seeded by splitmix64, cel/engine/oercel.c:143-174
    explicit HRng(uint64_t seed) {
        uint64_t z = seed;
        for (int k = 0; k < 4; k++) {
            z += 0x9e3779b97f4a7c15ULL;
            s[k] = mix13(z);
        }
    }
    uint64_t next() {
        const uint64_t out = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    double unif() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        double u1 = unif();
        const double u2 = unif();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};
double wrap_box(double v, double L) {
    // wrap1, cel/engine/oercel.c:761-772
    if (v < 0.0) {
        v += L;
        if (v >= L) v -= L;
    } else if (v >= L) {
        v -= L;
        if (v >= L || v < 0.0) v = fmod(v, L);
    }
    return v;
}
}  // namespace

OC_API int oc_soup_host(const oc_params *p, int32_t N, oc_state *o) {
    if (!p || !o || !o->x || !o->y || !o->vx || !o->vy || !o->type || !o->mpos || !o->nxt)
        return fail(OC_EINVAL, "oc_soup_host: missing output array");
    if (N < 2) return fail(OC_EINVAL, "N must be >= 2");
    MolTmpl M;
    const char *why = "";
    if (!build_template(*p, M, &why)) return fail(OC_EINVAL, why);
    const int ml = M.mlen;
    const double L = p->L;
    // composition, cel/engine/oercel.c:1775-1780
    const long nlip = lround(p->f_lip * N / ml);
    const long nKa = lround(p->f_Ka * N), nKb = lround(p->f_Kb * N), nN = lround(p->c_N * N);
    if (nlip * ml + nKa + nKb + nN > N) return fail(OC_EINVAL, "soup composition exceeds N");
    HRng rng(p->seed);
    int g = 0;
    for (long m = 0; m < nlip; m++) {
        // straight template chain from a random head position and angle
        const double hx = L * rng.unif(), hy = L * rng.unif();
        const double phi = 2.0 * M_PI * rng.unif();
        const double cph = cos(phi), sph = sin(phi);
        for (int b = 0; b < ml; b++, g++) {
            if (M.arch == OC_ARCH_LINEAR) {
                o->x[g] = wrap_box(hx + b * p->r0 * cos(phi), L);
                o->y[g] = wrap_box(hy + b * p->r0 * sin(phi), L);
            } else {
                // the template rotated by phi about bead 0
                o->x[g] = wrap_box(hx + (M.ix[b] * cph - M.iy[b] * sph), L);
                o->y[g] = wrap_box(hy + (M.ix[b] * sph + M.iy[b] * cph), L);
            }
            o->type[g] = b < M.nH ? OC_H : OC_T;
            o->mpos[g] = (int8_t)b;
            o->nxt[g] = b + 1 < ml ? g + 1 : -1;
            if (o->bond) o->bond[g] = -1;
        }
    }
    for (long k = 0; g < N; g++, k++) {
        o->x[g] = L * rng.unif();
        o->y[g] = L * rng.unif();
        o->type[g] = k < nKa ? OC_KA : (k < nKa + nKb ? OC_KB : (k < nKa + nKb + nN ? OC_N : OC_W));
        o->mpos[g] = -1;
        o->nxt[g] = -1;
        if (o->bond) o->bond[g] = -1;
    }
    // Maxwell-Boltzmann velocities, net momentum removed, scaled to exact kT
    double px = 0, py = 0;
    const double skT = sqrt(p->kT);
    for (int i = 0; i < N; i++) {
        o->vx[i] = rng.normal() * skT;
        o->vy[i] = rng.normal() * skT;
        px += o->vx[i];
        py += o->vy[i];
    }
    px /= N;
    py /= N;
    double k2 = 0;
    for (int i = 0; i < N; i++) {
        o->vx[i] -= px;
        o->vy[i] -= py;
        k2 += o->vx[i] * o->vx[i] + o->vy[i] * o->vy[i];
    }
    const double sc = sqrt(p->kT * (2.0 * N - 2.0) / k2);
    for (int i = 0; i < N; i++) {
        o->vx[i] *= sc;
        o->vy[i] *= sc;
    }
    if (o->fx)
        for (int i = 0; i < N; i++) o->fx[i] = 0.0;
    if (o->fy)
        for (int i = 0; i < N; i++) o->fy[i] = 0.0;
    return 0;
}

// --------------------------------------------------------------------------- create / destroy

OC_API const char *oc_last_error(void) { return g_err.c_str(); }
OC_API int oc_abi_version(void) { return OC_ABI_VERSION; }

static void adj_free(oc_engine *h);

OC_API void oc_destroy(oc_engine *h) {
    if (!h) return;
    adj_free(h);
    drop_graphs(h);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

OC_API int oc_create(const oc_params *p, const int32_t *N, int32_t nrep, oc_engine **out) {
    if (!p || !N || !out || nrep < 1) return fail(OC_EINVAL, "oc_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(OC_EHIP, "no HIP device");
    oc_engine *h = new oc_engine();
    h->nrep = nrep;
    h->prm.assign(p, p + nrep);
    h->ri.resize(nrep);
    int64_t off = 0, coff = 0, col0 = 0;
    for (int r = 0; r < nrep; r++) {
        if (int rc = make_repinfo(p[r], N[r], h->ri[r])) {
            delete h;
            return rc;
        }
        RepInfo &R = h->ri[r];
        R.off = (int32_t)off;
        R.coff = (int32_t)coff;
        R.col0 = (int32_t)col0;
        for (int cx = 0; cx < R.ncx; cx++) h->cols.push_back(make_int2(r, cx));
        off += R.N;
        coff += R.ncell;
        col0 += R.ncx;
        if (off >= INT32_MAX || coff + col0 >= INT32_MAX) {
            delete h;
            return fail(OC_EINVAL, "too many beads or cells for 32-bit slot indices");
        }
        R.bblk0 = (int32_t)h->bblk.size();
        for (int b = 0; b < grid_of(R.N, kBeadBlock); b++) h->bblk.push_back(r);
        // sort tiles of about 192 beads (at most kSortRows rows)
        {
            const double per_row = (double)R.N / ((double)R.ncx * R.ncy);
            const int rows = std::max(1, std::min(kSortRows, (int)lround(192.0 / std::max(per_row, 1e-9))));
            const int nt = (R.ncy + rows - 1) / rows;
            R.stile0 = (int32_t)h->stiles.size();
            R.ntc = nt;
            for (int cx = 0; cx < R.ncx; cx++)
                for (int k = 0; k < nt; k++) {
                    Tile T;
                    T.rep = r;
                    T.cx = cx;
                    T.k0 = (int)(((int64_t)k * R.ncy) / nt);
                    T.k1 = (int)(((int64_t)(k + 1) * R.ncy) / nt);
                    h->stiles.push_back(T);
                }
        }
        R.rblk0 = (int32_t)h->rblk.size();
        for (int b = 0; b < grid_of(R.N, kRedChunk); b++) h->rblk.push_back(r);
        h->any_stir |= R.has_stir != 0;
        h->any_chem |= R.chem != 0;
        h->any_angle |= R.ka != 0.0 && R.arch == 0;
        h->any_tmpl |= R.arch != 0;
        h->any_rchem |= R.rchem != 0;
        h->any_react |= R.react != 0;
        h->any_fcap |= R.fcap_on != 0;
    }
    h->nokeys = h->any_tmpl || h->any_rchem;
    h->Ntot = off;
    h->Ctot = coff;
    if ((int64_t)h->stiles.size() << kSortCapLog >= INT32_MAX) {
        delete h;
        return fail(OC_EINVAL, "too many sort tiles for 32-bit indices");
    }
    if (h->Ntot >= kPairSlotLimit) {  // the pair kernel packs a staged bead's slot in 28 bits
        delete h;
        return fail(OC_EINVAL, "more than 2^28 beads in one engine");
    }
    h->evq.resize(nrep);
    h->evcap = std::min<int64_t>(std::max<int64_t>(1 << 16, 4 * h->Ntot), 8 << 20);
    if (const char *ec = getenv("OC_EVENT_CAP")) {  // a smaller event buffer (tests of the overflow path)
        const long long v = atoll(ec);
        if (v > 0) h->evcap = std::min<int64_t>(h->evcap, v);
    }
    int rc = 0;
    const size_t Nt = (size_t)h->Ntot, Ct = (size_t)h->Ctot;
    rc = rc ? rc : dalloc(h, &h->d_ri, nrep);
    rc = rc ? rc : dalloc(h, &h->d_rs, nrep);
    rc = rc ? rc : dalloc(h, &h->d_bblk, h->bblk.size());
    rc = rc ? rc : dalloc(h, &h->d_brng, nrep);
    rc = rc ? rc : dalloc(h, &h->d_stiles, h->stiles.size());
    rc = rc ? rc : dalloc(h, &h->d_rblk, h->rblk.size());
    for (int k = 0; k < 2; k++) {
        rc = rc ? rc : dalloc(h, &h->st[k].xy, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].v, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].vt, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].gt, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].mp, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].lnk, Nt);
        rc = rc ? rc : dalloc(h, &h->st[k].prv, Nt);
        rc = rc ? rc : dalloc(h, &h->d_cs[k], Ct + 1);
    }
    rc = rc ? rc : dalloc(h, &h->d_f, Nt);
    rc = rc ? rc : dalloc(h, &h->d_key, Nt);
    rc = rc ? rc : dalloc(h, &h->d_tcount, h->stiles.size());
    rc = rc ? rc : dalloc(h, &h->d_toff, h->stiles.size());
    rc = rc ? rc : dalloc(h, &h->d_tsrc, h->stiles.size() << kSortCapLog);
    rc = rc ? rc : dalloc(h, &h->d_tslot, Nt);
    rc = rc ? rc : dalloc(h, &h->d_locst, Ct);
    rc = rc ? rc : dalloc(h, &h->d_moved, 1 + h->stiles.size());
    rc = rc ? rc : dalloc(h, &h->d_rmax, h->nrep);
    rc = rc ? rc : dalloc(h, &h->d_tpart, (h->stiles.size() + kTScanChunk - 1) / kTScanChunk);
    rc = rc ? rc : dalloc(h, &h->d_epot, Nt);
    rc = rc ? rc : dalloc(h, &h->d_evir, Nt);
    rc = rc ? rc : dalloc(h, &h->d_ebond, Nt);
    rc = rc ? rc : dalloc(h, &h->d_part, h->rblk.size() * 16);
    rc = rc ? rc : dalloc(h, &h->d_statout, (size_t)nrep * 16);
    rc = rc ? rc : dalloc(h, &h->d_ccls, h->any_chem ? Nt : 1);
    {
        const size_t nseg = h->any_chem ? std::max(h->stiles.size(), h->bblk.size()) : 1;
        rc = rc ? rc : dalloc(h, &h->d_cseg, nseg);
        rc = rc ? rc : dalloc(h, &h->d_cinl, nseg * kChemInl);
        rc = rc ? rc : dalloc(h, &h->d_cspill, h->any_chem ? Nt : 1);
    }
    rc = rc ? rc : dalloc(h, &h->d_ev, (size_t)h->evcap);
    rc = rc ? rc : dalloc(h, &h->d_evn, 1);
    rc = rc ? rc : dalloc(h, &h->d_err, 1);
    if (h->any_stir) rc = rc ? rc : dalloc(h, &h->d_stir, Nt);
    if (h->any_react) {
        // gather lists: three times the expected beads within the reach, plus a margin
        double want = 0.0;
        for (const RepInfo &R : h->ri)
            if (R.react) want = std::max(want, 3.0 * (R.N / (R.L * R.L)) * M_PI * R.rg2 + 128.0);
        h->gcap = (int)std::min(65536.0, std::ceil(want / 64.0) * 64.0);
        rc = rc ? rc : dalloc(h, &h->d_cat, (size_t)nrep * kCatCap);
        rc = rc ? rc : dalloc(h, &h->d_catn, nrep);
        rc = rc ? rc : dalloc(h, &h->d_ccell, (size_t)nrep * kCatCap);
        rc = rc ? rc : dalloc(h, &h->d_cdone, (size_t)nrep * kCatCap);
        rc = rc ? rc : dalloc(h, &h->d_gs, (size_t)(kFreeBlocks * 4 + (kRestThr / 64) * nrep) * h->gcap);
    }
    if (h->any_fcap) rc = rc ? rc : dalloc(h, &h->d_fkey, Nt);
    if (h->any_fcap) rc = rc ? rc : dalloc(h, &h->d_fcnt, nrep);
    if (h->any_fcap) rc = rc ? rc : dalloc(h, &h->d_npart, h->bblk.size());
    if (rc) {
        oc_destroy(h);
        return rc;
    }
    if (hipStreamCreate(&h->stream) != hipSuccess) {
        oc_destroy(h);
        return fail(OC_EHIP, "hipStreamCreate failed");
    }
    std::vector<RepStep> rs(nrep);
    memset(rs.data(), 0, rs.size() * sizeof(RepStep));
    for (int r = 0; r < nrep; r++) rs[r].step = p[r].step_anchor;
    hipError_t e = hipSuccess;
    e = e ? e : hipMemcpy(h->d_ri, h->ri.data(), nrep * sizeof(RepInfo), hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(h->d_rs, rs.data(), nrep * sizeof(RepStep), hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(h->d_bblk, h->bblk.data(), h->bblk.size() * 4, hipMemcpyHostToDevice);
    {
        std::vector<int2> br(nrep);
        for (int r = 0; r < nrep; r++)
            br[r] = make_int2(h->ri[r].bblk0 * kRdfSub, grid_of(h->ri[r].N, kBeadBlock) * kRdfSub);
        e = e ? e : hipMemcpy(h->d_brng, br.data(), nrep * sizeof(int2), hipMemcpyHostToDevice);
    }
    e = e ? e : hipMemcpy(h->d_stiles, h->stiles.data(), h->stiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(h->d_rblk, h->rblk.data(), h->rblk.size() * 4, hipMemcpyHostToDevice);
    e = e ? e : hipMemset(h->d_evn, 0, sizeof(unsigned long long));
    e = e ? e : hipMemset(h->d_err, 0, 4);
    e = e ? e : hipMemset(h->d_moved, 0, 4 * (1 + h->stiles.size()));
    e = e ? e : hipMemset(h->d_rmax, 0, h->nrep * sizeof(int2));
    e = e ? e : hipMemset(h->d_ebond, 0, Nt * sizeof(double));
    e = e ? e : hipMemset(h->d_epot, 0, Nt * sizeof(double));
    e = e ? e : hipMemset(h->d_evir, 0, Nt * sizeof(double));
    e = e ? e : hipMemset(h->d_f, 0, Nt * sizeof(double2));
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemset(h->st[k].lnk, 0xff, Nt * 4);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemset(h->st[k].prv, 0xff, Nt * 4);
    if (h->any_react) e = e ? e : hipMemset(h->d_catn, 0, nrep * sizeof(int32_t));
    if (h->any_fcap) e = e ? e : hipMemset(h->d_fcnt, 0, nrep * sizeof(int2));
    if (e != hipSuccess) {
        oc_destroy(h);
        return fail(OC_EHIP, std::string("setup copy failed: ") + hipGetErrorString(e));
    }
    if ((rc = build_tiles(h))) {
        oc_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

OC_API int oc_set_tuning(oc_engine *h, int32_t pair_block, int32_t tile_beads) {
    if (!h) return fail(OC_EINVAL, "null engine");
    if (pair_block != 0 && pair_block != 64 && pair_block != 128 && pair_block != 256)
        return fail(OC_EINVAL, "pair_block must be 64, 128 or 256");
    h->pair_block = pair_block ? pair_block : 256;
    h->tile_beads = tile_beads > 0 ? tile_beads : 0;
    drop_graphs(h);
    adj_free(h);  // the window gradient's per-tile sums follow the tiles
    return build_tiles(h);
}

OC_API void *oc_stream(oc_engine *h) { return h ? (void *)h->stream : nullptr; }

OC_API int oc_set_pair_variant(oc_engine *h, int32_t lists) {
    if (!h) return fail(OC_EINVAL, "null engine");
    if (lists < 0 || lists > 3) return fail(OC_EINVAL, "pair variant must be 0, 1, 2 or 3");
    h->pair_variant = lists;
    drop_graphs(h);
    return 0;
}

OC_API int oc_abi_sizes(int32_t *out, int32_t n) {
    const int32_t sz[4] = {(int32_t)sizeof(oc_params), (int32_t)sizeof(oc_state), (int32_t)sizeof(oc_event),
                           (int32_t)sizeof(oc_stats)};
    int k = 0;
    for (; out && k < n && k < 4; k++) out[k] = sz[k];
    return k;
}

// --------------------------------------------------------------------------- state in / out

// Load replica r from gid-order host arrays: stable counting sort by cell on
// the host (the order sort_cells gives a gid-ordered input), upload in slot order.
// order (optional): gids in the order the beads held their slots (a
// checkpoint's slot order); the stable sort by cell then keeps that order
// within each cell, as the reference's state does after a restart.
static int load_replica(oc_engine *h, int r, const oc_state *s, size_t base, bool vt_from_v,
                        const uint32_t *order = nullptr) {
    const RepInfo &R = h->ri[r];
    const int N = R.N;
    if (!s->x || !s->y || !s->vx || !s->vy || !s->type || !s->mpos) return fail(OC_EINVAL, "state: missing array");
    std::vector<int32_t> key(N), cnt(R.ncell + 1, 0), src(N), dest(N);
    if (order) {
        std::vector<uint8_t> seen(N, 0);
        for (int q = 0; q < N; q++) {
            if (order[q] >= (uint32_t)N || seen[order[q]]) return fail(OC_EINVAL, "slot order is not a permutation");
            seen[order[q]] = 1;
        }
    }
    for (int g = 0; g < N; g++) {
        const double x = s->x[base + g], y = s->y[base + g];
        if (!(x >= 0.0 && x < R.L && y >= 0.0 && y < R.L)) return fail(OC_EINVAL, "state: position outside [0, L)");
        int cx = (int)(x * R.icsx), cy = (int)(y * R.icsy);
        cx = cx >= R.ncx ? R.ncx - 1 : cx;
        cy = cy >= R.ncy ? R.ncy - 1 : cy;
        key[g] = cx * R.ncy + cy;
        cnt[key[g]]++;
    }
    std::vector<int32_t> cs(R.ncell + 1);
    int run = 0;
    for (int c = 0; c < R.ncell; c++) {
        cs[c] = run;
        run += cnt[c];
    }
    cs[R.ncell] = run;
    std::vector<int32_t> fill(cs.begin(), cs.end() - 1);
    for (int q = 0; q < N; q++) {
        const int g = order ? (int)order[q] : q;
        const int d = fill[key[g]]++;
        src[d] = g;
        dest[g] = d;
    }
    std::vector<double2> xy(N), v(N), f(N);
    std::vector<float2> vt(N);
    std::vector<uint32_t> gt(N);
    std::vector<int8_t> mp(N);
    std::vector<int32_t> lnk(N), csg(R.ncell + 1);
    for (int d = 0; d < N; d++) {
        const int g = src[d];
        const size_t k = base + g;
        const int ty = s->type[k];
        if (ty < 0 || ty >= OC_NTYPES) return fail(OC_EINVAL, "state: bad type");
        xy[d] = make_double2(s->x[k], s->y[k]);
        v[d] = make_double2(s->vx[k], s->vy[k]);
        f[d] = make_double2(s->fx ? s->fx[k] : 0.0, s->fy ? s->fy[k] : 0.0);
        vt[d] = make_float2((float)s->vx[k], (float)s->vy[k]);
        gt[d] = (uint32_t)g | ((uint32_t)ty << kGidBits);
        mp[d] = s->mpos[k];
        int32_t partner = -1;
        if (s->mpos[k] >= 0) partner = s->nxt ? s->nxt[k] : -1;
        else if (s->bond) partner = s->bond[k];
        if (partner >= N) return fail(OC_EINVAL, "state: link gid out of range");
        lnk[d] = partner >= 0 ? R.off + dest[partner] : -1;
    }
    for (int c = 0; c <= R.ncell; c++) csg[c] = cs[c] + R.off;
    (void)vt_from_v;
    const int c = h->cur;
    const size_t o = R.off;
    HIPCHK(hipMemcpy(h->st[c].xy + o, xy.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->st[c].v + o, v.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->st[c].vt + o, vt.data(), N * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->st[c].gt + o, gt.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->st[c].mp + o, mp.data(), N * sizeof(int8_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->st[c].lnk + o, lnk.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_f + o, f.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_cs[c] + R.coff, csg.data(), (R.ncell + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

// step counter of a loaded state; a state load also clears the device error
// word (an error a step reported leaves the engine usable after a reload)
static int set_step_fields(oc_engine *h, int r, int64_t step) {
    RepStep rs;
    HIPCHK(hipMemcpy(&rs, h->d_rs + r, sizeof rs, hipMemcpyDeviceToHost));
    rs.step = step;
    rs.maxbond_bits = 0;
    rs.disp = 0;
    HIPCHK(hipMemcpy(h->d_rs + r, &rs, sizeof rs, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(h->d_err, 0, sizeof(uint32_t)));
    return 0;
}

OC_API int oc_start(oc_engine *h, const oc_state *s) {
    if (!h || !s) return fail(OC_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(h->stream));
    size_t base = 0;
    for (int r = 0; r < h->nrep; r++) {
        if (int rc = load_replica(h, r, s, base, true)) return rc;
        if (int rc = set_step_fields(h, r, h->prm[r].step_anchor)) return rc;
        // a fresh start ramps the repulsion in over `warmup` steps
        // (g_warm_end = warmup, cel/engine/oercel.c:2197)
        h->ri[r].warm_end = h->prm[r].step_anchor + h->prm[r].warmup;
        base += h->ri[r].N;
    }
    HIPCHK(hipMemcpy(h->d_ri, h->ri.data(), h->nrep * sizeof(RepInfo), hipMemcpyHostToDevice));
    drop_graphs(h);
    h->have_state = true;
    launch_build_prv(h);
    // forces at the start step, then vh = v - dt/2 f (cel/engine/oercel.c:2203-2211)
    launch_step_params(h, 0, -1);
    launch_forces(h, true);
    k_half_kick<<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, h->st[h->cur].v, h->d_f);
    h->keys_fresh = false;
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_set_state(oc_engine *h, int32_t r, const oc_state *s, int64_t step) {
    return oc_set_state_slots(h, r, s, step, nullptr);
}

OC_API int oc_set_state_slots(oc_engine *h, int32_t r, const oc_state *s, int64_t step, const uint32_t *slot_gid) {
    if (!h || !s || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (int rc = load_replica(h, r, s, 0, false, slot_gid)) return rc;
    if (int rc = set_step_fields(h, r, step)) return rc;
    h->keys_fresh = false;
    launch_build_prv(h);
    launch_step_params(h, 0, -1);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_state = true;
    return 0;
}

OC_API int oc_get_state(oc_engine *h, int32_t r, oc_state *s) {
    if (!h || !s || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    const RepInfo &R = h->ri[r];
    const int N = R.N, c = h->cur;
    const size_t o = R.off;
    std::vector<double2> xy(N), v(N), f(N);
    std::vector<uint32_t> gt(N);
    std::vector<int8_t> mp(N);
    std::vector<int32_t> lnk(N);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(xy.data(), h->st[c].xy + o, N * sizeof(double2), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(v.data(), h->st[c].v + o, N * sizeof(double2), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(f.data(), h->d_f + o, N * sizeof(double2), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gt.data(), h->st[c].gt + o, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mp.data(), h->st[c].mp + o, N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lnk.data(), h->st[c].lnk + o, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int d = 0; d < N; d++) {
        const uint32_t g = gt[d] & kGidMask;
        if (s->x) s->x[g] = xy[d].x;
        if (s->y) s->y[g] = xy[d].y;
        if (s->vx) s->vx[g] = v[d].x;
        if (s->vy) s->vy[g] = v[d].y;
        if (s->fx) s->fx[g] = f[d].x;
        if (s->fy) s->fy[g] = f[d].y;
        if (s->type) s->type[g] = (uint8_t)(gt[d] >> kGidBits);
        if (s->mpos) s->mpos[g] = mp[d];
        const int32_t l = lnk[d];
        const int32_t lg = l < 0 ? -1 : (int32_t)(gt[l - R.off] & kGidMask);
        if (s->nxt) s->nxt[g] = mp[d] >= 0 ? lg : -1;
        if (s->bond) s->bond[g] = mp[d] < 0 ? lg : -1;
    }
    return 0;
}

// --------------------------------------------------------------------------- stepping

// Two steps, captured with the keys of the first taken from the force
// kernel before it (oc_step makes them fresh before a replay).
static int capture_graph(oc_engine *h, int parity) {
    const int saved = h->cur;
    const bool fresh = h->keys_fresh;
    h->cur = parity;
    h->keys_fresh = !h->nokeys;  // without keys from the force kernel the graph's first step computes them
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(h);
    enqueue_step(h);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->cur = saved;
    h->keys_fresh = fresh;
    if (e != hipSuccess) return fail(OC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    h->graph[parity] = g;
    HIPCHK(hipGraphInstantiate(&h->gexec[parity], g, nullptr, nullptr, 0));
    return 0;
}

// The n_chem steps from a check to the next one as one graph: the chemistry
// kernels run only in its last step (the check).
static int capture_period(oc_engine *h, int parity, int nc) {
    const int saved = h->cur;
    h->cur = parity;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 1; s <= nc; s++) enqueue_step(h, s == nc);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->cur = saved;
    if (e != hipSuccess) return fail(OC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    h->pgraph[parity] = g;
    HIPCHK(hipGraphInstantiate(&h->pexec[parity], g, nullptr, nullptr, 0));
    return 0;
}

// n_chem of every replica when they share it and their step counters agree
// modulo it (then the check steps are known here), else 0; step: replica 0's
static int common_period(oc_engine *h, int64_t *step) {
    const int nc = h->ri[0].n_chem;
    if (nc < 2 || nc > 1000) return 0;
    std::vector<RepStep> rs(h->nrep);
    if (hipMemcpy(rs.data(), h->d_rs, h->nrep * sizeof(RepStep), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    for (int r = 0; r < h->nrep; r++)
        if (h->ri[r].n_chem != nc || ((rs[r].step - rs[0].step) % nc) != 0) return 0;
    *step = rs[0].step;
    return nc;
}

OC_API int oc_step(oc_engine *h, int64_t n, int32_t use_graph) {
    if (!h) return fail(OC_EINVAL, "null engine");
    if (!h->have_state) return fail(OC_ESTATE, "no state: call oc_start or oc_set_state first");
    if (n <= 0) return 0;
    int64_t k = 0, step = 0;
    const int nc = use_graph && h->any_rchem && n >= 2 ? common_period(h, &step) : 0;
    if (nc) {
        // full steps up to the next check, then whole periods from the graph
        for (; k < n && step % nc != 0; k++, step++) enqueue_step(h);
        for (; k + nc <= n; k += nc) {
            const int p = h->cur;
            if (!h->pexec[p])
                for (int q = 0; q < 2; q++)  // both buffer parities, uploaded, before the first replay
                    if (!h->pexec[q]) {
                        if (int rc = capture_period(h, q, nc)) return rc;
                        HIPCHK(hipGraphUpload(h->pexec[q], h->stream));
                    }
            HIPCHK(hipGraphLaunch(h->pexec[p], h->stream));
            h->cur = p ^ (nc & 1);
        }
    } else if (use_graph && n >= 2) {
        const int p = h->cur;
        if (!h->gexec[p])
            for (int q = 0; q < 2; q++)  // both buffer parities, uploaded, before the first replay
                if (!h->gexec[q]) {
                    if (int rc = capture_graph(h, q)) return rc;
                    HIPCHK(hipGraphUpload(h->gexec[q], h->stream));
                }
        if (!h->keys_fresh && !h->nokeys) launch_keys(h, true);
        for (; k + 2 <= n; k += 2) HIPCHK(hipGraphLaunch(h->gexec[p], h->stream));
        h->keys_fresh = !h->nokeys;
    }
    for (; k < n; k++) enqueue_step(h);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_sort(oc_engine *h) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    launch_step_params(h, 0, -1);
    launch_sort(h, false, 0);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_forces(oc_engine *h, int64_t step, int32_t with_bonded) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    launch_step_params(h, 0, step);
    launch_forces(h, with_bonded != 0);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_launch_pair(oc_engine *h, int32_t reps) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    if (reps <= 0) return 0;
    const int c = h->cur;
    if (h->treps[c] != reps) {
        // the launches as one graph, so that back-to-back launches are timed
        // without the host's per-launch gaps; rebuilt when the count changes
        if (h->texec[c]) (void)hipGraphExecDestroy(h->texec[c]);
        if (h->tgraph[c]) (void)hipGraphDestroy(h->tgraph[c]);
        h->texec[c] = nullptr;
        h->tgraph[c] = nullptr;
        h->treps[c] = 0;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < reps; k++) launch_pair(h, false, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (e != hipSuccess) return fail(OC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        h->tgraph[c] = g;
        HIPCHK(hipGraphInstantiate(&h->texec[c], g, nullptr, nullptr, 0));
        h->treps[c] = reps;
    }
    HIPCHK(hipGraphLaunch(h->texec[c], h->stream));
    return 0;
}

OC_API int oc_chem(oc_engine *h, int64_t step) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    launch_step_params(h, 0, step);
    launch_chem(h);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

// Sorted (lo << 32 | hi) keys of a caller's pair table, with its checks.
static int pair_keys(const uint32_t *lo, const uint32_t *hi, int64_t n, std::vector<uint64_t> &k) {
    if (n < 0 || (n > 0 && (!lo || !hi))) return fail(OC_EINVAL, "bad pair table");
    k.resize((size_t)n);
    for (int64_t q = 0; q < n; q++) {
        if (lo[q] >= hi[q]) return fail(OC_EINVAL, "pair table: every entry needs lo < hi");
        k[q] = ((uint64_t)lo[q] << 32) | hi[q];
        if (q > 0 && k[q] <= k[q - 1]) return fail(OC_EINVAL, "pair table: keys (lo, hi) must strictly increase");
    }
    return 0;
}

// Device copies of a caller's tables, freed when the call ends.
struct CallMem {
    std::vector<void *> p;
    ~CallMem() {
        for (void *q : p) (void)hipFree(q);
    }
    template <class T>
    int put(const T *src, size_t n, T **out) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return fail(OC_ENOMEM, "hipMalloc failed");
        p.push_back(q);
        if (n && hipMemcpy(q, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return fail(OC_EHIP, "copy of a caller table failed");
        *out = (T *)q;
        return 0;
    }
};

OC_API int oc_forces_theta(oc_engine *h, int64_t step, int32_t with_bonded, const uint32_t *pair_lo,
                           const uint32_t *pair_hi, const float *theta, int64_t npair) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    if (h->nrep != 1) return fail(OC_EINVAL, "oc_forces_theta needs an engine of one replica");
    if (npair > 0 && !theta) return fail(OC_EINVAL, "null theta");
    std::vector<uint64_t> k;
    if (int rc = pair_keys(pair_lo, pair_hi, npair, k)) return rc;
    CallMem m;
    InjTheta inj{nullptr, nullptr, npair, h->d_err};
    if (int rc = m.put(k.data(), k.size(), (uint64_t **)&inj.k)) return rc;
    if (int rc = m.put(theta, (size_t)npair, (float **)&inj.v)) return rc;
    launch_step_params(h, 0, step);
    launch_forces(h, with_bonded != 0, &inj);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_chem_uniforms(oc_engine *h, int64_t step, const double *u_conv, const double *u_break,
                            const uint32_t *form_lo, const uint32_t *form_hi, const double *u_form, int64_t nform) {
    if (!h || !h->have_state) return fail(OC_ESTATE, "no state");
    if (h->nrep != 1) return fail(OC_EINVAL, "oc_chem_uniforms needs an engine of one replica");
    if (!h->any_chem || h->any_rchem) return fail(OC_EINVAL, "oc_chem_uniforms: precursor chemistry only (chem on, rchem off)");
    if (!u_conv || !u_break || (nform > 0 && !u_form)) return fail(OC_EINVAL, "null uniforms");
    std::vector<uint64_t> k;
    if (int rc = pair_keys(form_lo, form_hi, nform, k)) return rc;
    const size_t N = (size_t)h->ri[0].N;
    CallMem m;
    ChemInj inj{nullptr, nullptr, nullptr, nullptr, nform, h->d_err};
    if (int rc = m.put(u_conv, N, (double **)&inj.uconv)) return rc;
    if (int rc = m.put(u_break, N, (double **)&inj.ubreak)) return rc;
    if (int rc = m.put(k.data(), k.size(), (uint64_t **)&inj.fk)) return rc;
    if (int rc = m.put(u_form, (size_t)nform, (double **)&inj.fv)) return rc;
    launch_step_params(h, 0, step);
    launch_pchem(h, false, &inj);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_get_step(oc_engine *h, int32_t r, int64_t *step) {
    if (!h || !step || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    RepStep rs;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&rs, h->d_rs + r, sizeof rs, hipMemcpyDeviceToHost));
    *step = rs.step;
    return 0;
}

// --------------------------------------------------------------------------- cells, events, stats

OC_API int oc_get_cells(oc_engine *h, int32_t r, int32_t *cstart, uint32_t *gid_of_slot, int32_t *ncell) {
    if (!h || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    const RepInfo &R = h->ri[r];
    if (ncell) *ncell = R.ncell;
    HIPCHK(hipStreamSynchronize(h->stream));
    const int c = h->cur;
    if (cstart) {
        HIPCHK(hipMemcpy(cstart, h->d_cs[c] + R.coff, (R.ncell + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int k = 0; k <= R.ncell; k++) cstart[k] -= R.off;
    }
    if (gid_of_slot) {
        HIPCHK(hipMemcpy(gid_of_slot, h->st[c].gt + R.off, R.N * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (int k = 0; k < R.N; k++) gid_of_slot[k] &= kGidMask;
    }
    return 0;
}

static int drain_events(oc_engine *h) {
    unsigned long long n = 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&n, h->d_evn, sizeof n, hipMemcpyDeviceToHost));
    const int64_t m = std::min<int64_t>((int64_t)n, h->evcap);
    if (m > 0) {
        std::vector<Event> ev(m);
        HIPCHK(hipMemcpy(ev.data(), h->d_ev, m * sizeof(Event), hipMemcpyDeviceToHost));
        for (const Event &e : ev) {
            oc_event o;
            o.step = e.step;
            o.kind = e.kind;
            o.gid_a = e.gid_a;
            o.gid_b = e.gid_b;
            o.pad = 0;
            o.x = e.x;
            o.y = e.y;
            h->evq[e.rep].push_back(o);
        }
    }
    const unsigned long long z = 0;
    HIPCHK(hipMemcpy(h->d_evn, &z, sizeof z, hipMemcpyHostToDevice));
    return 0;
}

OC_API int oc_get_events(oc_engine *h, int32_t r, oc_event *ev, int64_t cap, int64_t *n_out) {
    if (!h || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    if (int rc = drain_events(h)) return rc;
    std::vector<oc_event> &q = h->evq[r];
    std::sort(q.begin(), q.end(), [](const oc_event &a, const oc_event &b) {
        if (a.step != b.step) return a.step < b.step;
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.gid_a != b.gid_a) return a.gid_a < b.gid_a;
        return a.gid_b < b.gid_b;
    });
    const int64_t n = (int64_t)q.size();
    if (n_out) *n_out = n;
    const int64_t m = ev ? std::min<int64_t>(n, std::max<int64_t>(cap, 0)) : 0;
    for (int64_t k = 0; k < m; k++) ev[k] = q[k];
    q.erase(q.begin(), q.begin() + m);
    if (h->events_dropped) {
        h->events_dropped = false;
        return fail(OC_EDEVICE, "event buffer overflowed: some events were dropped");
    }
    return 0;
}

OC_API int oc_get_stats(oc_engine *h, int32_t r, oc_stats *out) {
    if (!h || !out || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    if (!h->have_state) return fail(OC_ESTATE, "no state");
    const int c = h->cur;
    launch_pair(h, true);
    k_bonded<true><<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, h->st[c].xy,
                                                                      h->st[c].gt, h->st[c].mp, h->st[c].lnk, h->d_f,
                                                                      h->d_ebond, h->d_err);
    if (h->any_tmpl)
        k_mol<true><<<(int)h->bblk.size(), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_rs, h->d_bblk, h->st[c].xy,
                                                                      h->st[c].mp, h->st[c].lnk, h->d_f, h->d_ebond,
                                                                      h->d_err);
    k_stats_partial<<<(int)h->rblk.size(), 256, 0, h->stream>>>(h->d_ri, h->d_rblk, h->st[c].v, h->d_f, h->st[c].gt,
                                                               h->st[c].mp, h->st[c].lnk, h->d_epot, h->d_ebond,
                                                               h->d_evir, h->d_part);
    k_stats_final<<<h->nrep, 256, 0, h->stream>>>(h->d_ri, h->d_part, h->d_statout);
    HIPCHK(hipGetLastError());
    if (int rc = check_err(h)) return rc;
    double v[16];
    RepStep rs;
    HIPCHK(hipMemcpy(v, h->d_statout + (size_t)r * 16, sizeof v, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&rs, h->d_rs + r, sizeof rs, hipMemcpyDeviceToHost));
    const RepInfo &R = h->ri[r];
    memset(out, 0, sizeof *out);
    out->step = rs.step;
    out->time = R.t_anchor + (double)(rs.step - R.step_anchor) * R.dt;
    out->T_kin = v[0] / (2.0 * R.N - 2.0);
    out->px = v[1];
    out->py = v[2];
    out->E_kin = 0.5 * v[0];
    out->E_pair = v[3];
    out->E_bond = v[4];
    double mb;
    memcpy(&mb, &rs.maxbond_bits, sizeof mb);
    out->max_bond = mb;
    for (int k = 0; k < OC_NTYPES; k++) out->count[k] = (int64_t)v[5 + k];
    out->n_lipids = (int64_t)v[12];
    out->n_bonds = (int64_t)v[13];
    out->virial = v[15];
    out->pressure = (double)R.N / (R.L * R.L) * out->T_kin + v[15] / (2.0 * R.L * R.L);
    rs.maxbond_bits = 0;
    HIPCHK(hipMemcpy(h->d_rs + r, &rs, sizeof rs, hipMemcpyHostToDevice));
    return 0;
}

OC_API int oc_max_bond(oc_engine *h, int32_t r, double *get, const double *set) {
    if (!h || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    if (set && !(*set >= 0.0)) return fail(OC_EINVAL, "max bond must be >= 0");
    RepStep rs;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&rs, h->d_rs + r, sizeof rs, hipMemcpyDeviceToHost));
    if (get) memcpy(get, &rs.maxbond_bits, sizeof *get);
    if (set) {
        memcpy(&rs.maxbond_bits, set, sizeof *set);
        HIPCHK(hipMemcpy(h->d_rs + r, &rs, sizeof rs, hipMemcpyHostToDevice));
    }
    return 0;
}

// --------------------------------------------------------------------------- window gradient

OC_API int oc_set_pair_params(oc_engine *h, int32_t r, const double *a64, double gamma) {
    if (!h || !a64 || r < 0 || r >= h->nrep) return fail(OC_EINVAL, "bad arguments");
    if (!(gamma > 0.0)) return fail(OC_EINVAL, "gamma must be > 0");
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            const double v = (i < OC_NTYPES && j < OC_NTYPES) ? a64[8 * i + j] : 0.0;
            if (v != ((i < OC_NTYPES && j < OC_NTYPES) ? a64[8 * j + i] : 0.0))
                return fail(OC_EINVAL, "repulsion matrix must be symmetric");
            h->prm[r].a[i][j] = v;
            h->ri[r].a[8 * i + j] = v;
        }
    h->prm[r].gamma = gamma;
    h->ri[r].gamma = gamma;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->d_ri + r, &h->ri[r], sizeof(RepInfo), hipMemcpyHostToDevice));
    // the step parameters (a_eff, sigma) of the current step follow at once
    launch_step_params(h, 0, -1);
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

struct StateSnap {
    double2 *xy = nullptr, *v = nullptr, *f = nullptr;
    float2 *vt = nullptr;
    uint32_t *gt = nullptr;
    int8_t *mp = nullptr;
    int32_t *lnk = nullptr, *cs = nullptr, *src = nullptr;
    RepStep *rs = nullptr;
    // tape states: previous chain beads, the RDF gradient and the RDF's type
    // list (TapeRef)
    int32_t *prv = nullptr;
    double2 *rg = nullptr, *lxy = nullptr;
    uint32_t *lgt = nullptr;
    int32_t *lidx = nullptr, *lcs = nullptr;
};

struct Adj {
    int K = 0, nseg = 0, nbins = 0;
    std::vector<StateSnap> tape, ckpt;
    TapeRef *d_tape = nullptr;  // device copy of the tape's TapeRef
    int2 *d_cols = nullptr;     // device copy of h->cols
    StateSnap end;
    double2 *X[2] = {nullptr, nullptr}, *V[2] = {nullptr, nullptr}, *F[2] = {nullptr, nullptr};
    RepStep *rsa = nullptr;  // step parameters of a segment's reverse steps
    double *thp = nullptr, *gw = nullptr, *grad = nullptr;
    unsigned long long *rpart = nullptr, *racc = nullptr;
    int64_t *base = nullptr;
    std::vector<void *> mem;
    // the last forward window
    bool ready = false;
    int nsteps = 0;
    RdfArgs RA;
    std::vector<double> Cb;
};

static void adj_free(oc_engine *h) {
    if (!h->adj) return;
    for (void *p : h->adj->mem) (void)hipFree(p);
    delete h->adj;
    h->adj = nullptr;
}

template <class T>
static int aalloc(Adj *a, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(OC_ENOMEM, std::string("hipMalloc failed (window gradient): ") + hipGetErrorString(e));
    a->mem.push_back(q);
    *p = (T *)q;
    return 0;
}

// full: also v, f and the step counters (checkpoints); tape: also the sort
// map, prv, the RDF gradient and the type list
static int snap_alloc(oc_engine *h, Adj *a, StateSnap &s, bool full, bool tape) {
    const size_t N = h->Ntot, C = h->Ctot + 1;
    int rc = 0;
    rc = rc ? rc : aalloc(a, &s.xy, N);
    rc = rc ? rc : aalloc(a, &s.vt, N);
    rc = rc ? rc : aalloc(a, &s.gt, N);
    rc = rc ? rc : aalloc(a, &s.mp, N);
    rc = rc ? rc : aalloc(a, &s.lnk, N);
    rc = rc ? rc : aalloc(a, &s.cs, C);
    if (tape) {
        rc = rc ? rc : aalloc(a, &s.src, N);
        rc = rc ? rc : aalloc(a, &s.prv, N);
        rc = rc ? rc : aalloc(a, &s.rg, N);
        rc = rc ? rc : aalloc(a, &s.lxy, N);
        rc = rc ? rc : aalloc(a, &s.lgt, N);
        rc = rc ? rc : aalloc(a, &s.lidx, N);
        rc = rc ? rc : aalloc(a, &s.lcs, h->Ctot + h->cols.size());
    }
    if (full) {
        rc = rc ? rc : aalloc(a, &s.v, N);
        rc = rc ? rc : aalloc(a, &s.f, N);
        rc = rc ? rc : aalloc(a, &s.rs, h->nrep);
    }
    return rc;
}

// copy the engine's current state into (save) or out of (restore) a snapshot
static int snap_copy(oc_engine *h, StateSnap &s, bool save) {
    // all arrays in one launch (k_copy_multi)
    const size_t N = h->Ntot, C = h->Ctot + 1;
    const int c = h->cur;
    CopyJobs J;
    J.n = 0;
    auto add = [&](void *eng, void *snp, size_t bytes) {
        J.dst[J.n] = (char *)(save ? snp : eng);
        J.src[J.n] = (const char *)(save ? eng : snp);
        J.bytes[J.n] = (int64_t)bytes;
        J.n++;
    };
    add(h->st[c].xy, s.xy, N * sizeof(double2));
    add(h->st[c].vt, s.vt, N * sizeof(float2));
    add(h->st[c].gt, s.gt, N * sizeof(uint32_t));
    add(h->st[c].mp, s.mp, N);
    add(h->st[c].lnk, s.lnk, N * sizeof(int32_t));
    add(h->d_cs[c], s.cs, C * sizeof(int32_t));
    if (s.v) {
        add(h->st[c].v, s.v, N * sizeof(double2));
        add(h->d_f, s.f, N * sizeof(double2));
        add(h->d_rs, s.rs, h->nrep * sizeof(RepStep));
    }
    k_copy_multi<<<512, 256, 0, h->stream>>>(J);
    if (!save) h->keys_fresh = false;
    HIPCHK(hipGetLastError());
    return 0;
}

// The window's steps write their states straight into tape buffers: before a
// step, the next state's positions, gt words and cell starts (with full also
// vt, mp, lnk, prv and the sort map) are pointed at tape state t.  The engine's
// own buffers are put back with restore_bufs.
struct EngBufs {
    StateP st[2];
    int32_t *cs[2];
};

static EngBufs save_bufs(const oc_engine *h) {
    EngBufs b;
    for (int k = 0; k < 2; k++) {
        b.st[k] = h->st[k];
        b.cs[k] = h->d_cs[k];
    }
    return b;
}

static void restore_bufs(oc_engine *h, const EngBufs &b) {
    for (int k = 0; k < 2; k++) {
        h->st[k] = b.st[k];
        h->d_cs[k] = b.cs[k];
    }
    h->rec_src = nullptr;
}

static void step_into(oc_engine *h, const StateSnap &t, bool full) {
    const int n = h->cur ^ 1;
    h->st[n].xy = t.xy;
    h->st[n].gt = t.gt;
    h->d_cs[n] = t.cs;
    if (full) {
        h->st[n].vt = t.vt;
        h->st[n].mp = t.mp;
        h->st[n].lnk = t.lnk;
        h->st[n].prv = t.prv;
        h->rec_src = t.src;
    }
    enqueue_step(h);
}

// The reverse step runs over the pair tiles when they fill the chip (at
// least kAdjTileMin of them), else over bead blocks with four lanes per bead.
constexpr int kAdjTileMin = 1024;
static bool adj_tiled(const oc_engine *h) { return (int)h->tiles.size() >= kAdjTileMin; }

// one reverse step over the pair tiles (their shape follows the pair block)
static void launch_adj_tile(oc_engine *h, const AdjStep &A) {
    if (!adj_tiled(h)) {
        k_adj_rstep<<<(int)h->bblk.size() * kRdfSub, kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, A);
        return;
    }
    const dim3 grid(8 * h->per_xcd);
    const int nt = (int)h->tiles.size();
    switch (h->pair_block) {
    case 64:
        k_adj_tile<64, 224><<<grid, 64, 0, h->stream>>>(h->d_ri, h->d_tiles, nt, h->per_xcd, A, h->d_err);
        break;
    case 128:
        k_adj_tile<128, 448><<<grid, 128, 0, h->stream>>>(h->d_ri, h->d_tiles, nt, h->per_xcd, A, h->d_err);
        break;
    default:
        k_adj_tile<256, 896><<<grid, 256, 0, h->stream>>>(h->d_ri, h->d_tiles, nt, h->per_xcd, A, h->d_err);
        break;
    }
}

static int adj_setup(oc_engine *h, int K, int nseg, int nbins) {
    if (h->adj && h->adj->K == K && h->adj->nseg == nseg && h->adj->nbins == nbins) return 0;
    adj_free(h);
    Adj *a = new Adj();
    h->adj = a;
    a->K = K;
    a->nseg = nseg;
    a->nbins = nbins;
    const size_t N = h->Ntot;
    int rc = 0;
    a->tape.resize(K + 1);
    for (auto &t : a->tape) rc = rc ? rc : snap_alloc(h, a, t, false, true);
    rc = rc ? rc : aalloc(a, &a->d_tape, K + 1);
    rc = rc ? rc : aalloc(a, &a->d_cols, h->cols.size());
    if (!rc) {
        std::vector<TapeRef> tr(K + 1);
        for (int m = 0; m <= K; m++) {
            const StateSnap &t = a->tape[m];
            tr[m] = TapeRef{t.cs, t.xy, t.gt, t.rg, t.lxy, t.lgt, t.lidx, t.lcs};
        }
        if (hipMemcpy(a->d_tape, tr.data(), tr.size() * sizeof(TapeRef), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(a->d_cols, h->cols.data(), h->cols.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(OC_EHIP, "tape table copy failed");
    }
    a->ckpt.resize(nseg);
    for (auto &t : a->ckpt) rc = rc ? rc : snap_alloc(h, a, t, true, false);
    rc = rc ? rc : snap_alloc(h, a, a->end, true, false);
    for (int k = 0; k < 2; k++) {
        rc = rc ? rc : aalloc(a, &a->X[k], N);
        rc = rc ? rc : aalloc(a, &a->V[k], N);
        rc = rc ? rc : aalloc(a, &a->F[k], N);
    }
    rc = rc ? rc : aalloc(a, &a->rsa, (size_t)K * h->nrep);
    rc = rc ? rc : aalloc(a, &a->thp, std::max(h->tiles.size(), h->bblk.size() * (size_t)kRdfSub) * 8);
    rc = rc ? rc : aalloc(a, &a->gw, (size_t)h->nrep * nbins);
    rc = rc ? rc : aalloc(a, &a->grad, (size_t)h->nrep * 8);
    rc = rc ? rc : aalloc(a, &a->rpart, h->bblk.size() * (size_t)K * nbins);
    rc = rc ? rc : aalloc(a, &a->racc, (size_t)h->nrep * nbins);
    rc = rc ? rc : aalloc(a, &a->base, h->nrep);
    if (rc) adj_free(h);
    return rc;
}

static void zero_f64(oc_engine *h, double *p, size_t n) {
    k_zero_f64<<<grid_of((int64_t)n, 256), 256, 0, h->stream>>>(p, (int64_t)n);
}

OC_API int oc_window_forward(oc_engine *h, int32_t nsteps, int32_t ckpt_every, const oc_rdf_spec *spec,
                             double *gbar) {
    if (!h || !spec || nsteps < 1) return fail(OC_EINVAL, "bad arguments");
    if (!h->have_state) return fail(OC_ESTATE, "no state");
    if (spec->nbins < 1 || spec->nbins > 256 || !(spec->rmax > 0) || !(spec->width > 0))
        return fail(OC_EINVAL, "rdf spec: 1..256 bins, rmax > 0, width > 0");
    if (spec->ta < 0 || spec->ta >= OC_NTYPES || spec->tb < 0 || spec->tb >= OC_NTYPES)
        return fail(OC_EINVAL, "rdf spec: bad types");
    const float cut = (float)(spec->rmax + spec->width);
    int Rc = 0;
    for (int r = 0; r < h->nrep; r++) {
        const RepInfo &R = h->ri[r];
        const oc_params &P = h->prm[r];
        if (P.k_angle != 0.0 || P.stir_A != 0.0 || P.kT_amp != 0.0 || P.chem)
            return fail(OC_EINVAL, "window gradient needs k_angle = 0 and no stirring, kT cycle or chemistry");
        if (h->ri[r].arch != OC_ARCH_LINEAR) return fail(OC_EINVAL, "window gradient needs linear lipids");
        if (h->ri[r].rchem)
            return fail(OC_EINVAL, "window gradient: the reference chemistry (chem = 1) is not differentiated");
        Rc = std::max(Rc, (int)ceil(cut * R.icsx));
    }
    if (2 * Rc + 1 > kRdfMaxCol) return fail(OC_EINVAL, "rdf cutoff longer than four cell columns");
    for (int r = 0; r < h->nrep; r++)
        if (2 * Rc + 1 > h->ri[r].ncx || 2.0 * cut + 2.0 > h->ri[r].L)
            return fail(OC_EINVAL, "box too small for the RDF cutoff");
    const int K = ckpt_every > 0 ? std::min(ckpt_every, nsteps) : nsteps;
    const int nseg = (nsteps + K - 1) / K;
    const int nb = spec->nbins;
    if (int rc = adj_setup(h, K, nseg, nb)) return rc;
    Adj *a = h->adj;
    a->ready = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<RepStep> rs0(h->nrep);
    HIPCHK(hipMemcpy(rs0.data(), h->d_rs, h->nrep * sizeof(RepStep), hipMemcpyDeviceToHost));
    std::vector<int64_t> base(h->nrep);
    for (int r = 0; r < h->nrep; r++) {
        base[r] = rs0[r].step;
        if (rs0[r].step < h->ri[r].warm_end) return fail(OC_EINVAL, "window gradient inside the warm-up ramp");
    }
    HIPCHK(hipMemcpy(a->base, base.data(), h->nrep * sizeof(int64_t), hipMemcpyHostToDevice));
    // RDF normalisation per replica from its type counts
    std::vector<uint32_t> gth(h->Ntot);
    HIPCHK(hipMemcpy(gth.data(), h->st[h->cur].gt, h->Ntot * 4, hipMemcpyDeviceToHost));
    a->Cb.assign((size_t)h->nrep * nb, 0.0);
    const double dr = spec->rmax / nb;
    for (int r = 0; r < h->nrep; r++) {
        const RepInfo &R = h->ri[r];
        int64_t na = 0, nbb = 0;
        for (int k = 0; k < R.N; k++) {
            const uint32_t t = gth[R.off + k] >> kGidBits;
            na += (int)t == spec->ta;
            nbb += (int)t == spec->tb;
        }
        if (na == 0 || nbb == 0) return fail(OC_EINVAL, "rdf spec: a replica has no bead of the pair's types");
        const double area = R.L * R.L;
        const double norm = spec->ta == spec->tb ? 2.0 / (na * (na / area)) : 1.0 / (na * (nbb / area));
        for (int b = 0; b < nb; b++) {
            const double r0 = b * dr, r1 = (b + 1) * dr;
            a->Cb[(size_t)r * nb + b] = dr * norm / (M_PI * (r1 * r1 - r0 * r0));
        }
    }
    RdfArgs &RA = a->RA;
    RA.ta = spec->ta;
    RA.tb = spec->tb;
    RA.nbins = nb;
    RA.rmax = (float)spec->rmax;
    RA.width = (float)spec->width;
    RA.dr = (float)dr;
    RA.idr = (float)(1.0 / dr);
    RA.cut = cut;
    RA.Rc = Rc;
    const int nbk = (int)h->bblk.size(), ncol = (int)h->cols.size();
    // f0 = F(x0, v~0; theta), then nsteps steps, checkpoints at the segment
    // starts; each segment's states are stepped into the tape and their RDF
    // histograms taken in one launch
    drop_graphs(h);
    launch_step_params(h, 0, -1);
    launch_forces(h, true);
    HIPCHK(hipMemsetAsync(a->racc, 0, (size_t)h->nrep * nb * sizeof(unsigned long long), h->stream));
    const EngBufs eb = save_bufs(h);
    for (int seg = 0; seg < nseg; seg++) {
        const int len = std::min(K, nsteps - seg * K);
        if (int rc = snap_copy(h, a->ckpt[seg], true)) return rc;
        // whole states (with the sort maps) into the tape: the last segment's
        // tape is then the one the backward sweep starts from
        for (int m = 1; m <= len; m++) step_into(h, a->tape[m], true);
        restore_bufs(h, eb);
        // the engine's buffers take the segment's last state back
        {
            const StateSnap &t = a->tape[len];
            const int c = h->cur;
            const size_t N = h->Ntot;
            CopyJobs J;
            J.n = 0;
            auto add = [&](void *d, const void *sp, size_t bytes) {
                J.dst[J.n] = (char *)d;
                J.src[J.n] = (const char *)sp;
                J.bytes[J.n] = (int64_t)bytes;
                J.n++;
            };
            add(h->st[c].xy, t.xy, N * sizeof(double2));
            add(h->st[c].gt, t.gt, N * sizeof(uint32_t));
            add(h->d_cs[c], t.cs, (h->Ctot + 1) * sizeof(int32_t));
            add(h->st[c].vt, t.vt, N * sizeof(float2));
            add(h->st[c].mp, t.mp, N);
            add(h->st[c].lnk, t.lnk, N * sizeof(int32_t));
            add(h->st[c].prv, t.prv, N * sizeof(int32_t));
            k_copy_multi<<<512, 256, 0, h->stream>>>(J);
        }
        k_tlist<<<dim3(ncol, len), 256, 0, h->stream>>>(h->d_ri, a->d_cols, a->d_tape + 1, RA, seg == nseg - 1);
        k_rdf<<<dim3(nbk, len), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, a->d_tape + 1, RA, a->rpart);
        k_rdf_reduce<<<dim3(nb, h->nrep), 256, 0, h->stream>>>(h->d_ri, nbk, len, a->rpart, a->racc, nb);
    }
    if (int rc = snap_copy(h, a->end, true)) return rc;
    HIPCHK(hipGetLastError());
    if (int rc = check_err(h)) return rc;
    std::vector<unsigned long long> acc((size_t)h->nrep * nb);
    HIPCHK(hipMemcpy(acc.data(), a->racc, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (gbar)
        for (size_t k = 0; k < acc.size(); k++) gbar[k] = a->Cb[k] * ((double)acc[k] / kFix) / nsteps;
    a->nsteps = nsteps;
    a->ready = true;
    return 0;
}

OC_API int oc_window_backward(oc_engine *h, const double *dL_dgbar, double *grad) {
    if (!h || !dL_dgbar) return fail(OC_EINVAL, "bad arguments");
    Adj *a = h->adj;
    if (!a || !a->ready) return fail(OC_ESTATE, "no forward window to differentiate");
    const int nb = a->nbins, K = a->K, nseg = a->nseg, nsteps = a->nsteps;
    std::vector<double> gw((size_t)h->nrep * nb);
    for (size_t k = 0; k < gw.size(); k++) gw[k] = dL_dgbar[k] / nsteps * a->Cb[k];
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(a->gw, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice));
    const RdfArgs RA = a->RA;
    const int nbk = (int)h->bblk.size(), ncol = (int)h->cols.size();
    const size_t N = h->Ntot;
    zero_f64(h, a->thp, std::max(h->tiles.size(), h->bblk.size() * (size_t)kRdfSub) * 8);
    int ab = 0;  // adjoint buffer parity
    const EngBufs eb = save_bufs(h);
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int n0 = seg * K, n1 = std::min(n0 + K, nsteps), len = n1 - n0;
        // recompute the segment from its checkpoint (bit-identical) into the
        // tape; the forward window left the last segment's tape (with its type
        // lists) in place, so only segment 0 of a one-segment window is redone
        if (seg != nseg - 1 || seg == 0) {
            if (int rc = snap_copy(h, a->ckpt[seg], false)) return rc;
            launch_build_prv(h);
            if (seg == 0)
                if (int rc = snap_copy(h, a->tape[0], true)) return rc;
            for (int m = 1; m <= len; m++) step_into(h, a->tape[m], true);
            restore_bufs(h, eb);
            k_tlist<<<dim3(ncol, len), 256, 0, h->stream>>>(h->d_ri, a->d_cols, a->d_tape + 1, RA, 1);
        }
        // d loss / d x of the RDF at the segment's states n0+1..n1, one launch
        k_rdf_adj<<<dim3(nbk, len), kBeadBlock, 0, h->stream>>>(h->d_ri, h->d_bblk, a->d_tape + 1, RA, a->gw);
        if (seg == nseg - 1) {
            zero_f64(h, (double *)a->X[ab], 2 * N);
            zero_f64(h, (double *)a->V[ab], 2 * N);
            zero_f64(h, (double *)a->F[ab], 2 * N);
        }
        // one launch per reverse step (pair and bond terms, the RDF gradient, the
        // step back through the sort map)
        k_step_params_seg<<<dim3(h->nrep, len), 64, 0, h->stream>>>(h->d_ri, a->rsa, n0 + 1, a->base);
        for (int m = len; m >= 1; m--) {
            const StateSnap &t = a->tape[m];
            const AdjStep A = {a->rsa + (size_t)(m - 1) * h->nrep, t.cs, t.src, t.lnk, t.prv, t.xy, t.rg, t.vt, t.gt,
                               t.mp, a->F[ab], a->X[ab], a->V[ab], a->X[ab ^ 1], a->V[ab ^ 1], a->F[ab ^ 1], a->thp};
            launch_adj_tile(h, A);
            ab ^= 1;
        }
    }
    // the window's first force, f0 = F(x0, v~0; theta), carries parameter gradient too
    {
        const StateSnap &t = a->tape[0];
        k_step_params_seg<<<dim3(h->nrep, 1), 64, 0, h->stream>>>(h->d_ri, a->rsa, 0, a->base);
        const AdjStep A = {a->rsa, t.cs, nullptr, t.lnk, nullptr, t.xy, nullptr, t.vt, t.gt, t.mp, a->F[ab], a->X[ab],
                           a->V[ab], nullptr, nullptr, nullptr, a->thp};
        launch_adj_tile(h, A);
    }
    k_theta_reduce<<<h->nrep, 256, 0, h->stream>>>(adj_tiled(h) ? h->d_trng : h->d_brng, a->thp, a->grad);
    // the engine continues from the window end
    if (int rc = snap_copy(h, a->end, false)) return rc;
    launch_build_prv(h);
    HIPCHK(hipGetLastError());
    if (int rc = check_err(h)) return rc;
    std::vector<double> g8((size_t)h->nrep * 8);
    HIPCHK(hipMemcpy(g8.data(), a->grad, g8.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (grad)
        for (int r = 0; r < h->nrep; r++)
            for (int q = 0; q < 7; q++) grad[(size_t)r * 7 + q] = g8[(size_t)r * 8 + q];
    return 0;
}

OC_API int oc_window_rewind(oc_engine *h) {
    if (!h) return fail(OC_EINVAL, "null engine");
    Adj *a = h->adj;
    if (!a || !a->ready || a->ckpt.empty()) return fail(OC_ESTATE, "no forward window to rewind to");
    HIPCHK(hipStreamSynchronize(h->stream));
    // the window's first checkpoint: the state (and step counters) it started from
    if (int rc = snap_copy(h, a->ckpt[0], false)) return rc;
    launch_build_prv(h);
    HIPCHK(hipGetLastError());
    return check_err(h);
}

OC_API int oc_window_grad(oc_engine *h, int32_t nsteps, int32_t ckpt_every, const oc_rdf_spec *spec,
                          const double *target, int32_t target_stride, double *loss, double *grad, double *gbar) {
    if (!h || !spec || !target) return fail(OC_EINVAL, "bad arguments");
    const int nb = spec->nbins;
    std::vector<double> g((size_t)h->nrep * std::max(nb, 1));
    if (int rc = oc_window_forward(h, nsteps, ckpt_every, spec, g.data())) return rc;
    std::vector<double> dl(g.size());
    for (int r = 0; r < h->nrep; r++) {
        double Ls = 0.0;
        for (int b = 0; b < nb; b++) {
            const size_t k = (size_t)r * nb + b;
            const double d = g[k] - target[(size_t)r * target_stride + b];
            Ls += d * d;
            dl[k] = 2.0 * d;
        }
        if (loss) loss[r] = Ls;
    }
    if (gbar) std::copy(g.begin(), g.end(), gbar);
    return oc_window_backward(h, dl.data(), grad);
}

// --------------------------------------------------------------------------- kinetic_T on device arrays

// Kinetic temperature and momentum of n beads held by the caller in device
// memory (kinetic_T, cel/engine/oercel.c:2058-2071): v = vh + dt/2 f,
// out = (sum |v|^2 / (2n - 2), sum v_x, sum v_y).  FP64 over a fixed
// partition (kKinChunk beads per block, a fixed tree in the block and over the
// blocks), so the same inputs give the same bits.
constexpr int kKinChunk = 4096;

__global__ __launch_bounds__(256) void k_kin_partial(const double *__restrict__ vx, const double *__restrict__ vy,
                                                     const double *__restrict__ fx, const double *__restrict__ fy,
                                                     int64_t n, double hdt, double *__restrict__ part) {
    __shared__ double red[3][256];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int64_t b0 = (int64_t)blockIdx.x * kKinChunk;
    for (int q = 0; q < kKinChunk / 256; q++) {
        const int64_t i = b0 + q * 256 + threadIdx.x;
        if (i >= n) break;
        const double ux = vx[i] + hdt * fx[i], uy = vy[i] + hdt * fy[i];
        a0 += ux * ux + uy * uy;
        a1 += ux;
        a2 += uy;
    }
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    red[2][threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[(size_t)blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_kin_final(const double *__restrict__ part, int nb, int64_t n,
                                                   double *__restrict__ out) {
    __shared__ double red[3][256];
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 3; k++) a[k] += part[(size_t)b * 3 + k];
    for (int k = 0; k < 3; k++) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0] / (2.0 * (double)n - 2.0);
        out[1] = red[1][0];
        out[2] = red[2][0];
    }
}

static int64_t kin_blocks(int64_t n) { return (n + kKinChunk - 1) / kKinChunk; }

static void launch_kin(const double *vx, const double *vy, const double *fx, const double *fy, int64_t n, double dt,
                       double *out, double *part, hipStream_t st) {
    const int nb = (int)kin_blocks(n);
    k_kin_partial<<<nb, 256, 0, st>>>(vx, vy, fx, fy, n, 0.5 * dt, part);
    k_kin_final<<<1, 256, 0, st>>>(part, nb, n, out);
}

OC_API size_t oc_kinetic_dev_work(int64_t n) {
    return n < 2 ? 0 : (size_t)kin_blocks(n) * 3 * sizeof(double);
}

OC_API int oc_kinetic_dev(const double *vx, const double *vy, const double *fx, const double *fy, int64_t n,
                          double dt, double *out, void *work, size_t work_bytes, void *stream) {
    if (!vx || !vy || !fx || !fy || !out) return fail(OC_EINVAL, "null array");
    if (n < 2) return fail(OC_EINVAL, "kinetic temperature needs n >= 2 beads");
    if (kin_blocks(n) > INT32_MAX / 3) return fail(OC_EINVAL, "too many beads");
    if (!work || work_bytes < oc_kinetic_dev_work(n))
        return fail(OC_EINVAL, "workspace smaller than oc_kinetic_dev_work(n) bytes");
    launch_kin(vx, vy, fx, fy, n, dt, out, (double *)work, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

// --------------------------------------------------------------------------- allocation-free kinetic_T

// Block partials of oc_kinetic_dev_noalloc: up to kKinScratchBlocks blocks,
// i.e. n <= kKinScratchBlocks * kKinChunk = 67,108,864 beads, 384 KiB.  Each
// GPU the library runs on holds its own copy until the process ends; every
// call on a GPU uses that GPU's copy (calls that overlap on one GPU race).
// Every partial a call reads is written by the same call first: nothing
// needs clearing between calls.
constexpr int kKinScratchBlocks = 16384;
static __device__ __align__(256) double scratch[3 * kKinScratchBlocks];

// Device address of scratch on the current device, looked up once per device
// (host code uses this address, never the array's name).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

static int scratch_ptr(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OC_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(OC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

OC_API int oc_kinetic_dev_noalloc(const double *vx, const double *vy, const double *fx, const double *fy, int64_t n,
                                  double dt, double *out, void *stream) {
    if (!vx || !vy || !fx || !fy || !out) return fail(OC_EINVAL, "null array");
    if (n < 2) return fail(OC_EINVAL, "kinetic temperature needs n >= 2 beads");
    if (kin_blocks(n) > kKinScratchBlocks)
        return fail(OC_EINVAL, "oc_kinetic_dev_noalloc: n exceeds 67,108,864 beads (its fixed scratch); use "
                               "oc_kinetic_dev with a workspace");
    double *part = nullptr;
    if (int rc = scratch_ptr(&part)) return rc;
    launch_kin(vx, vy, fx, fy, n, dt, out, part, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}
