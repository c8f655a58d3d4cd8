/*
 * oercel_mi.h -- C ABI of the MI355X (gfx950) 2D DPD step library.
 *
 * The reference engine (cel/engine/oercel.c) has no FFI: its step is a set of
 * static C functions called from main()'s loop (cel/engine/oercel.c:2300-2362)
 * and, in its tests, in-process via "#define main ...; #include"
 * (cel/engine/tests/forcetest.c:3-5).  This header is the boundary a caller of
 * that step binds instead: plain C types, pointers and sizes, integer return
 * codes (0 = ok) plus oc_last_error(), one engine handle per set of replicas,
 * one HIP stream per handle, called from one host thread.
 *
 * A single system is an ensemble of one replica; every entry point that takes
 * a replica index accepts 0 for it.
 *
 * State exchanged with the caller is in gid (permanent bead id) order, the
 * order the reference's snapshots and checkpoints use
 * (cel/engine/oercel.c:1830-1936); inside the library it lives in cell-sorted
 * slot order, re-sorted every step.
 */
#ifndef OERCEL_MI_H
#define OERCEL_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_ABI_VERSION 4
#define OC_NTYPES 7
#define OC_MAXMLEN 12

/* bead types, cel/engine/oercel.c:47-49.  In the precursor chemistry the
 * precursor monomer P is type 3 (the reference's N row of a_ij). */
enum { OC_W = 0, OC_H = 1, OC_T = 2, OC_N = 3, OC_KA = 4, OC_KB = 5, OC_Z = 6 };

/* error codes */
enum {
    OC_OK = 0,
    OC_EINVAL = 1,   /* bad argument or parameter */
    OC_ENOMEM = 2,   /* device or host allocation failed */
    OC_EHIP = 3,     /* HIP runtime error */
    OC_EDEVICE = 4,  /* sticky device-side error (event overflow, broken chain, ...) */
    OC_ESTATE = 5    /* call made in the wrong state (e.g. step before any state) */
};

/* event kinds: the precursor chemistry (no counterpart in the reference), then
 * the reference chemistry's decay and food supply (its dlip, dN and feed
 * events, cel/engine/oercel.c:658-659; feed is reported per converted bead) */
enum { OC_EV_CONV = 0, OC_EV_FORM = 1, OC_EV_BREAK = 2, OC_EV_DLIP = 3, OC_EV_DN = 4, OC_EV_FEED = 5,
       OC_EV_SYN = 6, OC_EV_REP = 7, OC_EV_MUT = 8, OC_EV_DK = 9 };

/*
 * Parameters of one replica: the numeric subset of the reference's Params
 * (cel/engine/oercel.c:178-221) that the DPD step reads, after the ini/key=value
 * rules of set_param/finalize_params (cel/engine/oercel.c:338-395, 578-619) have
 * been applied by the host (the Python package mirrors those rules).
 * Lipids follow the template of lip_arch (linear, pi, y or bola); the beads
 * of a lipid are chained through nxt in template order (link_lipid,
 * cel/engine/oercel.c:801-806).
 */
typedef struct oc_params {
    /* physics */
    double L, rho, dt, lambda, gamma, kT;
    double k_bond, r0, k_angle;
    double a[8][8];            /* symmetric repulsion matrix; row/column 7 unused */
    double stir_A;             /* Kolmogorov stirring amplitude (0 = off) */
    double kT_amp, kT_period;  /* temperature cycle kT(t) = kT + kT_amp sin(2 pi t / kT_period) */
    int32_t stir_n;
    int32_t nsub;              /* y sub-cells per unit cell, 1..8 (performance knob) */
    int32_t lip_nH, lip_nT;
    int64_t warmup;            /* repulsion ramp length (steps), cel/engine/oercel.c:1122-1129 */
    int64_t warm_end;          /* step at which the ramp ends (0 = none) */
    uint64_t seed;
    /* soup composition (init only), cel/engine/oercel.c:1774-1814 */
    double f_lip, f_Ka, f_Kb, c_N;
    /* precursor chemistry: P -> T near a lipid tail, T-T bond form / break */
    int32_t chem;              /* 1 = on */
    int32_t n_chem;            /* check interval (steps); probabilities are per check */
    double p_conv, r_conv;     /* P within r_conv of a lipid T becomes a free T */
    double p_form, r_form;     /* two free T within r_form bond */
    double p_break;            /* a chemistry bond breaks */
    /* time axis: t(step) = t_anchor + (step - step_anchor) * dt */
    double t_anchor;
    int64_t step_anchor;
    /* lipid template, mol_build (cel/engine/oercel.c:479-576): OC_ARCH_*.
     * linear: lip_nH heads then lip_nT tails; pi: two bonded heads, each with a
     * tail of lip_nT beads; y: one head with two tails; bola: a head pair and a
     * single head joined by two tails (a macrocycle). */
    int32_t lip_arch;
    int32_t lip_tie;           /* pi / y / bola: ladder rungs between the two tails */
    double r0_head;            /* pi / bola: rest length of the head-head bond */
    double k_head, cos_head;   /* junction angle terms U = k_head (cos - cos_head)^2 (0 = none) */
    /* the reference's chemistry (its chem = 1, cel/engine/oercel.c:1565-1756):
     * at every check the catalyst reactions (below), then lipid decay (a lipid
     * becomes water, p_ld per molecule), catalyst decay (p_kd) and nutrient
     * decay (N -> W, p_nd), then the food supply (every eligible W -> N,
     * p_feed, capped by feed_cap) -- per-bead draws of hash_u01 with salts
     * 11, 12, 13, 14 */
    int32_t rchem;             /* 1 = on */
    int32_t feed_mode;         /* OC_FEED_NONE / UNIFORM / VENT */
    double p_ld, p_nd, p_feed;
    int32_t vent_shape;        /* OC_VENT_DISK / VSTRIP / HSTRIP */
    int32_t pad2;
    double vent_x, vent_y, vent_r, vent_w;  /* fixed vent region (feed = vent) */
    /* the catalyst chemistry (rchem = 1 with Ka / Kb beads), in the reference
     * GPU port's hash-decided form (cel/engine_gpu/oercel_gpu.cu:1645-1709,
     * 1746-1880): a catalyst draws synthesis / replication from per-catalyst
     * hashes (salts 21, 22, mutation 23); the active ones react one after the
     * other in the order of a hash priority (salt 24), each seeing the
     * reactions before it.  Synthesis turns the syn_nN (0 = the lipid length)
     * nearest N within r_cat (then the nearest W) into a new lipid through
     * their centroid, the orientation of least overlap among syn_tries
     * (salts 40 + try), relaxed by syn_relax descent steps; replication turns
     * the next two N into a catalyst of the same kind (the other with
     * probability mu) and a W.  Catalyst decay p_kd (salt 12); feed_cap > 0
     * keeps, when the feed would exceed feed_cap * N nutrients, the
     * conversions with the smallest hash (salt 25) keys. */
    double p_syn_a, p_syn_b, p_rep_a, p_rep_b, mu, p_kd;
    double r_cat, feed_cap;
    int32_t syn_tries, syn_relax, syn_nN, pad3;
} oc_params;

enum { OC_FEED_NONE = 0, OC_FEED_UNIFORM = 1, OC_FEED_VENT = 2 };
enum { OC_VENT_DISK = 0, OC_VENT_VSTRIP = 1, OC_VENT_HSTRIP = 2 };

enum { OC_ARCH_LINEAR = 0, OC_ARCH_PI = 1, OC_ARCH_Y = 2, OC_ARCH_BOLA = 3 };

/*
 * Per-bead state in gid order.  Any pointer may be NULL in oc_get_state (that
 * field is skipped); oc_set_state needs all but fx/fy (NULL = zero forces).
 *   vx, vy : half-step velocity vh = v - dt/2 f (cel/engine/oercel.c:2207-2211)
 *   mpos   : position in the lipid (0 = first head), -1 = not a lipid bead
 *   nxt    : gid of the next bead of the lipid chain, -1 = none
 *   bond   : gid of the chemistry-bond partner, -1 = none (only for mpos = -1)
 */
typedef struct oc_state {
    double *x, *y, *vx, *vy, *fx, *fy;
    uint8_t *type;
    int8_t *mpos;
    int32_t *nxt;
    int32_t *bond;
} oc_state;

/* one chemistry event.  conv / form / break: the two beads.  dlip / dN / feed:
 * gid_a = gid_b = the bead (dlip: the lipid's first bead), at (x, y).  syn:
 * gid_a = the catalyst, gid_b = its kind (0 = Ka, 1 = Kb), (x, y) = the
 * catalyst; rep / mut: gid_a = the new catalyst, gid_b = its kind, (x, y) =
 * the parent catalyst; dK: gid_a = the catalyst, gid_b = its kind, at (x, y). */
typedef struct oc_event {
    int64_t step;
    int32_t kind;
    uint32_t gid_a, gid_b;
    int32_t pad;
    double x, y;               /* 0 for conv / form / break */
} oc_event;

/* stats of one replica (cel/engine/oercel.c:2053-2071 kinetic_T/count_types,
 * plus an energy the reference does not compute) */
typedef struct oc_stats {
    int64_t step;
    double time;
    double T_kin;       /* sum (vh + dt/2 f)^2 / (2N - 2) */
    double px, py;      /* total momentum of the full velocity */
    double E_kin;       /* 1/2 sum v^2 */
    double E_pair;      /* sum_{i<j} a_ij/2 (1 - r)^2 at the current warm-up ramp */
    double E_bond;      /* harmonic bonds + angle terms */
    double max_bond;    /* running max bond length since the last oc_stats call */
    int64_t count[8];   /* beads per type */
    int64_t n_lipids;   /* beads with mpos == 0 */
    int64_t n_bonds;    /* chemistry bonds */
    double virial;      /* sum_{i<j} r_ij . F^C_ij of the conservative pair forces (current ramp) */
    double pressure;    /* N T_kin / L^2 + virial / (2 L^2): the 2D virial pressure of the
                         * reference's water check (cel/engine_gpu/tests/compare.py:82-127) */
} oc_stats;

typedef struct oc_engine oc_engine;

const char *oc_last_error(void);
int oc_abi_version(void);

/* Engine over nrep replicas; p[r] and N[r] describe replica r.  Allocates all
 * device memory up front (nothing is allocated by oc_step).  The chemistry
 * event buffer holds max(65536, 4 N) events (at most 8M; the environment
 * variable OC_EVENT_CAP lowers it); oc_get_events drains it, and events past
 * a full buffer are dropped and reported once by the next oc_get_events.
 * Replaces sys_alloc (cel/engine/oercel.c:714-759) for one or many systems. */
int oc_create(const oc_params *p, const int32_t *N, int32_t nrep, oc_engine **out);
void oc_destroy(oc_engine *h);

/* Tuning knobs (0 = library default): pair-kernel workgroup size (64/128/256)
 * and target i-beads per pair tile. */
int oc_set_tuning(oc_engine *h, int32_t pair_block, int32_t tile_beads);

/* Pair-kernel variant (no counterpart in the reference, whose pair loop has
 * one form): 0 evaluates every candidate in one pass per lane; 1 (the
 * default) tests every candidate first and evaluates each lane's marked
 * interacting pairs in a second pass; 2 queues the interacting pairs of all
 * lanes of a wavefront and evaluates them 64 at a time, each lane summing its
 * own pairs back in candidate order; 3 lists each lane's interacting pairs in
 * LDS.  0, 1 and 2 give bit-identical forces for windows of up to 64
 * candidates (1 sums the pairs past a full mask first, in another FP64
 * rounding order).  Other values: OC_EINVAL. */
int oc_set_pair_variant(oc_engine *h, int32_t lists);

/* Host-side soup of replica r in gid order, exactly as init_soup +
 * set_velocities (cel/engine/oercel.c:1757-1814) with rng_seed(seed)
 * (cel/engine/oercel.c:2193-2195); velocities are full velocities v.
 * Needs no GPU. */
int oc_soup_host(const oc_params *p, int32_t N, oc_state *out);

/* Fresh start of every replica from full velocities v (gid order arrays are
 * concatenated over replicas): v~ = v, sort, forces at step 0 (warm-up ramp
 * starts), vh = v - dt/2 f.  cel/engine/oercel.c:2196-2212. */
int oc_start(oc_engine *h, const oc_state *s);

/* Load / read replica r's state (gid order).  oc_set_state mirrors a restart
 * from a checkpoint (cel/engine/oercel.c:1942-2010): vh and f as stored. */
int oc_set_state(oc_engine *h, int32_t r, const oc_state *s, int64_t step);

/* oc_set_state with the slot order of a checkpoint: slot_gid[q] = gid of the
 * bead in slot q (a permutation of 0..N-1).  The reference keeps its
 * checkpoint in slot order and continues from it (read_checkpoint,
 * cel/engine/oercel.c:1942-2010), so the next stable sort, and with it every
 * force sum, is the one the uninterrupted run makes: continuation is bitwise. */
int oc_set_state_slots(oc_engine *h, int32_t r, const oc_state *s, int64_t step, const uint32_t *slot_gid);
int oc_get_state(oc_engine *h, int32_t r, oc_state *s);

/* n full steps of every replica (integrate, sort, chemistry, pair forces,
 * bonded forces): the loop body of cel/engine/oercel.c:2300-2362.
 * use_graph = 1 captures a step pair in a hipGraph and replays it. */
int oc_step(oc_engine *h, int64_t n, int32_t use_graph);

/* Single phases on the current state, for parity tests. */
int oc_sort(oc_engine *h);                       /* sort_cells, :828-933 */
int oc_forces(oc_engine *h, int64_t step, int32_t with_bonded); /* pair_forces (+ bonded_forces), :1198-1216, :1317-1343 */
int oc_chem(oc_engine *h, int64_t step);         /* precursor chemistry at this step (no counterpart) */

/* oc_forces with the caller's pair noise in place of the counter hash: the
 * reference's force test compares forces on identical positions, velocities
 * and noise (cel/engine/tests/forcetest.c:91-103).  theta[q] is the
 * unit-variance deviate of the pair (pair_lo[q], pair_hi[q]) of gids, lo < hi,
 * entries in strictly increasing (lo, hi) order; an interacting pair missing
 * from the table gets theta = 0 and the call returns OC_EDEVICE.  The
 * single-pass pair kernel; one replica. */
int oc_forces_theta(oc_engine *h, int64_t step, int32_t with_bonded, const uint32_t *pair_lo,
                    const uint32_t *pair_hi, const float *theta, int64_t npair);

/* oc_chem with the caller's uniforms in place of the counter hashes (no
 * counterpart in the reference): u_conv[gid] and u_break[gid] are the P
 * conversion and bond break draws of every bead, u_form[q] the symmetric
 * formation draw of the pair (form_lo[q], form_hi[q]) (lo < hi, strictly
 * increasing); a free-T pair within r_form missing from the table does not
 * bond and the call returns OC_EDEVICE.  Precursor chemistry only; one
 * replica. */
int oc_chem_uniforms(oc_engine *h, int64_t step, const double *u_conv, const double *u_break,
                     const uint32_t *form_lo, const uint32_t *form_hi, const double *u_form, int64_t nform);

/* Cell list of replica r: cstart[ncell+1] (slot indices relative to the
 * replica) and gid_of_slot[N]; ncell = floor(L) * floor(L) * nsub, column-major
 * (cel/engine/oercel.c:739-745, 779-786). */
int oc_get_cells(oc_engine *h, int32_t r, int32_t *cstart, uint32_t *gid_of_slot, int32_t *ncell);

/* Chemistry events of replica r not yet returned, sorted by (step, kind,
 * gid_a, gid_b).  n_out receives how many are pending; the first min(cap, n)
 * are copied to ev and removed (ev = NULL or cap = 0 only counts). */
int oc_get_events(oc_engine *h, int32_t r, oc_event *ev, int64_t cap, int64_t *n_out);

int oc_get_stats(oc_engine *h, int32_t r, oc_stats *out);

/* Running max bond length of replica r (the value oc_get_stats reports and
 * resets): *get receives it, then *set (>= 0) replaces it; either may be NULL.
 * A checkpoint carries it (the reference's write_ckpt / read_ckpt,
 * cel/engine/oercel.c:1891-2010, store s->maxbond), so the stats row after a
 * restart equals the uninterrupted run's. */
int oc_max_bond(oc_engine *h, int32_t r, double *get, const double *set);

/* Current step counter of replica r. */
int oc_get_step(oc_engine *h, int32_t r, int64_t *step);

/* Smooth radial distribution of the type pair (ta, tb): nbins bins on
 * [0, rmax), each pair spread over the bins by a triweight kernel
 * 35/(32 w) (1 - (u/w)^2)^3 of half-width w = width. */
typedef struct oc_rdf_spec {
    int32_t ta, tb, nbins, pad;
    double rmax, width;
} oc_rdf_spec;

/* Replace replica r's repulsion matrix (8x8 row-major, symmetric) and gamma;
 * sigma = sqrt(2 gamma kT) follows.  No counterpart in the reference (its
 * parameters are fixed for a run). */
int oc_set_pair_params(oc_engine *h, int32_t r, const double *a64, double gamma);

/* Forward half of the window gradient: recompute f at the current state, run
 * nsteps steps of every replica (checkpoint every ckpt_every steps) and return
 * the smooth RDF of spec averaged over steps 1..nsteps, gbar[r*nbins + b].
 * The engine ends at the window's last step.  Needs k_angle = 0, chemistry,
 * stirring and the kT cycle off, and the warm-up ramp over. */
int oc_window_forward(oc_engine *h, int32_t nsteps, int32_t ckpt_every, const oc_rdf_spec *spec, double *gbar);

/* Reverse half: given dL/dgbar (nrep * nbins) of any loss of the last forward
 * window, grad[r*7 + q] = dL/dtheta_q of replica r over q = (a_WW, a_WH,
 * a_WT, a_HH, a_HT, a_TT, gamma), the noise stream held fixed.  Recomputes the
 * window from its checkpoints; the engine stays at the window's last step. */
int oc_window_backward(oc_engine *h, const double *dL_dgbar, double *grad);

/* Put every replica back at the start of the last forward window (state, step
 * counters), so the next window runs from the same state: a parameter fit
 * whose windows all start there minimises one deterministic loss of theta
 * (the noise stream is fixed).  No counterpart in the reference. */
int oc_window_rewind(oc_engine *h);

/* Window gradient of every replica (no counterpart in the reference): with the
 * noise stream fixed and chemistry off, recompute f at the current state, run
 * nsteps steps (checkpoint every ckpt_every steps, recomputed in the reverse
 * sweep), average the smooth RDF of spec over steps 1..nsteps into gbar, and
 * return loss_r = sum_b (gbar_rb - target_rb)^2 and its gradient
 * grad[r*7 + q] over q = (a_WW, a_WH, a_WT, a_HH, a_HT, a_TT, gamma).
 * target_stride = nbins for per-replica targets, 0 for one shared target.
 * The engine ends at the window's last step.  Needs k_angle = 0. */
int oc_window_grad(oc_engine *h, int32_t nsteps, int32_t ckpt_every, const oc_rdf_spec *spec, const double *target,
                   int32_t target_stride, double *loss, double *grad, double *gbar);

/* Launch only the force kernel (pair forces with the bonded terms, as the
 * step runs it) reps times on the handle's stream, for the current state and
 * step, without synchronising: lets a caller time that kernel alone with
 * events on oc_stream().  The reps launches run as one captured graph, built
 * at the first call with that count (and after any change of the engine's
 * buffers), so a timing caller makes one untimed call with the same count
 * first; between the launches are only the graph's own gaps. */
int oc_launch_pair(oc_engine *h, int32_t reps);

/* sizeof the ABI structs, for binding checks: out[0..3] = oc_params,
 * oc_state, oc_event, oc_stats (n = room in out); returns how many it wrote. */
int oc_abi_sizes(int32_t *out, int32_t n);

/* HIP stream of the handle (hipStream_t as void*), for callers that overlap
 * their own work or time kernels with events on it. */
void *oc_stream(oc_engine *h);

/* Kinetic temperature and total momentum of n >= 2 beads whose half-step
 * velocities and forces the caller holds in device memory, in any order:
 * kinetic_T (cel/engine/oercel.c:2058-2071) with v = vh + dt/2 f, written to
 * the device array out[3] = (sum |v|^2 / (2n - 2), sum v_x, sum v_y).  FP64
 * over a fixed partition: the same inputs give the same bits.  work: device
 * memory of at least oc_kinetic_dev_work(n) bytes for the block partials.
 * Asynchronous on stream (a hipStream_t; NULL = the null stream); needs no
 * engine handle.  Errors: OC_EINVAL (null array, n < 2, small workspace). */
size_t oc_kinetic_dev_work(int64_t n);
int oc_kinetic_dev(const double *vx, const double *vy, const double *fx, const double *fy, int64_t n, double dt,
                   double *out, void *work, size_t work_bytes, void *stream);

/* oc_kinetic_dev without a workspace and without any allocation (no
 * hipMalloc, hipFree or allocating library call), for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops.  No
 * counterpart in the reference.  The block partials live in the library's
 * fixed device scratch of 384 KiB, one copy per GPU held until the process
 * ends, so n <= 67,108,864 (OC_EINVAL beyond); calls that overlap on one GPU
 * race on that scratch.  Same kernels and bits as oc_kinetic_dev. */
int oc_kinetic_dev_noalloc(const double *vx, const double *vy, const double *fx, const double *fy, int64_t n,
                           double dt, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
