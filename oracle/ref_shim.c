/*
 * ref_shim.c -- TEST INFRASTRUCTURE ONLY.  Compiles the reference engine
 * (cel/engine/oercel.c, included from where it lies; path given by the
 * Makefile as REF_ENGINE_SRC) into oracle/_ref/libocref.so and exposes a few of
 * its static step functions with plain C signatures, the way the reference's
 * own tests reach them (cel/engine/tests/forcetest.c:3-5).  Used only to pin
 * the oracle restatement (oracle.cpp) and to write tests/golden fixtures.
 *
 * Parameters not set here keep the reference's defaults (PDEFS,
 * cel/engine/oercel.c:237-299) and pass through its own finalize_params.
 */
#define main oercel_ref_main
#include REF_ENGINE_SRC
#undef main

#include "../include/oercel_mi.h"

#define REF_API __attribute__((visibility("default")))

static Params g_refP;

static void ref_params(const oc_params *p, int threads) {
    memset(&g_refP, 0, sizeof g_refP);
    for (int k = 0; k < NPDEFS; k++) set_param(&g_refP, PDEFS[k].key, PDEFS[k].def);
    default_amatrix(&g_refP);
    Params *P = &g_refP;
    P->L = p->L; P->rho = p->rho; P->dt = p->dt; P->lambda = p->lambda; P->gamma = p->gamma; P->kT = p->kT;
    P->k_bond = p->k_bond; P->r0 = p->r0; P->k_angle = p->k_angle;
    for (int i = 0; i < NTYPES; i++)
        for (int j = 0; j < NTYPES; j++) P->a[i][j] = p->a[i][j];
    P->stir_A = p->stir_A; P->stir_n = p->stir_n; P->kT_amp = p->kT_amp; P->kT_period = p->kT_period;
    P->nsub = p->nsub; P->lip_nH = p->lip_nH; P->lip_nT = p->lip_nT;
    P->warmup = p->warmup; P->seed = p->seed;
    P->f_lip = p->f_lip; P->f_Ka = p->f_Ka; P->f_Kb = p->f_Kb; P->c_N = p->c_N;
    P->threads = threads; P->chem = 0;
    {
        static const char *const arch[4] = {"linear", "pi", "y", "bola"};
        snprintf(P->lip_arch, sizeof P->lip_arch, "%s", arch[p->lip_arch >= 0 && p->lip_arch < 4 ? p->lip_arch : 0]);
        P->lip_tie = p->lip_tie; P->r0_head = p->r0_head; P->k_head = p->k_head; P->cos_head = p->cos_head;
    }
    finalize_params(P);
    g_warm_end = p->warm_end;
    omp_set_num_threads(threads);
}

static void ref_alloc(Sys *s, int N) {
    sys_alloc(s, N, g_refP.L, g_mol.mlen, g_mol.nH, g_refP.threads, g_refP.nsub);
}

REF_API uint64_t ref_mix13(uint64_t z) { return mix13(z); }
REF_API uint64_t ref_step_seed(uint64_t seed, int64_t step) { return step_seed(seed, step); }
REF_API double ref_pair_theta(uint64_t s, uint32_t gi, uint32_t gj) { return pair_theta(s, gi, gj); }
REF_API uint64_t ref_hash_key(uint64_t seed, int64_t step, uint64_t salt) { return hash_key(seed, step, salt); }
REF_API double ref_hash_u01(uint64_t key, uint32_t g) { return hash_u01(key, g); }

/* init_soup + set_velocities after rng_seed(seed), cel/engine/oercel.c:2190-2196; gid order */
REF_API int ref_soup(const oc_params *p, int N, double *x, double *y, double *vx, double *vy, uint8_t *type,
                     int8_t *mpos, int32_t *nxt) {
    ref_params(p, 1);
    Sys S, *s = &S;
    ref_alloc(s, N);
    s->seed = g_refP.seed;
    rng_seed(&s->rng, g_refP.seed);
    init_soup(s, &g_refP);
    set_velocities(s, &g_refP);
    for (int i = 0; i < N; i++) {
        x[i] = s->x[i]; y[i] = s->y[i]; vx[i] = s->vx[i]; vy[i] = s->vy[i];
        type[i] = s->type[i]; mpos[i] = s->mpos[i]; nxt[i] = s->nxt[i];
    }
    return 0;
}

/* sort_cells on the given positions (input order = gid): src[new slot] = input index */
REF_API int ref_sort(const oc_params *p, int N, int threads, const double *x, const double *y, int32_t *src,
                     int32_t *cstart) {
    ref_params(p, threads);
    Sys S, *s = &S;
    ref_alloc(s, N);
    for (int i = 0; i < N; i++) {
        s->x[i] = x[i]; s->y[i] = y[i]; s->vx[i] = s->vy[i] = s->vtx[i] = s->vty[i] = 0.0;
        s->type[i] = TW; s->gid[i] = (uint32_t)i; s->mpos[i] = -1; s->nxt[i] = -1;
    }
    sort_cells(s);
    for (int d = 0; d < N; d++) src[d] = (int32_t)s->gid[d];
    memcpy(cstart, s->cstart, sizeof(int32_t) * (s->ncell + 1));
    return s->ncell;
}

/* pair_forces (+ bonded_forces) of the reference on a gid-order state;
 * forces returned in gid order; nxt given as gid (slot = gid before the sort) */
REF_API int ref_forces(const oc_params *p, int N, int threads, const double *x, const double *y, const double *vtx,
                       const double *vty, const uint8_t *type, const int8_t *mpos, const int32_t *nxt, int64_t step,
                       double t, int with_pair, int with_bonded, double *fx, double *fy, double *maxbond) {
    ref_params(p, threads);
    Sys S, *s = &S;
    ref_alloc(s, N);
    for (int i = 0; i < N; i++) {
        s->x[i] = x[i]; s->y[i] = y[i]; s->vx[i] = s->vy[i] = 0.0; s->vtx[i] = vtx[i]; s->vty[i] = vty[i];
        s->type[i] = type[i]; s->gid[i] = (uint32_t)i; s->mpos[i] = mpos[i]; s->nxt[i] = nxt[i];
    }
    s->seed = g_refP.seed;
    sort_cells(s);
    if (with_pair) pair_forces(s, &g_refP, step, t);
    else
        for (int i = 0; i < N; i++) s->fx[i] = s->fy[i] = 0.0;
    s->maxbond = 0;
    if (with_bonded) bonded_forces(s, &g_refP);
    for (int d = 0; d < N; d++) { fx[s->gid[d]] = s->fx[d]; fy[s->gid[d]] = s->fy[d]; }
    if (maxbond) *maxbond = s->maxbond;
    return 0;
}

/* kinetic_T, cel/engine/oercel.c:2058-2071, on arrays in any order */
REF_API double ref_kinetic_T(int N, double dt, double *vx, double *vy, double *fx, double *fy, double *px, double *py) {
    Sys S;
    memset(&S, 0, sizeof S);
    S.N = N; S.vx = vx; S.vy = vy; S.fx = fx; S.fy = fy;
    return kinetic_T(&S, dt, px, py);
}
