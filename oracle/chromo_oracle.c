/*
 * chromo_oracle.c -- TEST INFRASTRUCTURE ONLY (see chromo_oracle.h).
 *
 * Faithful CPU restatement of chromo's MC inner loop.  Compiled with
 * -O2 -ffp-contract=off so that, like the reference's x86-64 build, no fused
 * multiply-adds appear, and every floating-point expression below keeps the
 * reference's evaluation order.  Each function names the reference lines it
 * restates (paths relative to the reference repository root).
 *
 * Two builds of the same arithmetic are selectable per simulation:
 *   faithful = 1 keeps the reference's per-move O(n_bins) reset of
 *                affected_bins_last_move (chromo/fields.pyx:1223-1224) and the
 *                O(n_bins) scan of update_affected_densities (fields.pyx:1971-1975);
 *   faithful = 0 ("lean") touches only the voxels of the move.  Results are identical.
 */
#define _GNU_SOURCE
#include "chromo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define RAND_MAX_I 2147483647
#define E_HUGE_FIELD 1E99   /* chromo/fields.pyx:32 */
#define E_HUGE_POLY 1E25    /* chromo/polymers.pyx:35 */

/* ------------------------------------------------------------------------- */
/* Random number streams                                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    int mode;
    int32_t g[31];
    int gf, gr;                 /* glibc TYPE_3 front / rear pointers */
    uint32_t mt[624];
    int mti;
    uint32_t key[2];
    uint64_t ctr;
    uint32_t pbuf[4];
    int pbi;
} rng_t;

/* glibc random_r TYPE_3: the generator behind libc rand() that the reference
 * calls unseeded (== srand(1)) at chromo/mc/mc_sim.pyx:171,
 * move_funcs.pyx:81,118,326,445,504,577,763,765, linalg.pyx:34-35,55-56,
 * bead_selection.pyx:58-65,144.  Additive feedback x[i] = x[i-3] + x[i-31]. */
static int32_t g3_next(rng_t *g)
{
    uint32_t v = (uint32_t)g->g[g->gf] + (uint32_t)g->g[g->gr];
    g->g[g->gf] = (int32_t)v;
    int32_t res = (int32_t)(v >> 1);
    if (++g->gf >= 31) {
        g->gf = 0;
        ++g->gr;
    } else if (++g->gr >= 31) {
        g->gr = 0;
    }
    return res;
}

static void g3_seed(rng_t *g, uint32_t seed)
{
    if (seed == 0) seed = 1;
    int32_t word = (int32_t)seed;
    g->g[0] = word;
    for (int i = 1; i < 31; ++i) {
        int32_t hi = word / 127773;
        int32_t lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        g->g[i] = word;
    }
    g->gf = 3;
    g->gr = 0;
    for (int i = 0; i < 310; ++i) (void)g3_next(g);
}

/* MT19937 as numpy's legacy RandomState (seeded through np.random.seed(int),
 * chromo/mc/mc_sim.pyx:81); only new binding states draw from it
 * (chromo/mc/move_funcs.pyx:819). */
static void mt_seed(rng_t *g, uint32_t s)
{
    g->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->mti = 624;
}

static uint32_t mt_next(rng_t *g)
{
    if (g->mti >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based. */
static void philox_round(uint32_t c[4], const uint32_t k[2])
{
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof(c));
}

static uint32_t philox_next(rng_t *g)
{
    if (g->pbi >= 4) {
        uint32_t c[4] = {(uint32_t)g->ctr, (uint32_t)(g->ctr >> 32), 0u, 0u};
        or_philox4x32_10(c, g->key, g->pbuf);
        g->ctr++;
        g->pbi = 0;
    }
    return g->pbuf[g->pbi++];
}

/* rand() replacement: 31-bit draw in [0, RAND_MAX]. */
static int32_t rng_rand(rng_t *g)
{
    if (g->mode == OR_RNG_GLIBC) return g3_next(g);
    return (int32_t)(philox_next(g) >> 1);
}

/* numpy legacy RandomState.randint(0, high) for small high: masked rejection
 * sampling over 32-bit draws. */
static int64_t rng_randint(rng_t *g, int64_t high)
{
    uint32_t rng = (uint32_t)(high - 1);
    if (rng == 0) return 0;
    uint32_t mask = rng;
    mask |= mask >> 1;
    mask |= mask >> 2;
    mask |= mask >> 4;
    mask |= mask >> 8;
    mask |= mask >> 16;
    for (;;) {
        uint32_t v = (g->mode == OR_RNG_GLIBC ? mt_next(g) : philox_next(g)) & mask;
        if (v <= rng) return (int64_t)v;
    }
}

void or_glibc_stream(uint32_t seed, int32_t *out, int64_t n)
{
    rng_t g;
    memset(&g, 0, sizeof(g));
    g3_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = g3_next(&g);
}

void or_mt_randint_stream(uint32_t seed, int64_t high, int64_t *out, int64_t n)
{
    rng_t g;
    memset(&g, 0, sizeof(g));
    g.mode = OR_RNG_GLIBC;
    mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = rng_randint(&g, high);
}

/* ------------------------------------------------------------------------- */
/* CPython set order (bins_found = set() at chromo/fields.pyx:1427-1522)      */
/* ------------------------------------------------------------------------- */

/* Emulates CPython 3 setobject.c for non-negative int keys (hash(i) == i):
 * open addressing with LINEAR_PROBES = 9, PERTURB_SHIFT = 5, growth to
 * used*4 when fill*5 >= mask*3.  Iteration order is table order. */
typedef struct {
    int64_t *key;
    uint8_t *used;
    uint64_t mask;
    int64_t fill, nused;
    int64_t cap;
} pyset;

static void pyset_init(pyset *s)
{
    s->cap = 8;
    s->mask = 7;
    s->key = (int64_t *)calloc(8, sizeof(int64_t));
    s->used = (uint8_t *)calloc(8, 1);
    s->fill = s->nused = 0;
}

static void pyset_clear(pyset *s)
{
    if (s->cap != 8) {
        free(s->key);
        free(s->used);
        s->key = (int64_t *)calloc(8, sizeof(int64_t));
        s->used = (uint8_t *)calloc(8, 1);
        s->cap = 8;
    } else {
        memset(s->used, 0, 8);
    }
    s->mask = 7;
    s->fill = s->nused = 0;
}

static void pyset_free(pyset *s)
{
    free(s->key);
    free(s->used);
}

static void pyset_insert_clean(int64_t *key, uint8_t *used, uint64_t mask, int64_t k)
{
    uint64_t perturb = (uint64_t)k;
    uint64_t i = (uint64_t)k & mask;
    uint64_t e;
    for (;;) {
        e = i;
        if (!used[e]) goto found;
        if (i + 9 <= mask) {
            for (int j = 0; j < 9; ++j) {
                ++e;
                if (!used[e]) goto found;
            }
        }
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & mask;
    }
found:
    used[e] = 1;
    key[e] = k;
}

static void pyset_resize(pyset *s, int64_t minused)
{
    uint64_t newsize = 8;
    while (newsize <= (uint64_t)minused) newsize <<= 1;
    int64_t *nk = (int64_t *)calloc(newsize, sizeof(int64_t));
    uint8_t *nu = (uint8_t *)calloc(newsize, 1);
    for (uint64_t i = 0; i <= s->mask; ++i)
        if (s->used[i]) pyset_insert_clean(nk, nu, newsize - 1, s->key[i]);
    free(s->key);
    free(s->used);
    s->key = nk;
    s->used = nu;
    s->mask = newsize - 1;
    s->cap = (int64_t)newsize;
    s->fill = s->nused;
}

/* returns 1 if newly added, 0 if present */
static int pyset_add(pyset *s, int64_t k)
{
    uint64_t perturb = (uint64_t)k;
    uint64_t i = (uint64_t)k & s->mask;
    uint64_t e;
    for (;;) {
        int probes = (i + 9 <= s->mask) ? 9 : 0;
        e = i;
        do {
            if (!s->used[e]) goto unused;
            if (s->key[e] == k) return 0;
            ++e;
        } while (probes--);
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & s->mask;
    }
unused:
    s->used[e] = 1;
    s->key[e] = k;
    s->fill++;
    s->nused++;
    if ((uint64_t)s->fill * 5 >= s->mask * 3)
        pyset_resize(s, s->nused > 50000 ? s->nused * 2 : s->nused * 4);
    return 1;
}

static int pyset_contains(const pyset *s, int64_t k)
{
    uint64_t perturb = (uint64_t)k;
    uint64_t i = (uint64_t)k & s->mask;
    uint64_t e;
    for (;;) {
        int probes = (i + 9 <= s->mask) ? 9 : 0;
        e = i;
        do {
            if (!s->used[e]) return 0;
            if (s->key[e] == k) return 1;
            ++e;
        } while (probes--);
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & s->mask;
    }
}

static int64_t pyset_list(const pyset *s, int64_t *out)
{
    int64_t n = 0;
    for (uint64_t i = 0; i <= s->mask; ++i)
        if (s->used[i]) out[n++] = s->key[i];
    return n;
}

int64_t or_pyset_order(const int64_t *keys, int64_t n, int64_t *out)
{
    pyset s;
    pyset_init(&s);
    for (int64_t i = 0; i < n; ++i) pyset_add(&s, keys[i]);
    int64_t m = pyset_list(&s, out);
    pyset_free(&s);
    return m;
}

/* ------------------------------------------------------------------------- */
/* Simulation state                                                           */
/* ------------------------------------------------------------------------- */

struct or_sim {
    int64_t N, B, Bp1;
    double *r, *t3, *t2, *r_trial, *t3_trial, *t2_trial;
    int64_t *states, *states_trial, *marks;
    double *eps_bend, *eps_par, *eps_perp, *gamma, *eta;
    int twist;
    double *eps_twist, *bead_length;
    int64_t max_binders;
    int64_t sites[OR_MAXB];
    double mu[OR_MAXB];
    int64_t n_mark_levels;
    double F[OR_MAXB][OR_MAXLEV][OR_MAXS + 1];
    double mu_adjust_factor;
    /* field */
    int field_on;
    int64_t nx, ny, nz, n_bins;
    double width[3], dxyz[3], half_width[3], half_step[3], vol_bin;
    int64_t n_xyz_m1[3];
    float vf_limit;
    double chi, bead_vol;
    double pref[OR_MAXB], intra[OR_MAXB], cross[OR_MAXB * OR_MAXB];
    int confine_type;
    double confine_length;
    int64_t *wrap;                 /* inds_xyz_to_super, (nx+1)(ny+1)(nz+1) */
    double *density, *density_trial;
    double *access_vols;           /* access_vols[bin] (fields.pyx:714-770) */
    int fast_field;                /* init_fast_field (fields.pyx:577-671) */
    int64_t n_points, n_sub[3];
    double sub_w[3];
    int64_t *affected;
    int faithful;
    /* move scratch */
    int64_t *inds;
    int64_t inds_cap;
    int64_t *bins;
    int64_t bins_cap;
    pyset found;
    double direction[3], point[3], T[16];
    rng_t rng;
};

static inline int64_t wrap_index(const or_sim *s, int64_t i, int64_t j, int64_t k)
{
    return s->wrap[(i * (s->ny + 1) + j) * (s->nz + 1) + k];
}

/* Cython float modulo with cdivision=False: fmod plus a sign fix
 * (__Pyx_mod_double); used at chromo/fields.pyx:1437-1445. */
static inline double pymod(double a, double b)
{
    double r = fmod(a, b);
    r += (double)((r != 0) & ((r < 0) ^ (b < 0))) * b;
    return r;
}

or_sim *or_new(const or_spec *sp, int32_t rng_mode, uint32_t seed, uint32_t stream,
               int32_t faithful)
{
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    int64_t N = sp->N, B = sp->B;
    s->N = N;
    s->B = B;
    s->Bp1 = B + 1;
    size_t v3 = (size_t)N * 3 * sizeof(double);
    s->r = (double *)malloc(v3);
    s->t3 = (double *)malloc(v3);
    s->t2 = (double *)malloc(v3);
    s->r_trial = (double *)malloc(v3);
    s->t3_trial = (double *)malloc(v3);
    s->t2_trial = (double *)malloc(v3);
    memcpy(s->r, sp->r, v3);
    memcpy(s->t3, sp->t3, v3);
    memcpy(s->t2, sp->t2, v3);
    memcpy(s->r_trial, sp->r, v3);
    memcpy(s->t3_trial, sp->t3, v3);
    memcpy(s->t2_trial, sp->t2, v3);
    size_t sb = (size_t)N * (size_t)(B > 0 ? B : 1) * sizeof(int64_t);
    s->states = (int64_t *)calloc(1, sb);
    s->states_trial = (int64_t *)calloc(1, sb);
    s->marks = (int64_t *)calloc(1, sb);
    if (B > 0) {
        memcpy(s->states, sp->states, (size_t)N * B * sizeof(int64_t));
        memcpy(s->states_trial, sp->states, (size_t)N * B * sizeof(int64_t));
        memcpy(s->marks, sp->marks, (size_t)N * B * sizeof(int64_t));
    }
    size_t nb = (size_t)(N > 1 ? N - 1 : 1) * sizeof(double);
    s->eps_bend = (double *)calloc(1, nb);
    s->eps_par = (double *)calloc(1, nb);
    s->eps_perp = (double *)calloc(1, nb);
    s->gamma = (double *)calloc(1, nb);
    s->eta = (double *)calloc(1, nb);
    if (N > 1) {
        memcpy(s->eps_bend, sp->eps_bend, (size_t)(N - 1) * sizeof(double));
        memcpy(s->eps_par, sp->eps_par, (size_t)(N - 1) * sizeof(double));
        memcpy(s->eps_perp, sp->eps_perp, (size_t)(N - 1) * sizeof(double));
        memcpy(s->gamma, sp->gamma, (size_t)(N - 1) * sizeof(double));
        memcpy(s->eta, sp->eta, (size_t)(N - 1) * sizeof(double));
    }
    s->twist = sp->twist != 0;
    s->eps_twist = (double *)calloc(1, nb);
    s->bead_length = (double *)calloc(1, nb);
    if (s->twist && N > 1) {
        memcpy(s->eps_twist, sp->eps_twist, (size_t)(N - 1) * sizeof(double));
        memcpy(s->bead_length, sp->bead_length, (size_t)(N - 1) * sizeof(double));
    }
    s->max_binders = sp->max_binders;
    memcpy(s->sites, sp->sites, sizeof(s->sites));
    memcpy(s->mu, sp->mu, sizeof(s->mu));
    s->n_mark_levels = sp->n_mark_levels;
    memcpy(s->F, sp->F, sizeof(s->F));
    s->mu_adjust_factor = 1.0;

    s->field_on = sp->field_on;
    s->nx = sp->nx > 0 ? sp->nx : 1;
    s->ny = sp->ny > 0 ? sp->ny : 1;
    s->nz = sp->nz > 0 ? sp->nz : 1;
    /* UniformDensityField.init_grid, chromo/fields.pyx:546-565 */
    s->dxyz[0] = sp->x_width / (double)s->nx;
    s->dxyz[1] = sp->y_width / (double)s->ny;
    s->dxyz[2] = sp->z_width / (double)s->nz;
    s->n_bins = s->nx * s->ny * s->nz;
    s->vol_bin = sp->x_width * sp->y_width * sp->z_width / (double)s->n_bins;
    s->fast_field = sp->fast_field;
    s->n_points = sp->n_points + (sp->n_points % 2);
    {
        const int64_t nn[3] = {s->nx, s->ny, s->nz};
        for (int j = 0; j < 3; ++j) {
            s->sub_w[j] = s->dxyz[j] / (double)s->n_points;
            s->n_sub[j] = s->n_points * nn[j];
        }
    }
    s->access_vols = (double *)malloc((size_t)s->n_bins * sizeof(double));
    for (int64_t i = 0; i < s->n_bins; ++i)
        s->access_vols[i] = sp->access_vols ? sp->access_vols[i] : s->vol_bin;
    s->width[0] = sp->x_width;
    s->width[1] = sp->y_width;
    s->width[2] = sp->z_width;
    for (int j = 0; j < 3; ++j) {
        s->half_width[j] = 0.5 * s->width[j];
        s->half_step[j] = 0.5 * s->dxyz[j];
    }
    s->n_xyz_m1[0] = s->nx - 1;
    s->n_xyz_m1[1] = s->ny - 1;
    s->n_xyz_m1[2] = s->nz - 1;
    s->vf_limit = sp->vf_limit;
    s->chi = sp->chi;
    s->bead_vol = sp->bead_vol;
    memcpy(s->pref, sp->pref, sizeof(s->pref));
    memcpy(s->intra, sp->intra, sizeof(s->intra));
    memcpy(s->cross, sp->cross, sizeof(s->cross));
    s->confine_type = sp->confine_type;
    s->confine_length = sp->confine_length;
    /* precompute_ind_xyz_to_super, chromo/fields.pyx:673-685 */
    s->wrap = (int64_t *)malloc((size_t)(s->nx + 1) * (s->ny + 1) * (s->nz + 1) * sizeof(int64_t));
    for (int64_t i = 0; i <= s->nx; ++i)
        for (int64_t j = 0; j <= s->ny; ++j)
            for (int64_t k = 0; k <= s->nz; ++k)
                s->wrap[(i * (s->ny + 1) + j) * (s->nz + 1) + k] =
                    (i % s->nx) + (j % s->ny) * s->nx + (k % s->nz) * s->nx * s->ny;
    s->density = (double *)calloc((size_t)s->n_bins * s->Bp1, sizeof(double));
    s->density_trial = (double *)calloc((size_t)s->n_bins * s->Bp1, sizeof(double));
    s->affected = (int64_t *)calloc((size_t)s->n_bins, sizeof(int64_t));
    s->faithful = faithful;
    s->inds_cap = N + 1;
    s->inds = (int64_t *)malloc((size_t)s->inds_cap * sizeof(int64_t));
    s->bins_cap = 16 * (N + 1) + 16;
    s->bins = (int64_t *)malloc((size_t)s->bins_cap * sizeof(int64_t));
    pyset_init(&s->found);

    s->rng.mode = rng_mode;
    if (rng_mode == OR_RNG_GLIBC) {
        g3_seed(&s->rng, seed);
        mt_seed(&s->rng, 0u);
    } else {
        s->rng.key[0] = seed;
        s->rng.key[1] = stream;
        s->rng.ctr = 0;
        s->rng.pbi = 4;
    }
    if (s->field_on) or_density_init(s);
    return s;
}

void or_free(or_sim *s)
{
    if (!s) return;
    free(s->r);
    free(s->t3);
    free(s->t2);
    free(s->r_trial);
    free(s->t3_trial);
    free(s->t2_trial);
    free(s->states);
    free(s->states_trial);
    free(s->marks);
    free(s->eps_bend);
    free(s->eps_par);
    free(s->eps_perp);
    free(s->gamma);
    free(s->eta);
    free(s->eps_twist);
    free(s->bead_length);
    free(s->wrap);
    free(s->density);
    free(s->access_vols);
    free(s->density_trial);
    free(s->affected);
    free(s->inds);
    free(s->bins);
    pyset_free(&s->found);
    free(s);
}

void or_np_seed(or_sim *s, uint32_t seed)
{
    if (s->rng.mode == OR_RNG_GLIBC) mt_seed(&s->rng, seed);
}

int64_t or_n_bins(const or_sim *s) { return s->n_bins; }

void or_get_state(const or_sim *s, double *r, double *t3, double *t2, int64_t *states)
{
    size_t v3 = (size_t)s->N * 3 * sizeof(double);
    if (r) memcpy(r, s->r, v3);
    if (t3) memcpy(t3, s->t3, v3);
    if (t2) memcpy(t2, s->t2, v3);
    if (states && s->B > 0) memcpy(states, s->states, (size_t)s->N * s->B * sizeof(int64_t));
}

void or_set_positions(or_sim *s, const double *r, const double *t3, const double *t2)
{
    size_t v3 = (size_t)s->N * 3 * sizeof(double);
    memcpy(s->r, r, v3);
    memcpy(s->r_trial, r, v3);
    memcpy(s->t3, t3, v3);
    memcpy(s->t3_trial, t3, v3);
    memcpy(s->t2, t2, v3);
    memcpy(s->t2_trial, t2, v3);
}

void or_get_density(const or_sim *s, double *density)
{
    memcpy(density, s->density, (size_t)s->n_bins * s->Bp1 * sizeof(double));
}

/* ------------------------------------------------------------------------- */
/* Geometry: chromo/util/linalg.pyx and transform_r_t3_t2                     */
/* ------------------------------------------------------------------------- */

/* uniform_sample_unit_sphere[_inplace], chromo/util/linalg.pyx:23-59 */
static void sample_unit_sphere(rng_t *g, double v[3])
{
    double phi = (double)rng_rand(g) / RAND_MAX_I * (2.0 * M_PI);
    double theta = acos(((double)rng_rand(g) / RAND_MAX_I) * 2 - 1);
    v[0] = cos(phi) * sin(theta);
    v[1] = sin(phi) * sin(theta);
    v[2] = cos(theta);
}

/* arbitrary_axis_rotation, chromo/util/linalg.pyx:62-139; axis[i]**2 is pow(x, 2.0)
 * in the Cython build, which glibc evaluates as x*x. */
void or_axis_rotation(const double a[3], const double p[3], double ang, double T[16])
{
    double c = cos(ang), s = sin(ang);
    for (int i = 0; i < 16; ++i) T[i] = 0.0;
    T[15] = 1.0;
    T[0] = a[0] * a[0] + (a[1] * a[1] + a[2] * a[2]) * c;
    T[1] = a[0] * a[1] * (1 - c) - a[2] * s;
    T[2] = a[0] * a[2] * (1 - c) + a[1] * s;
    T[4] = a[0] * a[1] * (1 - c) + a[2] * s;
    T[5] = a[1] * a[1] + (a[0] * a[0] + a[2] * a[2]) * c;
    T[6] = a[1] * a[2] * (1 - c) - a[0] * s;
    T[8] = a[0] * a[2] * (1 - c) - a[1] * s;
    T[9] = a[1] * a[2] * (1 - c) + a[0] * s;
    T[10] = a[2] * a[2] + (a[0] * a[0] + a[1] * a[1]) * c;
    double rv0 = (p[1] * a[2] - p[2] * a[1]) * s;
    double rv1 = (p[2] * a[0] - p[0] * a[2]) * s;
    double rv2 = (p[0] * a[1] - p[1] * a[0]) * s;
    rv0 += (p[0] * (1 - a[0] * a[0]) - a[0] * (p[1] * a[1] + p[2] * a[2])) * (1 - c);
    rv1 += (p[1] * (1 - a[1] * a[1]) - a[1] * (p[0] * a[0] + p[2] * a[2])) * (1 - c);
    rv2 += (p[2] * (1 - a[2] * a[2]) - a[2] * (p[0] * a[0] + p[1] * a[1])) * (1 - c);
    T[3] = rv0;
    T[7] = rv1;
    T[11] = rv2;
}

/* generate_translation_mat, chromo/util/linalg.pyx:172-199 */
void or_translation(const double d[3], double T[16])
{
    for (int i = 0; i < 16; ++i) T[i] = 0.0;
    T[0] = T[5] = T[10] = T[15] = 1.0;
    T[3] = d[0];
    T[7] = d[1];
    T[11] = d[2];
}

/* transform_r_t3_t2, chromo/mc/move_funcs.pyx:121-154: left fold from 0 over k. */
static inline void transform_bead(const double T[16], const double *r, const double *t3,
                                  const double *t2, double *ro, double *t3o, double *t2o)
{
    for (int j = 0; j < 3; ++j) {
        double er = 0, e3 = 0, e2 = 0;
        for (int k = 0; k < 3; ++k) {
            er += T[j * 4 + k] * r[k];
            e3 += T[j * 4 + k] * t3[k];
            e2 += T[j * 4 + k] * t2[k];
        }
        ro[j] = er + T[j * 4 + 3];
        t3o[j] = e3;
        t2o[j] = e2;
    }
}

void or_transform(const double T[16], const double *r, const double *t3, const double *t2,
                  int64_t n, double *ro, double *t3o, double *t2o)
{
    for (int64_t i = 0; i < n; ++i)
        transform_bead(T, r + 3 * i, t3 + 3 * i, t2 + 3 * i, ro + 3 * i, t3o + 3 * i, t2o + 3 * i);
}

/* ------------------------------------------------------------------------- */
/* Bead selection: chromo/util/bead_selection.pyx                             */
/* ------------------------------------------------------------------------- */

/* capped_exponential, bead_selection.pyx:19-66 */
static inline int64_t cexp_value(int32_t x, int64_t window)
{
    return (int64_t)(-log10(((double)x / RAND_MAX_I) + 0.00001) * window * 0.45 + 1.0001);
}

static int64_t capped_exponential(rng_t *g, int64_t window, int64_t cap)
{
    int64_t r = cexp_value(rng_rand(g), window);
    while (r > cap) r = cexp_value(rng_rand(g), window);
    return r;
}

/* from_left, bead_selection.pyx:69-90 */
static int64_t from_left(rng_t *g, int64_t window, int64_t nb, int *err)
{
    if (window > nb) {
        *err = 1;
        return 0;
    }
    return capped_exponential(g, window, window);
}

/* from_right, bead_selection.pyx:93-112 */
static int64_t from_right(rng_t *g, int64_t window, int64_t nb, int *err)
{
    int64_t d = from_left(g, window, nb, err);
    return nb - d;
}

static inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }

/* from_point, bead_selection.pyx:115-154 */
static int64_t from_point(rng_t *g, int64_t window, int64_t nb, int64_t ind0, int *err)
{
    if (window > nb) {
        *err = 1;
        return 0;
    }
    if (window < 1) return ind0;
    int64_t side = rng_rand(g) % 2;
    if (side == 0) {
        int64_t ws = imax(imin(window, ind0), 1);
        int64_t ub = imax(ind0, 1);
        return from_right(g, ws, ub, err);
    }
    int64_t ws = imax(imin(window, nb - ind0), 1);
    int64_t ub = imax(nb - ind0, 1);
    return from_left(g, ws, ub, err) + ind0;
}

/* check_bead_bounds, bead_selection.pyx:157-192 */
void or_check_bead_bounds(int64_t b0, int64_t b1, int64_t nb, int64_t *ind0, int64_t *indf)
{
    b0 = imin(b0, nb);
    b0 = imax(b0, 0);
    if (b1 > nb) {
        *ind0 = b0;
        *indf = nb;
    } else if (b1 < 0) {
        *ind0 = 0;
        *indf = b0 + 1;
    } else if (b0 == b1) {
        *ind0 = b0;
        *indf = b0 + 1;
    } else {
        *ind0 = imin(b0, b1);
        *indf = imax(b0, b1);
    }
}

int64_t or_select_stream(uint32_t seed, int32_t which, int64_t window, int64_t nb, int64_t ind0,
                         int64_t *out, int64_t n)
{
    rng_t g;
    memset(&g, 0, sizeof(g));
    g3_seed(&g, seed);
    int err = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (which == 0) out[i] = from_left(&g, window, nb, &err);
        else if (which == 1) out[i] = from_right(&g, window, nb, &err);
        else if (which == 2) out[i] = from_point(&g, window, nb, ind0, &err);
        else out[i] = capped_exponential(&g, window, nb);
        if (err) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Moves: chromo/mc/move_funcs.pyx                                            */
/* ------------------------------------------------------------------------- */

/* pick_random_amp, move_funcs.pyx:104-118 */
static inline double pick_random_amp(rng_t *g, double amp_move)
{
    return amp_move * (((double)rng_rand(g) / RAND_MAX_I) - 0.5);
}

static void apply_transform(or_sim *s, const int64_t *inds, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = inds[i];
        transform_bead(s->T, s->r + 3 * b, s->t3 + 3 * b, s->t2 + 3 * b, s->r_trial + 3 * b,
                       s->t3_trial + 3 * b, s->t2_trial + 3 * b);
    }
}

/* crank_shaft + get_crank_shaft_axis + get_crank_shaft_fulcrum,
 * move_funcs.pyx:39-101, 157-234, 237-280 */
static int64_t move_crank(or_sim *s, double amp_move, int64_t amp_bead, int *err)
{
    rng_t *g = &s->rng;
    int64_t N = s->N;
    double ang = pick_random_amp(g, amp_move);
    int64_t b0 = (int64_t)(int32_t)(((double)rng_rand(g) / RAND_MAX_I) * N);
    int64_t b1 = from_point(g, amp_bead, N, b0, err);
    if (*err) return 0;
    b1 = imax(b1, 1);
    int64_t ind0, indf;
    or_check_bead_bounds(b0, b1, N, &ind0, &indf);
    int64_t n = indf - ind0;
    for (int64_t i = 0; i < n; ++i) s->inds[i] = ind0 + i;
    /* axis */
    int64_t indf_m1 = indf - 1, ind0_m1 = ind0 - 1, nbm1 = N - 1;
    const double *ra, *rb;
    if (ind0 == indf_m1 && ind0 == 0) {
        ra = s->r + 3 * indf;
        rb = s->r + 3 * ind0;
    } else if (ind0 == indf_m1 && ind0 == nbm1) {
        ra = s->r + 3 * ind0;
        rb = s->r + 3 * ind0_m1;
    } else if (ind0 == 0 && indf == N) {
        ra = s->r + 3 * indf_m1;
        rb = s->r + 3 * ind0;
    } else if (ind0 == 0) {
        ra = s->r + 3 * indf;
        rb = s->r + 3 * ind0;
    } else if (indf == N) {
        ra = s->r + 3 * indf_m1;
        rb = s->r + 3 * ind0_m1;
    } else {
        ra = s->r + 3 * indf;
        rb = s->r + 3 * ind0_m1;
    }
    for (int i = 0; i < 3; ++i) s->direction[i] = ra[i] - rb[i];
    double mag = sqrt(pow(s->direction[0], 2.0) + pow(s->direction[1], 2.0) +
                      pow(s->direction[2], 2.0));
    if (mag < 1E-5) {
        sample_unit_sphere(g, s->direction);
    } else {
        double sc = 1.0 / mag;
        for (int i = 0; i < 3; ++i) s->direction[i] = s->direction[i] * sc;
    }
    /* fulcrum */
    const double *pf;
    if (ind0 == 0 && indf != N) pf = s->r + 3 * indf;
    else if (ind0 != 0 && indf == N) pf = s->r + 3 * ind0_m1;
    else if (ind0 == 0 && indf == N) pf = s->r + 3 * ind0;
    else pf = s->r + 3 * ind0_m1;
    for (int i = 0; i < 3; ++i) s->point[i] = pf[i];
    or_axis_rotation(s->direction, s->point, ang, s->T);
    apply_transform(s, s->inds, n);
    return n;
}

/* end_pivot + get_end_pivot_fulcrum, move_funcs.pyx:285-398 */
static int64_t move_end_pivot(or_sim *s, double amp_move, int64_t amp_bead, int *err)
{
    rng_t *g = &s->rng;
    int64_t N = s->N;
    double ang = pick_random_amp(g, amp_move);
    int64_t lhs = rng_rand(g) % 2;
    int64_t ind0, indf;
    if (lhs == 1) {
        ind0 = 0;
        indf = from_left(g, amp_bead, N, err) + 1;
    } else {
        ind0 = from_right(g, amp_bead, N, err);
        indf = N;
    }
    if (*err) return 0;
    int64_t n = indf - ind0;
    for (int64_t i = 0; i < n; ++i) s->inds[i] = ind0 + i;
    double axis[3];
    sample_unit_sphere(g, axis);
    const double *pf;
    if (ind0 == 0 && indf != N) pf = s->r + 3 * indf;
    else if (ind0 != 0 && indf == N) pf = s->r + 3 * (ind0 - 1);
    else if (ind0 == 0 && indf == N && lhs == 1) pf = s->r + 3 * (indf - 1);
    else pf = s->r + 3 * ind0;
    double fulcrum[3] = {pf[0], pf[1], pf[2]};
    or_axis_rotation(axis, fulcrum, ang, s->T);
    apply_transform(s, s->inds, n);
    return n;
}

/* slide, move_funcs.pyx:403-465 */
static int64_t move_slide(or_sim *s, double amp_move, int64_t amp_bead, int *err)
{
    rng_t *g = &s->rng;
    int64_t N = s->N;
    double amp = amp_move * ((double)rng_rand(g) / RAND_MAX_I);
    sample_unit_sphere(g, s->direction);
    for (int i = 0; i < 3; ++i) s->direction[i] *= amp;
    or_translation(s->direction, s->T);
    int64_t b0 = rng_rand(g) % N;
    int64_t b1 = from_point(g, amp_bead, N, b0, err);
    if (*err) return 0;
    int64_t ind0, indf;
    or_check_bead_bounds(b0, b1, N, &ind0, &indf);
    int64_t n = indf - ind0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = ind0 + i;
        s->inds[i] = b;
        for (int j = 0; j < 3; ++j) {
            s->r_trial[3 * b + j] = s->r[3 * b + j] + s->T[j * 4 + 3];
            s->t3_trial[3 * b + j] = s->t3[3 * b + j];
            s->t2_trial[3 * b + j] = s->t2[3 * b + j];
        }
    }
    return n;
}

/* tangent_rotation + rotate_select_beads + get_inds, move_funcs.pyx:470-582 */
static int64_t move_tangent(or_sim *s, double amp_move, int64_t amp_bead, int *err)
{
    rng_t *g = &s->rng;
    int64_t N = s->N;
    double ang = pick_random_amp(g, amp_move);
    int64_t k = rng_rand(g) % amp_bead + 1;
    if (k > N) {
        *err = 2;   /* the reference would loop forever drawing distinct indices */
        return 0;
    }
    pyset picked;
    pyset_init(&picked);
    for (int64_t i = 0; i < k; ++i) {
        for (;;) {
            int64_t c = rng_rand(g) % N;
            if (!pyset_contains(&picked, c)) {
                s->inds[i] = c;
                pyset_add(&picked, c);
                break;
            }
        }
    }
    pyset_free(&picked);
    static const double origin[3] = {0., 0., 0.};
    for (int64_t i = 0; i < k; ++i) {
        int64_t b = s->inds[i];
        sample_unit_sphere(g, s->direction);
        or_axis_rotation(s->direction, origin, ang, s->T);
        for (int j = 0; j < 3; ++j) {
            double e3 = 0, e2 = 0;
            for (int q = 0; q < 3; ++q) {
                e3 += s->T[j * 4 + q] * s->t3[3 * b + q];
                e2 += s->T[j * 4 + q] * s->t2[3 * b + q];
            }
            s->t3_trial[3 * b + j] = e3;
            s->t2_trial[3 * b + j] = e2;
        }
    }
    return k;
}

/* change_binding_state + conduct_change_binding_states + get_new_state,
 * move_funcs.pyx:717-820 */
static int64_t move_binding(or_sim *s, double amp_move, int64_t amp_bead, int *err)
{
    (void)amp_move;
    rng_t *g = &s->rng;
    int64_t N = s->N;
    int64_t binder = rng_rand(g) % s->B;
    int64_t n_tails = s->sites[binder];
    int64_t b0 = rng_rand(g) % N;
    int64_t b1 = from_point(g, amp_bead, N, b0, err);
    if (*err) return 0;
    int64_t ind0, indf;
    or_check_bead_bounds(b0, b1, N, &ind0, &indf);
    int64_t n = indf - ind0;
    for (int64_t i = 0; i < n; ++i) {
        s->inds[i] = ind0 + i;
        s->states_trial[(ind0 + i) * s->B + binder] = rng_randint(g, n_tails + 1);
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* Polymer energy: chromo/polymers.pyx (SSWLC)                                */
/* ------------------------------------------------------------------------- */

static inline double dot3(const double *a, const double *b)
{
    double d = 0;
    for (int i = 0; i < 3; ++i) d += a[i] * b[i];
    return d;
}

/* E_pair, polymers.pyx:1148-1175 */
static inline double E_pair(const or_sim *s, const double *bend, double dr_par, const double *dr_perp,
                            int64_t b)
{
    return 0.5 * s->eps_bend[b] * dot3(bend, bend) +
           0.5 * s->eps_par[b] * pow(dr_par - s->gamma[b], 2.0) +
           0.5 * s->eps_perp[b] * dot3(dr_perp, dr_perp);
}

/* ---- twist: SSTWLC, polymers.pyx:1889-2320 ---- */

/* polymers.pyx:38-40 */
static const double LENGTH_BP = 0.332;
static const double NATURAL_TWIST_BARE = 2 * M_PI / 10.5;

/* compute_twist_angle_omega, polymers.pyx:3427-3461 (np.dot of 3-vectors taken
 * as the left-to-right sum; np.arctan2 is glibc atan2) */
double or_twist_omega(const double *t2_0, const double *t3_0, const double *t2_1, const double *t3_1)
{
    double t1_0[3] = {t2_0[1] * t3_0[2] - t2_0[2] * t3_0[1], t2_0[2] * t3_0[0] - t2_0[0] * t3_0[2],
                      t2_0[0] * t3_0[1] - t2_0[1] * t3_0[0]};
    double t1_1[3] = {t2_1[1] * t3_1[2] - t2_1[2] * t3_1[1], t2_1[2] * t3_1[0] - t2_1[0] * t3_1[2],
                      t2_1[0] * t3_1[1] - t2_1[1] * t3_1[0]};
    return atan2(dot3(t2_0, t1_1) - dot3(t1_0, t2_1), dot3(t1_0, t1_1) + dot3(t2_0, t2_1));
}

/* Python float a // b (Objects/floatobject.c float_floor_div): E_pair_with_twist
 * evaluates its wrap with np.pi operands, i.e. on Python floats. */
static double py_floor_div(double vx, double wx)
{
    double mod = fmod(vx, wx);
    double div = (vx - mod) / wx;
    if (mod) {
        if ((wx < 0) != (mod < 0)) div -= 1.0;
    }
    double floordiv;
    if (div) {
        floordiv = floor(div);
        if (div - floordiv > 0.5) floordiv += 1.0;
    } else {
        floordiv = copysign(0.0, vx / wx);
    }
    return floordiv;
}

/* E_pair_with_twist, polymers.pyx:2050-2102 */
static inline double E_pair_tw(const or_sim *s, const double *bend, double dr_par, const double *dr_perp,
                               double omega, int64_t b)
{
    double d = omega - (s->bead_length[b] * NATURAL_TWIST_BARE / LENGTH_BP);
    d -= 2 * M_PI * py_floor_div(d + M_PI, 2 * M_PI);
    return 0.5 * s->eps_bend[b] * dot3(bend, bend) +
           0.5 * s->eps_par[b] * pow(dr_par - s->gamma[b], 2.0) +
           0.5 * s->eps_perp[b] * dot3(dr_perp, dr_perp) +
           0.5 * s->eps_twist[b] * pow(d, 2.0);
}

/* bead_pair_dE_poly_forward_with_twist, polymers.pyx:2104-2175 */
static double pair_dE_forward_tw(const or_sim *s, const double *r0, const double *r1, const double *tr1,
                                 const double *t30, const double *t31, const double *tt31,
                                 const double *t20, const double *t21, const double *tt21, int64_t b)
{
    double omega = or_twist_omega(t20, t30, t21, t31);
    double omega_t = or_twist_omega(t20, t30, tt21, tt31);
    double dr_t[3], dr[3], drp_t[3], drp[3], bend_t[3], bend[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = tr1[i] - r0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(t30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - t30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        bend_t[i] = tt31[i] - t30[i] - drp_t[i] * s->eta[b];
        bend[i] = t31[i] - t30[i] - drp[i] * s->eta[b];
    }
    return E_pair_tw(s, bend_t, par_t, drp_t, omega_t, b) - E_pair_tw(s, bend, par, drp, omega, b);
}

/* bead_pair_dE_poly_reverse_with_twist, polymers.pyx:2177-2248 */
static double pair_dE_reverse_tw(const or_sim *s, const double *r0, const double *tr0, const double *r1,
                                 const double *t30, const double *tt30, const double *t31,
                                 const double *t20, const double *tt20, const double *t21, int64_t b)
{
    double omega = or_twist_omega(t20, t30, t21, t31);
    double omega_t = or_twist_omega(tt20, tt30, t21, t31);
    double dr_t[3], dr[3], drp_t[3], drp[3], bend_t[3], bend[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = r1[i] - tr0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(tt30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - tt30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        bend_t[i] = t31[i] - tt30[i] - drp_t[i] * s->eta[b];
        bend[i] = t31[i] - t30[i] - drp[i] * s->eta[b];
    }
    return E_pair_tw(s, bend_t, par_t, drp_t, omega_t, b) - E_pair_tw(s, bend, par, drp, omega, b);
}

/* bead_pair_dE_poly_forward, polymers.pyx:1177-1277 */
static double pair_dE_forward(const or_sim *s, const double *r0, const double *r1, const double *tr1,
                              const double *t30, const double *t31, const double *tt31, int64_t b)
{
    double dr_t[3], dr[3], drp_t[3], drp[3], bend_t[3], bend[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = tr1[i] - r0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(t30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - t30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        bend_t[i] = tt31[i] - t30[i] - drp_t[i] * s->eta[b];
        bend[i] = t31[i] - t30[i] - drp[i] * s->eta[b];
    }
    return E_pair(s, bend_t, par_t, drp_t, b) - E_pair(s, bend, par, drp, b);
}

/* bead_pair_dE_poly_reverse, polymers.pyx:1279-1346 */
static double pair_dE_reverse(const or_sim *s, const double *r0, const double *tr0, const double *r1,
                              const double *t30, const double *tt30, const double *t31, int64_t b)
{
    double dr_t[3], dr[3], drp_t[3], drp[3], bend_t[3], bend[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = r1[i] - tr0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(tt30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - tt30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        bend_t[i] = t31[i] - tt30[i] - drp_t[i] * s->eta[b];
        bend[i] = t31[i] - t30[i] - drp[i] * s->eta[b];
    }
    return E_pair(s, bend_t, par_t, drp_t, b) - E_pair(s, bend, par, drp, b);
}

/* continuous_dE_poly, polymers.pyx:1084-1146 (SSTWLC: polymers.pyx:2003-2044) */
static double continuous_dE(const or_sim *s, int64_t ind0, int64_t indf)
{
    double dE = 0;
    if (s->twist) {
        if (ind0 != 0) {
            int64_t a = ind0 - 1;
            dE += pair_dE_forward_tw(s, s->r + 3 * a, s->r + 3 * ind0, s->r_trial + 3 * ind0,
                                     s->t3 + 3 * a, s->t3 + 3 * ind0, s->t3_trial + 3 * ind0,
                                     s->t2 + 3 * a, s->t2 + 3 * ind0, s->t2_trial + 3 * ind0, a);
        }
        if (indf != s->N) {
            int64_t a = indf - 1;
            dE += pair_dE_reverse_tw(s, s->r + 3 * a, s->r_trial + 3 * a, s->r + 3 * indf,
                                     s->t3 + 3 * a, s->t3_trial + 3 * a, s->t3 + 3 * indf,
                                     s->t2 + 3 * a, s->t2_trial + 3 * a, s->t2 + 3 * indf, a);
        }
        return dE;
    }
    if (ind0 != 0) {
        int64_t a = ind0 - 1;
        dE += pair_dE_forward(s, s->r + 3 * a, s->r + 3 * ind0, s->r_trial + 3 * ind0, s->t3 + 3 * a,
                              s->t3 + 3 * ind0, s->t3_trial + 3 * ind0, a);
    }
    if (indf != s->N) {
        int64_t a = indf - 1;
        dE += pair_dE_reverse(s, s->r + 3 * a, s->r_trial + 3 * a, s->r + 3 * indf, s->t3 + 3 * a,
                              s->t3_trial + 3 * a, s->t3 + 3 * indf, a);
    }
    return dE;
}

/* bead_binding_dE, polymers.pyx:1408-1538; the -log(single-site Z) values come
 * from host-built tables F[b][marks][state] (same numpy/scipy calls as the
 * reference, computed once). */
static double bead_binding_dE(const or_sim *s, int64_t ind, const int64_t *st_trial)
{
    const int64_t *mods = s->marks + ind * s->B;
    const int64_t *st = s->states + ind * s->B;
    double dE = 0;
    if (s->max_binders != -1) {
        int64_t tot = 0;
        for (int64_t h = 0; h < s->B; ++h) tot += st_trial[h];
        if (tot > s->max_binders) dE += E_HUGE_POLY * (double)(tot - s->max_binders);
        tot = 0;
        for (int64_t h = 0; h < s->B; ++h) tot += st[h];
        if (tot > s->max_binders) dE -= E_HUGE_POLY * (double)(tot - s->max_binders);
    }
    double f = s->mu_adjust_factor;
    for (int64_t b = 0; b < s->B; ++b) {
        dE += s->F[b][mods[b]][st_trial[b]];
        dE -= s->F[b][mods[b]][st[b]];
        double mu = s->mu[b];
        if (mu > 0) {
            dE -= (double)st_trial[b] * (mu * (-f + 2));
            dE += (double)st[b] * (mu * (-f + 2));
        } else {
            dE -= (double)st_trial[b] * mu * f;
            dE += (double)st[b] * mu * f;
        }
    }
    return dE;
}

/* SSWLC.compute_dE, polymers.pyx:1024-1082 */
static double poly_dE(or_sim *s, int move, const int64_t *inds, int64_t n)
{
    double dE = 0;
    if (move == OR_BINDING) {
        int64_t ind0 = inds[0];
        for (int64_t i = 0; i < n; ++i)
            dE += bead_binding_dE(s, ind0 + i, s->states_trial + (ind0 + i) * s->B);
    } else if (move == OR_TANGENT) {
        for (int64_t i = 0; i < n; ++i) dE += continuous_dE(s, inds[i], inds[i] + 1);
    } else {
        dE += continuous_dE(s, inds[0], inds[n - 1] + 1);
    }
    return dE;
}

/* ------------------------------------------------------------------------- */
/* Field: chromo/fields.pyx (UniformDensityField)                             */
/* ------------------------------------------------------------------------- */

/* get_confinement_dE, chromo/fields.pyx:121-202 */
static double confinement_E(const or_sim *s, const double *pos, const int64_t *inds, int64_t n,
                            int trial)
{
    int64_t out = 0;
    if (s->confine_type == 0) return 0.;
    if (s->confine_type == 1) {
        for (int64_t i = 0; i < n; ++i) {
            const double *p = pos + 3 * inds[i];
            double dist = sqrt(dot3(p, p));
            if (dist > s->confine_length) out += 1;
        }
        return (double)out * E_HUGE_FIELD;
    }
    /* Cubical: the current-configuration branch never counts (fields.pyx:188-193) */
    if (trial == 1) {
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < 3; ++j)
                if (fabs(pos[3 * inds[i] + j]) > s->confine_length / 2) out += 1;
    }
    return (double)out * E_HUGE_FIELD;
}

/* One bead's 8 lower/upper voxel indices and trilinear weights for one position
 * (fields.pyx:1434-1462, _generate_weight_vector_with_trial :1524-1598,
 * _generate_index_vector_with_trial :1600-1673). */
static void bead_voxels(const or_sim *s, const double *pos, int64_t idx8[8], double wt8[8])
{
    int64_t ix[3];
    double w[3];
    for (int j = 0; j < 3; ++j) {
        double x = pymod(pos[j] + s->half_width[j], s->width[j]) - s->half_step[j];
        int64_t ind = (int64_t)floor(x / s->dxyz[j]);
        ix[j] = (ind == -1) ? s->n_xyz_m1[j] : ind;
        w[j] = 1 - (x / s->dxyz[j] - (double)ind);
    }
    wt8[0] = w[0] * w[1] * w[2];
    wt8[1] = (1 - w[0]) * w[1] * w[2];
    wt8[2] = w[0] * (1 - w[1]) * w[2];
    wt8[3] = (1 - w[0]) * (1 - w[1]) * w[2];
    wt8[4] = w[0] * w[1] * (1 - w[2]);
    wt8[5] = (1 - w[0]) * w[1] * (1 - w[2]);
    wt8[6] = w[0] * (1 - w[1]) * (1 - w[2]);
    wt8[7] = (1 - w[0]) * (1 - w[1]) * (1 - w[2]);
    idx8[0] = wrap_index(s, ix[0], ix[1], ix[2]);
    idx8[1] = wrap_index(s, 1 + ix[0], ix[1], ix[2]);
    idx8[2] = wrap_index(s, ix[0], 1 + ix[1], ix[2]);
    idx8[3] = wrap_index(s, 1 + ix[0], 1 + ix[1], ix[2]);
    idx8[4] = wrap_index(s, ix[0], ix[1], 1 + ix[2]);
    idx8[5] = wrap_index(s, 1 + ix[0], ix[1], 1 + ix[2]);
    idx8[6] = wrap_index(s, ix[0], 1 + ix[1], 1 + ix[2]);
    idx8[7] = wrap_index(s, 1 + ix[0], 1 + ix[1], 1 + ix[2]);
}

/* bead_voxels under the fast field: each axis' sub-bin floor((x + W/2) / sbw)
 * mod (n_points n) (get_change_in_density_quickly, fields.pyx:1255-1300) and that
 * sub-bin's lower voxel and weight as init_fast_field tabulates them
 * (fields.pyx:614-671), evaluated from the same expressions. */
static void bead_voxels_fast(const or_sim *s, const double *pos, int64_t idx8[8], double wt8[8])
{
    const int64_t nm1[3] = {s->nx - 1, s->ny - 1, s->nz - 1};
    const double half = (double)(s->n_points / 2);
    int64_t ix[3];
    double w[3];
    for (int j = 0; j < 3; ++j) {
        double sb = floor((pos[j] + s->half_width[j]) / s->sub_w[j]);
        int64_t i = (int64_t)sb % s->n_sub[j];
        if (i < 0) i += s->n_sub[j];
        double ind = floor((((double)i - half) * s->sub_w[j]) / s->dxyz[j]);
        ix[j] = (ind == -1) ? nm1[j] : (int64_t)ind;
        double upper = (((double)i - half) * s->sub_w[j]) / s->dxyz[j] - ind;
        w[j] = 1 - upper;
    }
    wt8[0] = w[0] * w[1] * w[2];
    wt8[1] = (1 - w[0]) * w[1] * w[2];
    wt8[2] = w[0] * (1 - w[1]) * w[2];
    wt8[3] = (1 - w[0]) * (1 - w[1]) * w[2];
    wt8[4] = w[0] * w[1] * (1 - w[2]);
    wt8[5] = (1 - w[0]) * w[1] * (1 - w[2]);
    wt8[6] = w[0] * (1 - w[1]) * (1 - w[2]);
    wt8[7] = (1 - w[0]) * (1 - w[1]) * (1 - w[2]);
    idx8[0] = wrap_index(s, ix[0], ix[1], ix[2]);
    idx8[1] = wrap_index(s, 1 + ix[0], ix[1], ix[2]);
    idx8[2] = wrap_index(s, ix[0], 1 + ix[1], ix[2]);
    idx8[3] = wrap_index(s, 1 + ix[0], 1 + ix[1], ix[2]);
    idx8[4] = wrap_index(s, ix[0], ix[1], 1 + ix[2]);
    idx8[5] = wrap_index(s, 1 + ix[0], ix[1], 1 + ix[2]);
    idx8[6] = wrap_index(s, ix[0], 1 + ix[1], 1 + ix[2]);
    idx8[7] = wrap_index(s, 1 + ix[0], 1 + ix[1], 1 + ix[2]);
}

/* get_change_in_density, fields.pyx:1370-1522.  Fills density_trial rows of the
 * touched voxels and s->bins in CPython set order; returns the bin count. */
static int64_t change_in_density(or_sim *s, const double *r_trial, const int64_t *st_trial,
                                 const int64_t *inds, int64_t n, int state_change)
{
    int64_t Bp1 = s->Bp1, B = s->B;
    pyset_clear(&s->found);
    double dt[2][OR_MAXB + 1][8];
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = inds[i];
        int64_t idx[2][8];
        double wt[2][8];
        if (s->fast_field) {
            bead_voxels_fast(s, s->r + 3 * b, idx[0], wt[0]);
            if (state_change == 0) bead_voxels_fast(s, r_trial + 3 * b, idx[1], wt[1]);
        } else {
            bead_voxels(s, s->r + 3 * b, idx[0], wt[0]);
            if (state_change == 0) bead_voxels(s, r_trial + 3 * b, idx[1], wt[1]);
        }
        if (state_change != 0) {
            memcpy(idx[1], idx[0], sizeof(idx[0]));
            memcpy(wt[1], wt[0], sizeof(wt[0]));
        }
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 8; ++l) {
                dt[k][0][l] = wt[k][l] / s->access_vols[idx[k][l]];
                for (int64_t m = 1; m < Bp1; ++m) {
                    int64_t st = (k == 0 || state_change == 0) ? s->states[b * B + m - 1]
                                                               : st_trial[b * B + m - 1];
                    dt[k][m][l] = dt[k][0][l] * (double)st;
                }
            }
        for (int k = 0; k < 2; ++k) {
            double pre = (k == 0) ? -1. : 1.;
            for (int l = 0; l < 8; ++l) {
                int64_t bin = idx[k][l];
                double *row = s->density_trial + bin * Bp1;
                /* the fast path (fields.pyx:1346-1366) keeps every contribution */
                const int thr = !s->fast_field;
                if (pyset_contains(&s->found, bin)) {
                    for (int64_t m = 0; m < Bp1; ++m) {
                        double temp = pre * dt[k][m][l];
                        if (!thr || fabs(temp) > 1E-18) row[m] += temp;
                    }
                } else {
                    pyset_add(&s->found, bin);
                    for (int64_t m = 0; m < Bp1; ++m) {
                        double temp = pre * dt[k][m][l];
                        row[m] = (!thr || fabs(temp) > 1E-18) ? temp : 0;
                    }
                }
            }
        }
    }
    if (s->found.nused > s->bins_cap) {
        s->bins_cap = s->found.nused * 2;
        s->bins = (int64_t *)realloc(s->bins, (size_t)s->bins_cap * sizeof(int64_t));
    }
    return pyset_list(&s->found, s->bins);
}

/* get_dE_binders_and_beads + count_doubly_bound + nonspecific_interact_dE +
 * get_volume_fractions_with_trial, fields.pyx:1675-1937 */
static double dE_binders_and_beads(const or_sim *s, const int64_t *st_trial, const int64_t *inds,
                                   int64_t n, const int64_t *bins, int64_t nb, int state_change)
{
    int64_t B = s->B, Bp1 = s->Bp1;
    double tot[OR_MAXB];
    double totx[OR_MAXB * OR_MAXB];
    int totx_set[OR_MAXB * OR_MAXB];
    for (int64_t j = 0; j < B; ++j) tot[j] = 0.0;
    for (int64_t j = 0; j < B * B; ++j) {
        totx[j] = 0.0;
        totx_set[j] = 0;
    }
    /* delta_rho_squared / delta_rho_interact_squared, then the per-binder sums
     * in bin order (the reference fills full arrays first; sums are the same
     * left folds in the same bin order). */
    for (int64_t i = 0; i < nb; ++i) {
        const double *rho = s->density + bins[i] * Bp1;
        const double *drho = s->density_trial + bins[i] * Bp1;
        for (int64_t j = 0; j < B; ++j) {
            double d2 = pow(rho[j + 1] + drho[j + 1], 2.0) - pow(rho[j + 1], 2.0);
            if (fabs(d2) < 1E-18) d2 = 0;
            tot[j] += d2;
            for (int64_t k = 0; k < B; ++k) {
                double dx = ((rho[j + 1] + drho[j + 1]) * (rho[k + 1] + drho[k + 1])) -
                            (rho[j + 1] * rho[k + 1]);
                if (fabs(dx) < 1E-18) dx = 0;
                totx[j * B + k] += dx;
                totx_set[j * B + k] = 1;
            }
        }
    }
    int64_t db_trial[OR_MAXB], db[OR_MAXB];
    for (int64_t j = 0; j < B; ++j) {
        db_trial[j] = 0;
        db[j] = 0;
        for (int64_t i = 0; i < n; ++i) {
            int64_t b = inds[i];
            int64_t stt = state_change ? st_trial[b * B + j] : s->states[b * B + j];
            if (stt == 2) db_trial[j] += 1;
            if (s->states[b * B + j] == 2) db[j] += 1;
        }
    }
    double dE = 0.0;
    for (int64_t i = 0; i < B; ++i) {
        dE += s->pref[i] * tot[i];
        dE += s->intra[i] * (double)(db_trial[i] - db[i]);
    }
    for (int64_t i = 0; i < B; ++i)
        for (int64_t j = 0; j < B; ++j) {
            /* tot_density_change_interact starts as the Python int 0 */
            double t = totx_set[i * B + j] ? totx[i * B + j] : 0.0;
            dE += s->cross[i * B + j] * t;
        }
    /* nonspecific_interact_dE */
    double ns = 0;
    double V = s->bead_vol;
    for (int64_t i = 0; i < nb; ++i) {
        double av = s->access_vols[bins[i]];
        double vf0 = s->density[bins[i] * Bp1] * V;
        double vf1 = vf0 + (s->density_trial[bins[i] * Bp1] * V);
        if (vf1 > (double)s->vf_limit) ns += E_HUGE_FIELD * vf1;
        else ns += s->chi * (av / V) * pow(vf1, 2.0);
        if (vf0 > (double)s->vf_limit) ns -= E_HUGE_FIELD * vf0;
        else ns -= s->chi * (av / V) * pow(vf0, 2.0);
    }
    dE += ns;
    return dE;
}

/* UniformDensityField.compute_dE, fields.pyx:1149-1233 */
static double field_dE(or_sim *s, const int64_t *inds, int64_t n, int state_change, int64_t *nb_out)
{
    double dE = 0;
    if (state_change == 0) {
        dE += confinement_E(s, s->r_trial, inds, n, 1);
        dE -= confinement_E(s, s->r, inds, n, 0);
    }
    int64_t nb = change_in_density(s, s->r_trial, s->states_trial, inds, n, state_change);
    if (s->faithful) {
        for (int64_t i = 0; i < s->n_bins; ++i) s->affected[i] = 0;
        for (int64_t i = 0; i < nb; ++i) s->affected[s->bins[i]] = 1;
    }
    dE += dE_binders_and_beads(s, s->states_trial, inds, n, s->bins, nb, state_change);
    *nb_out = nb;
    return dE;
}

/* update_affected_densities, fields.pyx:1968-1975 */
static void update_affected_densities(or_sim *s, int64_t nb)
{
    int64_t Bp1 = s->Bp1;
    if (s->faithful) {
        for (int64_t i = 0; i < s->n_bins; ++i)
            if (s->affected[i] == 1)
                for (int64_t j = 0; j < Bp1; ++j) {
                    s->density[i * Bp1 + j] += s->density_trial[i * Bp1 + j];
                    s->density_trial[i * Bp1 + j] = 0;
                }
    } else {
        for (int64_t q = 0; q < nb; ++q) {
            int64_t i = s->bins[q];
            for (int64_t j = 0; j < Bp1; ++j) {
                s->density[i * Bp1 + j] += s->density_trial[i * Bp1 + j];
                s->density_trial[i * Bp1 + j] = 0;
            }
        }
    }
}

/* update_all_densities_for_all_polymers, fields.pyx:2041-2106 */
void or_density_init(or_sim *s)
{
    int64_t Bp1 = s->Bp1, B = s->B;
    memset(s->density, 0, (size_t)s->n_bins * Bp1 * sizeof(double));
    memset(s->density_trial, 0, (size_t)s->n_bins * Bp1 * sizeof(double));
    for (int64_t i = 0; i < s->N; ++i) {
        int64_t idx[8];
        double wt[8];
        bead_voxels(s, s->r + 3 * i, idx, wt);
        for (int l = 0; l < 8; ++l) {
            double d = wt[l] / s->access_vols[idx[l]];
            s->density[idx[l] * Bp1] += d;
            for (int64_t m = 1; m < Bp1; ++m)
                s->density[idx[l] * Bp1 + m] += d * (double)s->states[i * B + m - 1];
        }
    }
    for (int64_t i = 0; i < s->n_bins * Bp1; ++i)
        if (fabs(s->density[i]) < 1E-18) s->density[i] = 0;
}

/* ------------------------------------------------------------------------- */
/* mc_step / mc_sim / controllers                                             */
/* ------------------------------------------------------------------------- */

/* AcceptanceTracker.update_acceptance_rate, chromo/util/mc_stat.py:190-208 */
static inline void ewma(or_ctrl *c, double accept)
{
    c->rate = (c->alpha * accept) + (1 - c->alpha) * c->rate;
}

/* MCAdapter.accept, chromo/mc/moves.pyx:156-239 */
static void accept_move(or_sim *s, int move, const int64_t *inds, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = inds[i];
        if (move == OR_BINDING) {
            for (int64_t j = 0; j < s->B; ++j) s->states[b * s->B + j] = s->states_trial[b * s->B + j];
        } else if (move == OR_SLIDE) {
            for (int j = 0; j < 3; ++j) {
                s->r[3 * b + j] = s->r_trial[3 * b + j];
                s->t3_trial[3 * b + j] = s->t3[3 * b + j];
                s->t2_trial[3 * b + j] = s->t2[3 * b + j];
            }
        } else if (move == OR_TANGENT) {
            for (int j = 0; j < 3; ++j) {
                s->t3[3 * b + j] = s->t3_trial[3 * b + j];
                s->t2[3 * b + j] = s->t2_trial[3 * b + j];
                s->r_trial[3 * b + j] = s->r[3 * b + j];
            }
        } else {
            for (int j = 0; j < 3; ++j) {
                s->r[3 * b + j] = s->r_trial[3 * b + j];
                s->t3[3 * b + j] = s->t3_trial[3 * b + j];
                s->t2[3 * b + j] = s->t2_trial[3 * b + j];
            }
        }
    }
}

/* MCAdapter.reject, chromo/mc/moves.pyx:241-299 */
static void reject_move(or_sim *s, int move, const int64_t *inds, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = inds[i];
        if (move == OR_BINDING) {
            for (int64_t j = 0; j < s->B; ++j) s->states_trial[b * s->B + j] = s->states[b * s->B + j];
        } else {
            for (int j = 0; j < 3; ++j) {
                s->r_trial[3 * b + j] = s->r[3 * b + j];
                s->t3_trial[3 * b + j] = s->t3[3 * b + j];
                s->t2_trial[3 * b + j] = s->t2[3 * b + j];
            }
        }
    }
}

/* mc_step, chromo/mc/mc_sim.pyx:106-182 (+ MCAdapter.propose, moves.pyx:137-154) */
int or_mc_step(or_sim *s, or_ctrl *c, or_log *log)
{
    int err = 0;
    int move = c->move;
    int check_field = s->field_on && move != OR_TANGENT;
    c->num_attempt += 1;
    int64_t n;
    switch (move) {
    case OR_CRANK: n = move_crank(s, c->amp_move, c->amp_bead, &err); break;
    case OR_END_PIVOT: n = move_end_pivot(s, c->amp_move, c->amp_bead, &err); break;
    case OR_SLIDE: n = move_slide(s, c->amp_move, c->amp_bead, &err); break;
    case OR_TANGENT: n = move_tangent(s, c->amp_move, c->amp_bead, &err); break;
    default: n = move_binding(s, c->amp_move, c->amp_bead, &err); break;
    }
    if (err) return -err;
    if (n == 0) return 0;
    const int64_t *inds = s->inds;
    double dE = 0;
    double dEp = poly_dE(s, move, inds, n);
    dE += dEp;
    double dEf = 0;
    int64_t nb = 0;
    if (check_field) {
        dEf = field_dE(s, inds, n, move == OR_BINDING, &nb);
        dE += dEf;
    }
    double exp_dE = exp(-dE);
    double u = (double)rng_rand(&s->rng) / RAND_MAX_I;
    int acc = u < exp_dE;
    if (acc) {
        accept_move(s, move, inds, n);
        c->num_success += 1;
        ewma(c, 1.0);
        if (check_field) update_affected_densities(s, nb);
    } else {
        reject_move(s, move, inds, n);
        ewma(c, 0.0);
    }
    if (log) {
        log->move = move;
        log->accepted = acc;
        log->n = n;
        /* a tangent rotation's picks are not a window: its record holds the
         * first picked bead, as the device's */
        log->ind0 = inds[0];
        log->indf = move == OR_TANGENT ? inds[0] + 1 : inds[n - 1] + 1;
        log->dE_poly = dEp;
        log->dE_field = dEf;
        log->u = u;
        log->exp_dE = exp_dE;
        log->n_bins = nb;
    }
    return 1;
}

/* SimpleControl.update_move_amplitude / update_bead_amplitude,
 * chromo/mc/mc_controller.py:148-213 (setpoint 0.5, factor 0.95, 1 bead). */
void or_update_amplitudes(or_ctrl *c)
{
    if (!c->controlled) return;
    double acc = c->rate;
    if (acc < 0.5) {
        double prop = c->amp_move * 0.95;
        if (prop > c->move_lo) c->amp_move = prop;
        else {
            c->amp_move = c->move_hi;
            double nb = (double)(c->amp_bead - 1);
            c->amp_bead = (int64_t)(c->bead_lo >= nb ? c->bead_lo : nb);
        }
    } else if (acc > 0.5) {
        double prop = c->amp_move / 0.95;
        if (prop < c->move_hi) c->amp_move = prop;
        else {
            c->amp_move = c->move_lo;
            double nb = (double)(c->amp_bead + 1);
            c->amp_bead = (int64_t)(c->bead_hi <= nb ? c->bead_hi : nb);
        }
    }
}

/* mc_sim, chromo/mc/mc_sim.pyx:26-103 (single polymer).  The caller reseeds the
 * numpy stream (or_np_seed) as mc_sim does at :81. */
int64_t or_mc_sim(or_sim *s, int64_t num_steps, or_ctrl *ctrls, int32_t n_ctrl,
                  double mu_adjust_factor, or_log *logs, int64_t max_logs, int64_t *n_logs)
{
    int64_t nl = 0;
    s->mu_adjust_factor = mu_adjust_factor;
    for (int64_t k = 0; k < num_steps; ++k) {
        for (int32_t ci = 0; ci < n_ctrl; ++ci) {
            or_ctrl *c = &ctrls[ci];
            if (c->on == 1) {
                for (int32_t j = 0; j < c->num_per_cycle; ++j) {
                    or_log *lg = (logs && nl < max_logs) ? &logs[nl] : NULL;
                    int rc = or_mc_step(s, c, lg);
                    if (rc < 0) {
                        if (n_logs) *n_logs = nl;
                        return rc;
                    }
                    if (rc == 1) ++nl;
                }
            }
            or_update_amplitudes(c);
        }
    }
    if (n_logs) *n_logs = nl;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Tape entry points (test vectors for the HIP kernels)                       */
/* ------------------------------------------------------------------------- */

double or_field_dE_tape(or_sim *s, const int64_t *inds, int64_t n, const double *r_trial,
                        const int64_t *states_trial, int32_t state_change, double *delta_rho,
                        int64_t *bins, int64_t *n_bins_out)
{
    /* trial arrays are full-length; copy into scratch trial buffers */
    double *rt = (double *)malloc((size_t)s->N * 3 * sizeof(double));
    int64_t *stt = (int64_t *)malloc((size_t)s->N * (s->B > 0 ? s->B : 1) * sizeof(int64_t));
    memcpy(rt, s->r, (size_t)s->N * 3 * sizeof(double));
    if (s->B > 0) memcpy(stt, s->states, (size_t)s->N * s->B * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        int64_t b = inds[i];
        for (int j = 0; j < 3; ++j) rt[3 * b + j] = r_trial[3 * i + j];
        for (int64_t j = 0; j < s->B; ++j) stt[b * s->B + j] = states_trial[i * s->B + j];
    }
    double dE = 0;
    if (state_change == 0) {
        dE += confinement_E(s, rt, inds, n, 1);
        dE -= confinement_E(s, s->r, inds, n, 0);
    }
    int64_t nb = change_in_density(s, rt, stt, inds, n, state_change);
    dE += dE_binders_and_beads(s, stt, inds, n, s->bins, nb, state_change);
    if (delta_rho) memset(delta_rho, 0, (size_t)s->n_bins * s->Bp1 * sizeof(double));
    for (int64_t q = 0; q < nb; ++q) {
        int64_t i = s->bins[q];
        if (bins) bins[q] = i;
        for (int64_t m = 0; m < s->Bp1; ++m) {
            if (delta_rho) delta_rho[i * s->Bp1 + m] = s->density_trial[i * s->Bp1 + m];
            s->density_trial[i * s->Bp1 + m] = 0;
        }
    }
    *n_bins_out = nb;
    free(rt);
    free(stt);
    return dE;
}

double or_elastic_dE_tape(or_sim *s, int32_t move, const int64_t *inds, int64_t n,
                          const double *r_trial, const double *t3_trial, const double *t2_trial)
{
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            s->r_trial[3 * inds[i] + j] = r_trial[3 * i + j];
            s->t3_trial[3 * inds[i] + j] = t3_trial[3 * i + j];
            if (t2_trial) s->t2_trial[3 * inds[i] + j] = t2_trial[3 * i + j];
        }
    double dE = poly_dE(s, move, inds, n);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            s->r_trial[3 * inds[i] + j] = s->r[3 * inds[i] + j];
            s->t3_trial[3 * inds[i] + j] = s->t3[3 * inds[i] + j];
            s->t2_trial[3 * inds[i] + j] = s->t2[3 * inds[i] + j];
        }
    return dE;
}

double or_binding_dE_tape(or_sim *s, int64_t ind0, int64_t indf, const int64_t *states_trial)
{
    double dE = 0;
    for (int64_t i = ind0; i < indf; ++i)
        dE += bead_binding_dE(s, i, states_trial + (i - ind0) * s->B);
    return dE;
}

/* SSTWLC.compute_E, polymers.pyx:2250-2286 */
static double elastic_E_tw(const or_sim *s)
{
    double E = 0;
    for (int64_t i = 1; i < s->N; ++i) {
        int64_t a = i - 1;
        const double *r0 = s->r + 3 * a, *r1 = s->r + 3 * i, *t30 = s->t3 + 3 * a,
                     *t31 = s->t3 + 3 * i, *t20 = s->t2 + 3 * a, *t21 = s->t2 + 3 * i;
        double omega = or_twist_omega(t20, t30, t21, t31);
        double dr[3], drp[3], bend[3];
        for (int j = 0; j < 3; ++j) dr[j] = r1[j] - r0[j];
        double par = dot3(t30, dr);
        for (int j = 0; j < 3; ++j) drp[j] = dr[j] - t30[j] * par;
        for (int j = 0; j < 3; ++j) {
            bend[j] = t31[j];
            bend[j] += -t30[j] - s->eta[a] * drp[j];
        }
        E += E_pair_tw(s, bend, par, drp, omega, a);
    }
    return E;
}

/* SSWLC.compute_E, polymers.pyx:1348-1381 */
double or_elastic_E(const or_sim *s)
{
    if (s->twist) return elastic_E_tw(s);
    double E = 0;
    for (int64_t i = 1; i < s->N; ++i) {
        int64_t a = i - 1;
        const double *r0 = s->r + 3 * a, *r1 = s->r + 3 * i, *t30 = s->t3 + 3 * a,
                     *t31 = s->t3 + 3 * i;
        double dr[3], drp[3], bend[3];
        for (int j = 0; j < 3; ++j) dr[j] = r1[j] - r0[j];
        double par = dot3(t30, dr);
        for (int j = 0; j < 3; ++j) drp[j] = dr[j] - t30[j] * par;
        for (int j = 0; j < 3; ++j) {
            bend[j] = t31[j];
            bend[j] += -t30[j] - s->eta[a] * drp[j];
        }
        E += E_pair(s, bend, par, drp, a);
    }
    return E;
}

/* UniformDensityField.compute_E -> update_all_densities (no 1e-18 clean-up) +
 * get_E_binders_and_beads + nonspecific_interact_E, fields.pyx:1939-2039, 2208-2315.
 * round(x, 2) is Python's correctly-rounded decimal rounding; the comparison below
 * reproduces it for the vf_limit test. */
static double py_round2(double x)
{
    /* Python round(x, 2) rounds the exact binary value half-to-even at two
     * decimals and returns the double nearest k/100.  x * 100 is exact in the
     * x87 64-bit significand (53 + 7 bits), so nearbyintl sees the exact value. */
    long double y = (long double)x * 100.0L;
    long double z = nearbyintl(y);
    return (double)z / 100.0;
}

/* The truncated exponential of capped_exponential (bead_selection.pyx:58-65) for
 * given draws and windows, glibc log10: the reference side of the exhaustive
 * comparison with the device arithmetic near every integer crossing. */
void or_cexp_trunc(const int32_t *draws, const int64_t *windows, int64_t n, int64_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = cexp_value(draws[i], windows[i]);
}

/* Python's round(x, 2), exposed for its known-answer test against CPython. */
double or_py_round2(double x) { return py_round2(x); }

/* mu_adjust_factor of the next tape evaluation (mc_sim sets it from its
 * argument, chromo/mc/mc_sim.pyx:97). */
void or_set_mu_adjust(or_sim *s, double f) { s->mu_adjust_factor = f; }

double or_field_E(or_sim *s)
{
    int64_t Bp1 = s->Bp1, B = s->B;
    memset(s->density, 0, (size_t)s->n_bins * Bp1 * sizeof(double));
    memset(s->density_trial, 0, (size_t)s->n_bins * Bp1 * sizeof(double));
    for (int64_t i = 0; i < s->N; ++i) {
        int64_t idx[8];
        double wt[8];
        bead_voxels(s, s->r + 3 * i, idx, wt);
        for (int l = 0; l < 8; ++l) {
            double d = wt[l] / s->access_vols[idx[l]];
            s->density[idx[l] * Bp1] += d;
            for (int64_t m = 1; m < Bp1; ++m)
                s->density[idx[l] * Bp1 + m] += d * (double)s->states[i * B + m - 1];
        }
    }
    double E = 0;
    for (int64_t i = 0; i < B; ++i) {
        double tot = 0;
        for (int64_t j = 0; j < s->n_bins; ++j) tot += pow(s->density[j * Bp1 + i + 1], 2.0);
        E += s->pref[i] * tot;
        int64_t nd = 0;
        for (int64_t q = 0; q < s->N; ++q)
            if (s->states[q * B + i] == 2) nd += 1;
        E += s->intra[i] * (double)nd;
    }
    double ns = 0, V = s->bead_vol;
    for (int64_t i = 0; i < s->n_bins; ++i) {
        double vf = s->density[i * Bp1] * V;
        if (py_round2(vf) > (double)s->vf_limit) ns += E_HUGE_FIELD * vf;
        else ns += s->chi * (s->access_vols[i] / V) * vf * (1 - vf);
    }
    E += ns;
    return E;
}

/* ------------------------------------------------------------------------- */
/* Batched CPU baseline                                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    or_sim **sims;
    or_ctrl *ctrls;
    int32_t n_sims, n_ctrl;
    int64_t steps;
    int64_t next;
    int64_t proposals;
    pthread_mutex_t mu;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *jb = (batch_job *)arg;
    int64_t local = 0;
    for (;;) {
        pthread_mutex_lock(&jb->mu);
        int64_t i = jb->next++;
        pthread_mutex_unlock(&jb->mu);
        if (i >= jb->n_sims) break;
        or_sim *s = jb->sims[i];
        or_ctrl *c = jb->ctrls + i * jb->n_ctrl;
        for (int64_t k = 0; k < jb->steps; ++k)
            for (int32_t ci = 0; ci < jb->n_ctrl; ++ci) {
                if (c[ci].on == 1)
                    for (int32_t j = 0; j < c[ci].num_per_cycle; ++j) {
                        int rc = or_mc_step(s, &c[ci], NULL);
                        if (rc == 1) ++local;
                    }
                or_update_amplitudes(&c[ci]);
            }
    }
    pthread_mutex_lock(&jb->mu);
    jb->proposals += local;
    pthread_mutex_unlock(&jb->mu);
    return NULL;
}

int64_t or_run_batch(or_sim **sims, int32_t n_sims, or_ctrl *ctrls, int32_t n_ctrl,
                     int64_t num_steps, int32_t threads, double *seconds_out)
{
    batch_job jb;
    jb.sims = sims;
    jb.ctrls = ctrls;
    jb.n_sims = n_sims;
    jb.n_ctrl = n_ctrl;
    jb.steps = num_steps;
    jb.next = 0;
    jb.proposals = 0;
    pthread_mutex_init(&jb.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, batch_worker, &jb);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(th);
    pthread_mutex_destroy(&jb.mu);
    if (seconds_out) *seconds_out = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return jb.proposals;
}
