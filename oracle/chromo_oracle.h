/*
 * chromo_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C) of the Monte Carlo inner loop of JosephWakim/chromo:
 * move proposal (chromo/mc/move_funcs.pyx, chromo/util/bead_selection.pyx,
 * chromo/util/linalg.pyx), SSWLC elastic and binding dE (chromo/polymers.pyx),
 * UniformDensityField field dE (chromo/fields.pyx), the Metropolis step and the
 * move schedule (chromo/mc/mc_sim.pyx, chromo/mc/moves.pyx,
 * chromo/mc/mc_controller.py, chromo/util/mc_stat.py).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline.  The product
 * (chromo_amd) never links or calls it.
 */
#ifndef CHROMO_ORACLE_H
#define CHROMO_ORACLE_H
#include <stdint.h>

#define OR_MAXB 8          /* binder types */
#define OR_MAXLEV 8        /* mark levels (chemical_mods values 0..OR_MAXLEV-1) */
#define OR_MAXS 8          /* sites_per_bead upper bound; F tables hold s = 0..OR_MAXS */

enum { OR_RNG_GLIBC = 0, OR_RNG_PHILOX = 1 };
enum { OR_CRANK = 0, OR_END_PIVOT = 1, OR_SLIDE = 2, OR_TANGENT = 3, OR_BINDING = 4 };

/* Inputs that describe one replica: chain, binders and field.  All arrays are
 * row-major and copied by or_new(). */
typedef struct {
    int64_t N, B;
    const double *r, *t3, *t2;          /* N x 3 */
    const int64_t *states, *marks;      /* N x B */
    const double *eps_bend, *eps_par, *eps_perp, *gamma, *eta;   /* N-1 each */
    int64_t max_binders;                /* -1 = unlimited */
    int64_t sites[OR_MAXB];
    double mu[OR_MAXB];                 /* chemical_potential per binder */
    int64_t n_mark_levels;              /* F table rows */
    double F[OR_MAXB][OR_MAXLEV][OR_MAXS + 1];  /* -log(single-site Z), host-built */
    /* field */
    int32_t field_on;
    int64_t nx, ny, nz;
    double x_width, y_width, z_width;
    float vf_limit;
    double chi, bead_vol;
    double pref[OR_MAXB], intra[OR_MAXB], cross[OR_MAXB * OR_MAXB];
    int32_t confine_type;               /* 0 none, 1 spherical, 2 cubical */
    double confine_length;
    const double *access_vols;          /* n_bins accessible volumes (copied), or NULL: vol_bin */
    int32_t fast_field;                 /* 1: sub-bin lookup weights in the move dE */
    int64_t n_points;                   /* sub-bins per voxel side (made even) */
    /* SSTWLC (polymers.pyx:1889-2320): twist modulus per bond and the bead lengths
     * the natural twist is taken from; twist = 0 for SSWLC */
    int32_t twist;
    const double *eps_twist, *bead_length;      /* N-1 each */
} or_spec;

typedef struct {
    int32_t move;            /* OR_CRANK .. OR_BINDING */
    int32_t on;              /* move_on */
    int32_t controlled;      /* 1 = SimpleControl, 0 = NoControl */
    int32_t num_per_cycle;
    double amp_move;
    int64_t amp_bead;
    double rate;             /* AcceptanceTracker.acceptance_rate */
    double alpha;            /* 2 / (moves_in_average + 1) */
    double bead_lo, bead_hi, move_lo, move_hi;
    int64_t num_attempt, num_success;
} or_ctrl;

/* One record per proposal that reached dE evaluation (n_inds > 0). */
typedef struct {
    int32_t move;
    int32_t accepted;
    int64_t n, ind0, indf;
    double dE_poly, dE_field, u, exp_dE;
    int64_t n_bins;          /* unique voxels touched (0 for tangent rotation) */
} or_log;

typedef struct or_sim or_sim;

or_sim *or_new(const or_spec *spec, int32_t rng_mode, uint32_t seed, uint32_t stream,
               int32_t faithful);
void or_free(or_sim *s);
/* np.random.seed(inner_seed) equivalent (MT19937 init_genrand). */
void or_np_seed(or_sim *s, uint32_t seed);
/* mc_sim(): returns 0 on success, <0 on error.  logs may be NULL. */
int64_t or_mc_sim(or_sim *s, int64_t num_steps, or_ctrl *ctrls, int32_t n_ctrl,
                  double mu_adjust_factor, or_log *logs, int64_t max_logs, int64_t *n_logs);
/* One mc_step() for controller c (no controller update). */
int or_mc_step(or_sim *s, or_ctrl *c, or_log *log);
void or_update_amplitudes(or_ctrl *c);

/* state access */
void or_get_state(const or_sim *s, double *r, double *t3, double *t2, int64_t *states);
void or_set_positions(or_sim *s, const double *r, const double *t3, const double *t2);
void or_get_density(const or_sim *s, double *density);
int64_t or_n_bins(const or_sim *s);
/* update_all_densities_for_all_polymers() */
void or_density_init(or_sim *s);
/* Field dE for host-supplied trial positions/states (tape).  Fills the per-voxel
 * delta density (row = voxel, n_bins x (B+1)), the bins in reference set order and
 * returns dE.  The persistent density is not changed. */
double or_field_dE_tape(or_sim *s, const int64_t *inds, int64_t n, const double *r_trial,
                        const int64_t *states_trial, int32_t state_change,
                        double *delta_rho, int64_t *bins, int64_t *n_bins_out);
double or_elastic_dE_tape(or_sim *s, int32_t move, const int64_t *inds, int64_t n,
                          const double *r_trial, const double *t3_trial, const double *t2_trial);
/* compute_twist_angle_omega (polymers.pyx:3427-3461) */
double or_twist_omega(const double *t2_0, const double *t3_0, const double *t2_1, const double *t3_1);
double or_binding_dE_tape(or_sim *s, int64_t ind0, int64_t indf, const int64_t *states_trial);
/* Absolute energies: SSWLC.compute_E (SSTWLC.compute_E for a twisting chain) and
 * UniformDensityField.compute_E (as in the reference). */
double or_elastic_E(const or_sim *s);
double or_field_E(or_sim *s);

/* capped_exponential's truncated value for given draws and windows (glibc log10). */
void or_cexp_trunc(const int32_t *draws, const int64_t *windows, int64_t n, int64_t *out);
/* Python round(x, 2) as compute_E uses it (fields.pyx:2282). */
double or_py_round2(double x);
/* Sets mu_adjust_factor for tape evaluations (mc_sim sets it per call). */
void or_set_mu_adjust(or_sim *s, double f);
/* RNG primitives, exposed for known-answer tests. */
void or_glibc_stream(uint32_t seed, int32_t *out, int64_t n);
void or_mt_randint_stream(uint32_t seed, int64_t high, int64_t *out, int64_t n);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Transform / geometry known answers (linalg.pyx + move_funcs.transform_r_t3_t2). */
void or_axis_rotation(const double axis[3], const double point[3], double ang, double T[16]);
void or_translation(const double delta[3], double T[16]);
void or_transform(const double T[16], const double *r, const double *t3, const double *t2,
                  int64_t n, double *r_out, double *t3_out, double *t2_out);
/* Bead selection (bead_selection.pyx) driven by a glibc stream seeded with `seed`. */
int64_t or_select_stream(uint32_t seed, int32_t which, int64_t window, int64_t nbeads,
                         int64_t ind0, int64_t *out, int64_t n);
void or_check_bead_bounds(int64_t b0, int64_t b1, int64_t nb, int64_t *ind0, int64_t *indf);
/* CPython set iteration order for small non-negative ints inserted in order. */
int64_t or_pyset_order(const int64_t *keys, int64_t n, int64_t *out);

/* Batched CPU baseline: run replicas on `threads` threads, one replica per thread
 * at a time; returns total proposals evaluated.  seconds_out receives wall time. */
int64_t or_run_batch(or_sim **sims, int32_t n_sims, or_ctrl *ctrls /* n_sims x n_ctrl */,
                     int32_t n_ctrl, int64_t num_steps, int32_t threads, double *seconds_out);
#endif
