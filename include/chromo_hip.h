/*
 * chromo_hip.h -- C ABI of libchromo_hip.so, the MI355X (gfx950) Monte Carlo
 * inner loop of chromo.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on failure;
 *    chromo_last_error() copies the message of the last failure on this thread.
 *  - Pointers named d_* are device pointers (e.g. torch tensor data_ptr()); the
 *    library never frees caller memory.  `stream` is a hipStream_t (NULL = default).
 *  - Calls that share a replica table must be ordered by the caller (one stream).
 *
 * The reference has no C interface: its hot path is Cython (cdef / cpdef) reached
 * from Python.  Each entry point below names the reference function whose
 * behaviour it takes over; the Python layer in chromo_amd keeps the reference's
 * class and function names on top of these calls.
 */
#ifndef CHROMO_HIP_H
#define CHROMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHROMO_MAX_BINDERS 4   /* binder (reader protein) types per chain */
#define CHROMO_MAX_MARK 4      /* chemical_mods values 0..3 */
#define CHROMO_MAX_SITES 3     /* sites_per_bead 0..3 */
#define CHROMO_N_MOVES 5

enum chromo_move {
    CHROMO_CRANK_SHAFT = 0,
    CHROMO_END_PIVOT = 1,
    CHROMO_SLIDE = 2,
    CHROMO_TANGENT_ROTATION = 3,
    CHROMO_CHANGE_BINDING_STATE = 4
};

enum chromo_rng_mode {
    CHROMO_RNG_GLIBC = 0,   /* glibc rand() stream (srand(seed)) + numpy MT19937 */
    CHROMO_RNG_PHILOX = 1   /* Philox4x32-10, key = (seed, stream) */
};

/* RNG state block: CHROMO_RNG_WORDS uint32 words per replica (device). */
#define CHROMO_RNG_WORDS 704
#define CHROMO_RNG_GLIBC_STATE 0    /* 31 words */
#define CHROMO_RNG_GLIBC_F 31
#define CHROMO_RNG_GLIBC_R 32
#define CHROMO_RNG_MT_STATE 64      /* 624 words */
#define CHROMO_RNG_MT_INDEX 688
#define CHROMO_RNG_PHILOX_KEY 690   /* 2 words */
#define CHROMO_RNG_PHILOX_CTR 692   /* 2 words (64-bit block counter) */
#define CHROMO_RNG_PHILOX_BUF 694   /* 4 words */
#define CHROMO_RNG_PHILOX_IDX 698

/* One adaptable move and its SimpleControl / AcceptanceTracker state
 * (chromo/mc/moves.pyx:58-130, chromo/mc/mc_controller.py:97-213,
 * chromo/util/mc_stat.py:36-208).  Lives in device memory. */
typedef struct {
    int32_t move;           /* enum chromo_move */
    int32_t on;             /* MCAdapter.move_on */
    int32_t controlled;     /* 1 = SimpleControl, 0 = NoControl */
    int32_t num_per_cycle;
    double amp_move;
    int64_t amp_bead;
    double rate;            /* acceptance_tracker.acceptance_rate */
    double alpha;           /* 2 / (moves_in_average + 1) */
    double bead_lo, bead_hi, move_lo, move_hi;
    int64_t num_attempt, num_success;
} chromo_ctrl;

/* Optional per-proposal record (written only when chromo_replica.log != NULL). */
typedef struct {
    int32_t move;
    int32_t accepted;
    int64_t n, ind0, indf;
    double dE_poly, dE_field, u, exp_dE;
    int64_t n_bins;
} chromo_move_log;

/* Per-replica counters (int64, device). */
enum chromo_counter {
    CHROMO_CNT_PROPOSALS = 0,   /* proposals that reached dE (n_inds > 0) */
    CHROMO_CNT_ACCEPTED = 1,
    CHROMO_CNT_SUM_N = 2,       /* sum of n_inds */
    CHROMO_CNT_SUM_BINS = 3,    /* sum of unique voxels touched */
    CHROMO_CNT_ERROR = 4,       /* first error code (0 = none) */
    CHROMO_CNT_NEAR_TIES = 5,   /* |u - exp(-dE)| < 1e-12 */
    CHROMO_CNT_BYTES = 6,       /* algorithmic HBM bytes of all proposals */
    CHROMO_CNT_LOGGED = 7,      /* records written to log */
    CHROMO_CNT_TIE_VOXEL = 8,   /* bead coordinates within 1e-11 voxel of a voxel plane, where an
                                   ulp of difference between device and glibc sin/cos/acos/sqrt
                                   could change the voxel floor (every evaluation, speculative
                                   ones included) */
    CHROMO_CNT_TIE_HUGE = 9,    /* touched voxels over vf_limit both before and after a move: the
                                   reference's running sum of +-1e99 volume-fraction terms then
                                   drops or rounds away the terms summed before them in CPython
                                   set order, so its dE depends on that order (every evaluation,
                                   speculative ones included) */
    CHROMO_CNT_WORDS = 10
};

/* Values of counters[CHROMO_CNT_ERROR].  A replica whose counter is non-zero is
 * left unchanged by later launches. */
enum chromo_replica_error {
    CHROMO_ERR_WINDOW = 1,      /* bead window larger than the chain (the reference's ValueError) */
    CHROMO_ERR_SEGCAP = 2,      /* move longer than the replica's scratch arena (bead_cap / seg_cap) */
    CHROMO_ERR_TANGENT_K = 3,   /* more distinct tangent beads than the chain has */
    CHROMO_ERR_NBINDERS = 4,    /* binder count outside 1..4, or not in the launch's binder_mask */
    CHROMO_ERR_SPIN = 5,        /* a rejection or probe loop passed its bound (amplitude below one bead:
                                   the reference loops for ever) */
    CHROMO_ERR_ARENA = 6,       /* a scratch-arena index check failed */
    CHROMO_ERR_TWIST_WAVES = 7  /* a twisting chain (SSTWLC) launched with more than one wavefront
                                   per replica: it runs with one */
};

/* Everything the kernels need about one replica (one chain in one field).  The
 * host fills an array of these and copies it to the device; the kernels read it
 * there.  Layout: chain arrays are structure-of-arrays for coalesced access. */
typedef struct {
    double *pos;             /* 9 x N: rx ry rz t3x t3y t3z t2x t2y t2z */
    int8_t *states;          /* B x N binding states */
    const int8_t *marks;     /* B x N chemical modifications */
    const double *bonds;     /* 5 x (N-1): eps_bend eps_par eps_perp gamma eta */
    double *density;         /* n_bins x (B+1), row = voxel */
    uint32_t *rng;           /* CHROMO_RNG_WORDS */
    chromo_ctrl *ctrl;       /* n_ctrl entries, schedule order */
    chromo_move_log *log;    /* NULL = no log */
    int64_t *counters;       /* CHROMO_CNT_WORDS */
    void *scratch;           /* >= chromo_scratch_bytes(...) bytes, device */
    int64_t N;
    int64_t log_cap;
    int64_t seg_cap;         /* largest n_inds of a field move the scratch arena holds */
    int64_t bead_cap;        /* largest n_inds of any move the scratch arena holds */
    int32_t B;
    int32_t n_ctrl;
    int32_t rng_mode;
    int32_t field_on;
    int64_t nx, ny, nz;
    double width[3], dxyz[3], half_width[3], half_step[3];
    double vol_bin;
    double chi, bead_vol;
    double vf_limit;         /* the reference stores a C float; widened here */
    double pref[CHROMO_MAX_BINDERS];
    double intra[CHROMO_MAX_BINDERS];
    double cross[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS];
    double mu[CHROMO_MAX_BINDERS];
    double mu_adjust;
    int64_t max_binders;
    int32_t sites[CHROMO_MAX_BINDERS];
    int32_t n_mark_levels;
    int32_t confine_type;    /* 0 none, 1 spherical, 2 cubical */
    double confine_length;
    double F[CHROMO_MAX_BINDERS][CHROMO_MAX_MARK][CHROMO_MAX_SITES + 1];
    const double *access_vol;   /* n_bins accessible volumes (get_accessible_volumes,
                                   chromo/fields.pyx:714-770), or NULL: every voxel vol_bin */
    int32_t fast_field;         /* 1: move dE from sub-bin weights (init_fast_field and
                                   get_change_in_density_quickly, chromo/fields.pyx:577-671,
                                   1235-1368): no 1e-18 rules in the fold */
    int32_t pad_ff;
    int64_t n_sub[3];           /* sub-bins per axis: n_points * n (n_points even) */
    double sub_width[3];        /* voxel side / n_points */
    double half_points;         /* n_points / 2 */
    int32_t twist;              /* 1: SSTWLC (chromo/polymers.pyx:1889-2320): bonds has 7 rows, the
                                   last two eps_twist and the natural twist angle of each bond */
    int32_t pad_tw;
} chromo_replica;

/* ---- library ---------------------------------------------------------- */

/* Copies the last error message of this thread into buf (NUL-terminated). */
int chromo_last_error(char *buf, size_t len);
/* Library / ABI version and sizeof(chromo_replica) for layout checks. */
int chromo_abi_info(int64_t *version, int64_t *replica_bytes, int64_t *ctrl_bytes,
                    int64_t *log_bytes);
/* Bytes of device scratch one replica needs for moves of up to bead_cap beads,
 * of which field moves (all but tangent rotation) reach at most seg_cap beads.
 * Moves of at most 32 beads run in LDS and need no scratch. */
int64_t chromo_scratch_bytes(int64_t bead_cap, int64_t seg_cap, int32_t B);

/* Prepares each replica's scratch arena (must run once after allocation). */
int chromo_scratch_init(const chromo_replica *d_reps, int32_t n_reps, void *stream);

/* ---- RNG ---------------------------------------------------------------- */

/* Seeds the replica RNG blocks on the device.  GLIBC: srand(seeds[i]) for the
 * rand() stream (the reference never seeds it: == srand(1)); PHILOX: key =
 * (seeds[i], streams[i]).  Replaces the reference's implicit libc state. */
int chromo_rng_seed(const chromo_replica *d_reps, int32_t n_reps, const uint32_t *d_seeds,
                    const uint32_t *d_streams, int32_t mode, void *stream);
/* np.random.seed(seed) for the binding-state stream (chromo/mc/mc_sim.pyx:81). */
int chromo_np_seed(const chromo_replica *d_reps, int32_t n_reps, const uint32_t *d_seeds,
                   void *stream);

/* ---- field -------------------------------------------------------------- */

/* UniformDensityField.update_all_densities_for_all_polymers
 * (chromo/fields.pyx:2041-2106): deterministic, bitwise equal to the
 * reference's sequential bead-order scatter, including the 1e-18 clean-up. */
int chromo_density_init(const chromo_replica *d_reps, int32_t n_reps, void *stream);

/* ---- Monte Carlo -------------------------------------------------------- */

/* mc_sim (chromo/mc/mc_sim.pyx:26-103) for every replica in one persistent
 * launch: num_steps x (controllers x num_per_cycle) mc_step calls per replica,
 * then update_amplitudes per controller.  Chains advance independently. */
int chromo_mc_run(const chromo_replica *d_reps, int32_t n_reps, int64_t num_steps,
                  void *stream);
/* chromo_mc_run with an explicit launch shape.  waves_per_replica = 1 runs each
 * chain on one wavefront, one proposal at a time; 2 wavefronts evaluate two
 * proposals of a controller block at once and resolve them in order,
 * re-evaluating any proposal an earlier accepted one in the same round touched,
 * so the chains (decisions, states, densities, RNG streams) are the same for
 * every waves_per_replica.  binder_mask has bit B set for each binder count B
 * (1..4) among the replicas, and bit 8 + B for each B with twisting chains
 * (SSTWLC, one wavefront only); a replica whose B is not in it fails with the
 * binder-count error.  No reference counterpart: it is a launch-shape variant
 * of mc_sim (chromo/mc/mc_sim.pyx:26-103). */
int chromo_mc_run_tuned(const chromo_replica *d_reps, int32_t n_reps, int64_t num_steps,
                        int32_t waves_per_replica, int32_t binder_mask, void *stream);
/* The same schedule for one replica with one kernel launch per mc_step and one
 * per controller update (the "plain" path). */
int chromo_mc_run_plain(const chromo_replica *d_reps, int32_t rep, int64_t num_steps,
                        void *stream);
/* mc_step (chromo/mc/mc_sim.pyx:106-182) once, for controller index ctrl. */
int chromo_mc_step(const chromo_replica *d_reps, int32_t rep, int32_t ctrl, void *stream);
/* SimpleControl.update_amplitudes (chromo/mc/mc_controller.py:126-213). */
int chromo_update_amplitudes(const chromo_replica *d_reps, int32_t rep, int32_t ctrl,
                             void *stream);

/* ---- checkpoint / restore ----------------------------------------------- */

/* Writes every replica's state to `path`: positions, binder states, voxel
 * densities, the RNG block (glibc / MT19937 / Philox position), controllers and
 * counters.  Continuing from chromo_restore gives the same chains bit for bit as
 * running on.  The reference resumes from its CSV snapshots with fresh streams
 * (continue_polymer_in_field_simulation, chromo/mc/__init__.py:167-227); this
 * entry has no single counterpart and adds the exact continuation. */
int chromo_checkpoint(const chromo_replica *d_reps, int32_t n_reps, const char *path, void *stream);
/* Reads a chromo_checkpoint file into a table describing the same chains,
 * fields, controllers and RNG modes (checked; nothing is written on a
 * mismatch). */
int chromo_restore(const chromo_replica *d_reps, int32_t n_reps, const char *path, void *stream);

/* ---- tape entry points (parity testing) -------------------------------- */

/* UniformDensityField.compute_dE (chromo/fields.pyx:1149-1233) for one move
 * whose trial positions / states are given: d_inds[n] bead indices, d_r_trial
 * (n x 3), d_states_trial (n x B).  Writes dE, the unique voxels d_bins[*n_bins]
 * (first-occurrence order) and their delta densities d_delta (n_bins x (B+1)).
 * The persistent density is not modified. */
int chromo_field_dE_tape(const chromo_replica *d_reps, int32_t rep, const int64_t *d_inds,
                         int64_t n, const double *d_r_trial, const int8_t *d_states_trial,
                         int32_t state_change, double *d_dE, int64_t *d_bins,
                         double *d_delta, int64_t *d_n_bins, void *stream);

/* (int)(-log10(x / RAND_MAX + 1e-5) * window * 0.45 + 1.0001) for n draws x and
 * windows, with the device arithmetic every bead-window draw uses
 * (capped_exponential, chromo/util/bead_selection.pyx:58-65).  Parity testing:
 * the draws next to each integer crossing are compared with glibc's log10. */
int chromo_capped_exponential_tape(const int32_t *d_draws, const int64_t *d_windows, int64_t n,
                                   int64_t *d_out, void *stream);

/* ---- absolute energies and adjoint -------------------------------------- */

/* One chain in one field, for the energy entry points.  Device pointers:
 * t3 (N x 3), states / marks (B x N), bonds (5 x (N-1); 7 x (N-1) when twist). */
typedef struct {
    int64_t N;
    int32_t B;
    int32_t twist;      /* 1: SSTWLC bonds; chromo_energy_reference adds the twist term of
                           SSTWLC.compute_E (polymers.pyx:2250-2286), the adjoint entry
                           points reject it */
    const double *t3;
    const int8_t *states;
    const int8_t *marks;
    const double *bonds;
    int64_t nx, ny, nz;
    double x_width, y_width, z_width;
    double bead_vol;
    double vf_limit;
    double mu_adjust;
    double v_int[CHROMO_MAX_BINDERS];                        /* 4/3 pi r_int^3 */
    int32_t cross_mask[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS];  /* [i*4+j]: i lists j */
    double F[CHROMO_MAX_BINDERS][CHROMO_MAX_MARK][CHROMO_MAX_SITES + 1];
} chromo_energy_model;

/* New (no reference counterpart): the energy whose differences are the move
 * dE of SSWLC.compute_dE + UniformDensityField.compute_dE + binding_dE, and its
 * gradient.  d_r: bead positions (N x 3).  theta = (chi, eps_ij for i <= j in
 * row order, mu_j): 1 + B(B+1)/2 + B values.  d_E[4] = total, elastic, field,
 * binding.  d_grad_r (N x 3) and d_grad_theta may be NULL.  Tangents fixed. */
int chromo_energy_adjoint(const chromo_energy_model *model, const double *d_r,
                          const double *d_theta, double *d_E, double *d_grad_r,
                          double *d_grad_theta, void *stream);

/* No counterpart in the reference.  chromo_energy_adjoint for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments, same results bit for bit.
 * Its scratch (the density scatter's buckets, the voxel densities and their
 * adjoint, the block partials) lives in one file-scope device array of 32 MiB
 * per GPU, held until the process ends; it calls no hipMalloc / hipFree.  An
 * evaluation that needs more is rejected (-1, the bytes it needs in the message):
 * 72 B per bead plus 12 + 8 (B+1) B per voxel (12 + 16 (B+1) with gradients),
 * e.g. 20,000 beads and 48^3 voxels with B = 3 take 10 MiB.  Every call on one GPU shares that
 * array and it keeps the last call's contents: calls that overlap on one GPU
 * (two streams, two host threads) race. */
int chromo_energy_adjoint_noalloc(const chromo_energy_model *model, const double *d_r,
                                  const double *d_theta, double *d_E, double *d_grad_r,
                                  double *d_grad_theta, void *stream);

/* SSWLC.compute_E (chromo/polymers.pyx:1348-1381) and
 * UniformDensityField.compute_E (chromo/fields.pyx:1939-2286): d_out[0] =
 * elastic, d_out[1] = field (chi phi (1 - phi), round(phi, 2) > vf_limit, no
 * cross-talk, as the reference).  d_pos holds r, t3, t2 (soa = 1: 9 x N rows;
 * soa = 0: N x 9).  d_pref / d_intra: B field prefactors.  The recomputed
 * densities (no 1e-18 clean-up, as update_all_densities) are written to
 * d_density_out when it is not NULL. */
int chromo_energy_reference(const chromo_energy_model *model, const double *d_pos, int32_t soa,
                            const double *d_pref, const double *d_intra, double chi,
                            double *d_density_out, double *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
