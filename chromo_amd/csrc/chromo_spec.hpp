// chromo_spec.hpp -- device code of the speculative persistent kernel.
#pragma once

#include "chromo_mc.hpp"

namespace {

// ===========================================================================
// Speculative rounds: W proposals of one controller block evaluated at once.
//
// Inside a controller block the amplitudes are fixed, so the random draws of each
// proposal -- the proposal itself and its Metropolis draw -- do not depend on the
// accept/reject outcomes before it.  (The exception, a crank-shaft axis shorter
// than 1e-5 which draws a random axis, is detected and redone one at a time.)
// Wavefront 0 generates up to W proposals; wavefront w evaluates proposal w
// against the current state, in registers.  Resolution walks the proposals in
// order: proposal s keeps its evaluated dE unless an accepted proposal s' < s of
// the same round wrote a bead or a voxel that s reads; then s and all later
// proposals go to the next round and are evaluated again.  Each chain therefore
// is the reference's sequential chain; only the number of rounds depends on
// conflicts.  Moves of more than SPEC_NMAX beads run one at a time through
// mc_step_dev.
// ===========================================================================

constexpr int SPEC_NMAX = 8;          // beads (tangent picks) one wavefront evaluates
constexpr int SPEC_WMAX = 8;          // wavefronts per replica
constexpr int SPEC_HASH = 1024;       // conflict-table entries
constexpr int SPEC_HASH_SHIFT = 22;   // 32 - log2(SPEC_HASH)
constexpr int SPEC_HASH_SOFT = 640;   // insertions by wavefronts 1.. (wavefront 0 adds <= 128)

enum : int { SLOT_SMALL = 0, SLOT_EMPTY = 1, SLOT_SEQ = 2 };
constexpr int CK_WORDS = 40;

struct SpecSlot {
    int32_t kind, move, binder, lhs;
    int32_t ud, s1, s2, pad0;
    int64_t n, ind0, indf;
    double ang;                               // angle; slide: amplitude
    int64_t picks[SPEC_NMAX];                 // tangent rotation beads
    int32_t ps1[SPEC_NMAX], ps2[SPEC_NMAX];   // their unit-sphere draws
    int32_t newst[SPEC_NMAX];                 // binding: new states
    uint32_t ck[CK_WORDS];                    // RNG state before the proposal
};

struct SpecRes {
    double dEp, dEf, u, expdE;
    int32_t accepted, nbins, resample, overflow;
    int64_t rlo, rhi, wlo, whi;               // beads read [rlo, rhi], written [wlo, whi)
    uint32_t cmask;                           // earlier slots whose writes this one reads
    int32_t pad1;
};

struct SpecCtl {
    int32_t nq, cnt, remaining, err;
    int32_t used;                             // conflict-table insertions this round
    int32_t pad[3];
};

// LDS exchange area of the one-at-a-time path (seq_step).
struct SeqIO {
    uint32_t ck[CK_WORDS];
    int32_t mti, pad0;
    chromo_ctrl C;
    MoveOut out;
};

struct SpecLayout {
    size_t mt, slots, res, hkey, hmask, ctl, seq, total;
};

// LDS: the move arena of the one-at-a-time path, the MT19937 block, two slot
// buffers, per-wavefront results, the conflict table and the control words.
__host__ __device__ inline SpecLayout spec_layout(int Bp1, int W)
{
    SpecLayout S;
    ArenaLayout L = lds_layout(Bp1);
    S.mt = L.total;
    size_t o = align_up(L.total + MT_LDS_BYTES, 16);
    S.slots = o;
    o += sizeof(SpecSlot) * 2 * (size_t)W;
    S.res = o;
    o += sizeof(SpecRes) * (size_t)W;
    S.hkey = o;
    o += sizeof(int32_t) * SPEC_HASH;
    S.hmask = o;
    o += sizeof(uint32_t) * SPEC_HASH;
    S.ctl = o;
    o += sizeof(SpecCtl);
    o = align_up(o, 16);
    S.seq = o;
    o += sizeof(SeqIO);
    S.total = align_up(o, 256);
    return S;
}

struct SpecLds {
    SpecSlot *slots[2];
    SpecRes *res;
    int32_t *hkey;
    uint32_t *hmask;
    SpecCtl *ctl;
    SeqIO *seq;
    uint32_t *mt;
};

// Registers a wavefront keeps from its evaluation to its commit.
struct SpecKeep {
    double w[9];
    int64_t b;
    double nr[2][CHROMO_MAX_BINDERS + 1];
    int32_t vox[2], hs[2];
    int own[2], ins[2];
};

__device__ __forceinline__ void rng_save(const WaveRng &g, uint32_t *ck)
{
    const int lane = lane_id();
    if (lane < 31) ck[lane] = g.g;
    if (lane == 0) {
        ck[31] = (uint32_t)g.gf;
        ck[32] = (uint32_t)g.gr;
        ck[33] = (uint32_t)g.ctr;
        ck[34] = (uint32_t)(g.ctr >> 32);
        ck[35] = g.b0;
        ck[36] = g.b1;
        ck[37] = g.b2;
        ck[38] = g.b3;
        ck[39] = (uint32_t)g.bi;
    }
    wave_sync();
}

// Restores the glibc and Philox streams (the MT19937 stream is never rewound:
// no proposal that is redone consumed MT19937 draws).
__device__ __forceinline__ void rng_restore(WaveRng &g, const uint32_t *ck)
{
    wave_sync();
    const int lane = lane_id();
    g.g = lane < 31 ? ck[lane] : 0u;
    g.gf = (int)ck[31];
    g.gr = (int)ck[32];
    g.ctr = (uint64_t)ck[33] | ((uint64_t)ck[34] << 32);
    g.b0 = ck[35];
    g.b1 = ck[36];
    g.b2 = ck[37];
    g.b3 = ck[38];
    g.bi = (int)ck[39];
}

// Draws one proposal into slot S (wavefront 0), in the draw order of
// mc_step_dev.  A proposal that is too large for one wavefront, or that fails,
// is marked SLOT_SEQ with its draws given back: mc_step_dev redoes it.
template <int NB>
__device__ __forceinline__ int gen_slot(const chromo_replica *R, WaveRng &rng, const chromo_ctrl &C, SpecSlot &S)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    rng_save(rng, S.ck);
    const double amp_move = C.amp_move;
    const int64_t amp_bead = C.amp_bead;
    const int mv = C.move;
    int err = 0;
    bool seq = false;
    int64_t n = 0, ind0 = 0, indf = 0;
    double ang = 0;
    int32_t s1 = 0, s2 = 0, lhs = 0, binder = 0;
    if (mv == CHROMO_CRANK_SHAFT) {
        ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t b0 = (int64_t)(int32_t)(((double)rnd31(rng) / RMAX) * (double)N);
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (!err) {
            b1 = imax64(b1, 1);
            check_bead_bounds(b0, b1, N, ind0, indf);
            n = indf - ind0;
            // an empty window has coinciding axis beads (ind0 == indf == N): the
            // reference draws a random axis before it returns
            if (n == 0) {
                rnd31(rng);
                rnd31(rng);
            }
        }
        seq = err || n > SPEC_NMAX;
    } else if (mv == CHROMO_END_PIVOT) {
        ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        lhs = (int32_t)(rnd31(rng) % 2);
        if (lhs == 1) {
            ind0 = 0;
            indf = from_left(rng, amp_bead, N, err) + 1;
        } else {
            ind0 = from_right(rng, amp_bead, N, err);
            indf = N;
        }
        if (!err && indf > N) err = ERR_WINDOW;
        n = indf - ind0;
        seq = err || n > SPEC_NMAX;
        if (!seq) {
            s1 = rnd31(rng);
            s2 = rnd31(rng);
        }
    } else if (mv == CHROMO_SLIDE) {
        ang = amp_move * ((double)rnd31(rng) / RMAX);
        s1 = rnd31(rng);
        s2 = rnd31(rng);
        int64_t b0 = rnd31(rng) % N;
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (!err) {
            check_bead_bounds(b0, b1, N, ind0, indf);
            n = indf - ind0;
        }
        seq = err || n > SPEC_NMAX;
    } else if (mv == CHROMO_TANGENT_ROTATION) {
        ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t k = rnd31(rng) % amp_bead + 1;
        n = k;
        seq = (k > N) || (k > SPEC_NMAX);
        if (!seq) {
            int64_t mine = 0;
            for (int64_t i = 0; i < k && !seq; ++i) {
                for (int it = 0;; ++it) {
                    // a pick that does not end is left to mc_step_dev, which reports it
                    if (it >= SPIN_LIMIT) {
                        seq = true;
                        break;
                    }
                    int64_t c = rnd31(rng) % N;
                    if (!__any(lane < i && mine == c)) {
                        if (lane == i) mine = c;
                        break;
                    }
                }
            }
            if (lane < k) S.picks[lane] = mine;
            for (int64_t i = 0; i < k && !seq; ++i) {
                int32_t a = rnd31(rng);
                int32_t b = rnd31(rng);
                if (lane == 0) {
                    S.ps1[i] = a;
                    S.ps2[i] = b;
                }
            }
        }
    } else {
        binder = (int32_t)(rnd31(rng) % NB);
        const int64_t ntails = R->sites[binder];
        int64_t b0 = rnd31(rng) % N;
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (!err) {
            check_bead_bounds(b0, b1, N, ind0, indf);
            n = indf - ind0;
        }
        seq = err || n > SPEC_NMAX;
        if (!seq) {
            for (int64_t i = 0; i < n; ++i) {
                int st = (int)rnd_randint(rng, ntails + 1);
                if (lane == 0) S.newst[i] = st;
            }
        }
    }
    int kind;
    int32_t ud = 0;
    if (seq) {
        rng_restore(rng, S.ck);
        kind = SLOT_SEQ;
    } else if (n == 0) {
        kind = SLOT_EMPTY;
    } else {
        ud = rnd31(rng);
        kind = SLOT_SMALL;
    }
    if (lane == 0) {
        S.kind = kind;
        S.move = mv;
        S.binder = binder;
        S.lhs = lhs;
        S.ud = ud;
        S.s1 = s1;
        S.s2 = s2;
        S.n = n;
        S.ind0 = ind0;
        S.indf = indf;
        S.ang = ang;
    }
    wave_sync();
    return kind;
}

// mc_step_dev for one proposal, run by one wavefront; the RNG, the controller and
// the result pass through LDS (io).  Inlined: a call boundary here makes the
// callee save most of the register file on every call and the caller spill
// around it.
template <int NB, int W>
__device__ __forceinline__ void seq_step(const chromo_replica *R, SeqIO *io, uint32_t *mt)
{
    WaveRng rng;
    rng.mode = R->rng_mode;
    rng_restore(rng, io->ck);
    rng.mt = mt;
    rng.mti = io->mti;
    rng.k0 = R->rng[CHROMO_RNG_PHILOX_KEY];
    rng.k1 = R->rng[CHROMO_RNG_PHILOX_KEY + 1];
    chromo_ctrl C = io->C;
    MoveOut o;
    mc_step_dev<NB>(R, rng, C, o);
    rng_save(rng, io->ck);
    if (lane_id() == 0) {
        io->mti = rng.mti;
        io->C = C;
        io->out = o;
    }
    wave_sync();
}

// One (position, corner) term of the per-voxel fold (fields.pyx:1491-1520).
template <int NB>
__device__ __forceinline__ void fold_term(int l, const double *wq, double pre, int32_t sq, double vol_bin,
                                          double *acc, int &first, int ebase)
{
    const double a = (l & 1) ? (1 - wq[0]) : wq[0];
    const double b = (l & 2) ? (1 - wq[1]) : wq[1];
    const double c = (l & 4) ? (1 - wq[2]) : wq[2];
    const double wt = a * b * c;
    const double c0 = wt / vol_bin;
    const double t0 = pre * c0;
    if (fabs(t0) > 1E-18) acc[0] += t0;
#pragma unroll
    for (int m = 1; m <= NB; ++m) {
        const int st = (sq >> (8 * (m - 1))) & 0xff;
        const double t = pre * (c0 * (double)st);
        if (fabs(t) > 1E-18) acc[m] += t;
    }
    if (first < 0) first = ebase + l;
}

__device__ __forceinline__ bool axis_hit(int c, int d, int64_t nn)
{
    return nn == 1 ? true : (c == 0 ? d == 0 : d == 1);
}

__device__ __forceinline__ int64_t pack_ix(const int64_t ix[3])
{
    return ix[0] | (ix[1] << 21) | (ix[2] << 42);
}

// Evaluates one proposal (n <= SPEC_NMAX beads) with one wavefront and no LDS
// arena: dE, the decision, the footprint, and the trial state kept in registers.
// Lane i owns bead i; lane l owns contributions e = l and e = l + 64 (bead e/16,
// configuration (e/8)%2, corner e%8) and folds, for the voxel of each, every
// contribution in contribution order.  The per-voxel delta densities are the
// reference's left folds bit for bit, and lanes sum their voxels' terms in the
// same order as move_body, so dE equals the one-at-a-time path's.
template <int NB>
__device__ __forceinline__ void eval_small(const chromo_replica *R, const SpecSlot &S, SpecRes &res, SpecKeep &K,
                           const SpecLds &L, int slot)
{
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int64_t N = R->N;
    const double *Pp = R->pos;
    const int mv = S.move;
    const int64_t n = S.n, ind0 = S.ind0, indf = S.indf;
    const bool tangent = mv == CHROMO_TANGENT_ROTATION;
    const bool state_change = mv == CHROMO_CHANGE_BINDING_STATE;
    const bool check_field = R->field_on && !tangent;
    const bool act = lane < n;
    K.own[0] = K.own[1] = 0;
    K.ins[0] = K.ins[1] = 0;
    if (lane == 0) {
        res.cmask = 0;
        res.resample = 0;
        res.overflow = 0;
        res.nbins = 0;
        res.rlo = ind0 - 1;
        res.rhi = indf;
        res.wlo = ind0;
        res.whi = indf;
    }
    int64_t b = ind0 + lane;
    if (tangent) b = act ? S.picks[lane] : 0;
    if (!act) b = tangent ? 0 : ind0;
    K.b = b;

    double o[9];
    int8_t so[NB], sn[NB];
#pragma unroll
    for (int c = 0; c < 9; ++c) o[c] = 0;
#pragma unroll
    for (int m = 0; m < NB; ++m) so[m] = sn[m] = 0;
    if (act) {
#pragma unroll
        for (int c = 0; c < 9; ++c) o[c] = Pp[c * N + b];
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            so[m] = R->states[(int64_t)m * N + b];
            sn[m] = so[m];
        }
        if (state_change) {
#pragma unroll
            for (int m = 0; m < NB; ++m)
                if (m == S.binder) sn[m] = (int8_t)S.newst[lane];
        }
    }

    // transformation (move_funcs.pyx: crank 157-280, end pivot 349-398, slide 440-444)
    double T[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) T[q] = 0;
    if (mv == CHROMO_CRANK_SHAFT) {
        int64_t ia, ib;
        const int64_t indf_m1 = indf - 1, ind0_m1 = ind0 - 1;
        if (ind0 == indf_m1 && ind0 == 0) { ia = indf; ib = ind0; }
        else if (ind0 == indf_m1 && ind0 == N - 1) { ia = ind0; ib = ind0_m1; }
        else if (ind0 == 0 && indf == N) { ia = indf_m1; ib = ind0; }
        else if (ind0 == 0) { ia = indf; ib = ind0; }
        else if (indf == N) { ia = indf_m1; ib = ind0_m1; }
        else { ia = indf; ib = ind0_m1; }
        int64_t ifc;
        if (ind0 == 0 && indf != N) ifc = indf;
        else if (ind0 != 0 && indf == N) ifc = ind0_m1;
        else if (ind0 == 0 && indf == N) ifc = ind0;
        else ifc = ind0_m1;
        double ra[3], rb[3], pt[3], dir[3];
        load_pos3(Pp, N, ia, ra);
        load_pos3(Pp, N, ib, rb);
        load_pos3(Pp, N, ifc, pt);
        for (int i = 0; i < 3; ++i) dir[i] = ra[i] - rb[i];
        double mag = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        if (mag < 1E-5) {
            if (lane == 0) res.resample = 1;
            return;
        }
        double sc = 1.0 / mag;
        for (int i = 0; i < 3; ++i) dir[i] = dir[i] * sc;
        axis_rotation(dir, pt, S.ang, T);
    } else if (mv == CHROMO_END_PIVOT) {
        int64_t ifc;
        if (ind0 == 0 && indf != N) ifc = indf;
        else if (ind0 != 0 && indf == N) ifc = ind0 - 1;
        else if (ind0 == 0 && indf == N && S.lhs == 1) ifc = indf - 1;
        else ifc = ind0;
        double pt[3], axis[3];
        load_pos3(Pp, N, ifc, pt);
        sphere_from(S.s1, S.s2, axis);
        axis_rotation(axis, pt, S.ang, T);
    } else if (mv == CHROMO_SLIDE) {
        double dir[3];
        sphere_from(S.s1, S.s2, dir);
        for (int i = 0; i < 3; ++i) dir[i] *= S.ang;
        T[3] = dir[0];
        T[7] = dir[1];
        T[11] = dir[2];
    } else if (tangent && act) {
        double dir[3];
        const double origin[3] = {0., 0., 0.};
        sphere_from(S.ps1[lane], S.ps2[lane], dir);
        axis_rotation(dir, origin, S.ang, T);
    }
    double w[9];
    if (mv == CHROMO_CRANK_SHAFT || mv == CHROMO_END_PIVOT) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            w[j] = rowdot(&T[4 * j], o[0], o[1], o[2]) + T[4 * j + 3];
            w[3 + j] = rowdot(&T[4 * j], o[3], o[4], o[5]);
            w[6 + j] = rowdot(&T[4 * j], o[6], o[7], o[8]);
        }
    } else if (mv == CHROMO_SLIDE) {
#pragma unroll
        for (int j = 0; j < 3; ++j) w[j] = o[j] + T[4 * j + 3];
#pragma unroll
        for (int c = 3; c < 9; ++c) w[c] = o[c];
    } else if (tangent) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            w[j] = o[j];
            w[3 + j] = rowdot(&T[4 * j], o[3], o[4], o[5]);
            w[6 + j] = rowdot(&T[4 * j], o[6], o[7], o[8]);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) w[c] = o[c];
    }

    // polymer dE (SSWLC.compute_dE, polymers.pyx:1024-1082)
    double dEp = 0;
    if (tangent || state_change) {
        double e = 0;
        if (act && tangent) {
            if (b != 0) {
                double r0[3], t30[3];
                for (int j = 0; j < 3; ++j) {
                    r0[j] = Pp[j * N + b - 1];
                    t30[j] = Pp[(3 + j) * N + b - 1];
                }
                Bond bd = load_bond(R->bonds, N, b - 1);
                e += pair_dE_forward(bd, r0, &o[0], &w[0], t30, &o[3], &w[3]);
            }
            if (b + 1 != N) {
                double r1[3], t31[3];
                for (int j = 0; j < 3; ++j) {
                    r1[j] = Pp[j * N + b + 1];
                    t31[j] = Pp[(3 + j) * N + b + 1];
                }
                Bond bd = load_bond(R->bonds, N, b);
                e += pair_dE_reverse(bd, &o[0], &w[0], r1, &o[3], &w[3], t31);
            }
        } else if (act) {
            int8_t mods[NB];
#pragma unroll
            for (int m = 0; m < NB; ++m) mods[m] = R->marks[(int64_t)m * N + b];
            e = bead_binding_dE<NB>(R, mods, so, sn);
        }
        for (int i = 0; i < n; ++i) dEp += readlane_f64(e, i);
    } else {
        double ef = 0, er = 0;
        if (lane == 0 && ind0 != 0) {
            double r0[3], t30[3];
            for (int j = 0; j < 3; ++j) {
                r0[j] = Pp[j * N + ind0 - 1];
                t30[j] = Pp[(3 + j) * N + ind0 - 1];
            }
            Bond bd = load_bond(R->bonds, N, ind0 - 1);
            ef = pair_dE_forward(bd, r0, &o[0], &w[0], t30, &o[3], &w[3]);
        }
        if (lane == n - 1 && indf != N) {
            double r1[3], t31[3];
            for (int j = 0; j < 3; ++j) {
                r1[j] = Pp[j * N + indf];
                t31[j] = Pp[(3 + j) * N + indf];
            }
            Bond bd = load_bond(R->bonds, N, indf - 1);
            er = pair_dE_reverse(bd, &o[0], &w[0], r1, &o[3], &w[3], t31);
        }
        if (ind0 != 0) dEp += readlane_f64(ef, 0);
        if (indf != N) dEp += readlane_f64(er, (int)(n - 1));
    }

    // field dE (UniformDensityField.compute_dE, fields.pyx:1149-1233)
    double dEf = 0;
    int U = 0;
    if (check_field) {
        int64_t P0 = 0, P1 = 0;
        double wt0[3] = {0, 0, 0}, wt1[3] = {0, 0, 0};
        int32_t st0 = 0, st1 = 0;
        if (act) {
            int64_t ix[3];
            voxel_lower(R, &o[0], ix, wt0);
            P0 = pack_ix(ix);
            double pt[3];
            for (int j = 0; j < 3; ++j) pt[j] = state_change ? o[j] : w[j];
            voxel_lower(R, pt, ix, wt1);
            P1 = pack_ix(ix);
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                st0 |= ((int32_t)(uint8_t)so[m]) << (8 * m);
                st1 |= ((int32_t)(uint8_t)(state_change ? sn[m] : so[m])) << (8 * m);
            }
        }
        const int64_t nx = R->nx, ny = R->ny, nz = R->nz;
        const int E = (int)(16 * n);
        int X[2], Y[2], Z[2], first[2], valid[2];
        int32_t vox[2];
        double rho[2][Bp1], acc[2][Bp1];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = lane + WAVE * j;
            valid[j] = e < E;
            const int p = e >> 3, l = e & 7;
            const int i = (p >> 1) & (WAVE - 1), k = p & 1;
            const int64_t q0 = __shfl(P0, i, WAVE);
            const int64_t q1 = __shfl(P1, i, WAVE);
            const int64_t q = k ? q1 : q0;
            int64_t x = (q & 0x1fffff) + (l & 1), y = ((q >> 21) & 0x1fffff) + ((l >> 1) & 1),
                    z = ((q >> 42) & 0x1fffff) + ((l >> 2) & 1);
            if (x >= nx) x -= nx;
            if (y >= ny) y -= ny;
            if (z >= nz) z -= nz;
            X[j] = (int)x;
            Y[j] = (int)y;
            Z[j] = (int)z;
            vox[j] = (int32_t)(x + y * nx + z * nx * ny);
            first[j] = -1;
#pragma unroll
            for (int m = 0; m < Bp1; ++m) {
                acc[j][m] = 0;
                rho[j][m] = valid[j] ? R->density[(int64_t)vox[j] * Bp1 + m] : 0.0;
            }
        }
        const bool degen = nx == 1 || ny == 1 || nz == 1;
        const double vol_bin = R->vol_bin;
        for (int pq = 0; pq < 2 * n; ++pq) {
            const int iq = pq >> 1, kq = pq & 1;
            const int64_t qp = readlane_i64(kq ? P1 : P0, iq);
            double wq[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) wq[c] = readlane_f64(kq ? wt1[c] : wt0[c], iq);
            const int32_t sq = (int32_t)readlane_u32((uint32_t)(kq ? st1 : st0), iq);
            const double pre = kq ? 1. : -1.;
            const int qx = (int)(qp & 0x1fffff), qy = (int)((qp >> 21) & 0x1fffff),
                      qz = (int)((qp >> 42) & 0x1fffff);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (!valid[j]) continue;
                int dx = X[j] - qx, dy = Y[j] - qy, dz = Z[j] - qz;
                if (dx < 0) dx += (int)nx;
                if (dy < 0) dy += (int)ny;
                if (dz < 0) dz += (int)nz;
                if (!degen) {
                    if (dx <= 1 && dy <= 1 && dz <= 1)
                        fold_term<NB>(dx | (dy << 1) | (dz << 2), wq, pre, sq, vol_bin, acc[j], first[j],
                                      8 * pq);
                } else {
                    for (int l = 0; l < 8; ++l)
                        if (axis_hit(l & 1, dx, nx) && axis_hit((l >> 1) & 1, dy, ny) &&
                            axis_hit((l >> 2) & 1, dz, nz))
                            fold_term<NB>(l, wq, pre, sq, vol_bin, acc[j], first[j], 8 * pq);
                }
            }
        }
        double S2[NB], SX[NB * NB], NS = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) S2[j] = 0;
#pragma unroll
        for (int j = 0; j < NB * NB; ++j) SX[j] = 0;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            K.own[j] = valid[j] && first[j] == lane + WAVE * j;
            if (K.own[j]) voxel_dE_terms<NB>(R, rho[j], acc[j], S2, SX, NS);
            K.vox[j] = vox[j];
#pragma unroll
            for (int m = 0; m < Bp1; ++m) K.nr[j][m] = rho[j][m] + acc[j][m];
        }
        const double share = wave_sum(lane_field_share<NB>(R, S2, SX, NS));
        U = __popcll(__ballot(K.own[0])) + __popcll(__ballot(K.own[1]));
        // count_doubly_bound (fields.pyx:1877-1937) and confinement (:121-202)
        int dbl_t[NB], dbl[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            dbl_t[j] = __popcll(__ballot(act && (state_change ? sn[j] : so[j]) == 2));
            dbl[j] = __popcll(__ballot(act && so[j] == 2));
        }
        int out_t = 0, out_c = 0;
        const int ct = R->confine_type;
        if (ct != 0 && !state_change) {
            const double cl = R->confine_length;
            if (ct == 1) {
                out_t = __popcll(__ballot(act && sqrt(dot3(&w[0], &w[0])) > cl));
                out_c = __popcll(__ballot(act && sqrt(dot3(&o[0], &o[0])) > cl));
            } else {
#pragma unroll
                for (int j = 0; j < 3; ++j) out_t += __popcll(__ballot(act && fabs(w[j]) > cl / 2));
            }
        }
        dEf = field_dE_total<NB>(R, share, dbl_t, dbl, out_t, out_c, state_change);

        // publish the voxels to the conflict table
        int base = 0;
        if (lane == 0) base = atomicAdd(&L.ctl->used, U);
        base = bcast0(base);
        if (slot > 0 && base + U > SPEC_HASH_SOFT) {
            if (lane == 0) res.overflow = 1;
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (!K.own[j]) continue;
                uint32_t h = ((uint32_t)vox[j] * 2654435761u) >> SPEC_HASH_SHIFT;
                int probes = 0;
                bool full = false;
                for (;;) {
                    int old = atomicCAS(&L.hkey[h], HKEY_EMPTY, vox[j]);
                    if (old == HKEY_EMPTY || old == vox[j]) break;
                    h = (h + 1) & (uint32_t)(SPEC_HASH - 1);
                    if (++probes >= SPEC_HASH) {
                        full = true;
                        break;
                    }
                }
                if (full) {
                    res.overflow = 1;
                    continue;
                }
                atomicOr(&L.hmask[h], 1u << slot);
                K.hs[j] = (int32_t)h;
                K.ins[j] = 1;
            }
        }
    }

    // Metropolis (mc_sim.pyx:152-182)
    double dE = 0;
    dE += dEp;
    if (check_field) dE += dEf;
    const double expdE = exp(-dE);
    const double u = (double)S.ud / RMAX;
    const int accepted = u < expdE;
#pragma unroll
    for (int c = 0; c < 9; ++c) K.w[c] = w[c];
    if (lane == 0) {
        res.dEp = dEp;
        res.dEf = dEf;
        res.u = u;
        res.expdE = expdE;
        res.accepted = accepted;
        res.nbins = U;
    }
}

// Earlier slots of this round whose writes slot s reads.
__device__ __forceinline__ void spec_conflicts(const SpecSlot *slots, SpecRes *res, const SpecKeep &K, const SpecLds &L,
                               int s)
{
    const int lane = lane_id();
    const uint32_t low = (1u << s) - 1u;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (K.ins[j]) m |= L.hmask[K.hs[j]] & low;
    const SpecSlot &S = slots[s];
    if (S.move == CHROMO_TANGENT_ROTATION) {
        const int a = lane >> 3, c = lane & 7;
        if (a < s && slots[a].kind == SLOT_SMALL && c < S.n) {
            const int64_t mine = S.picks[c];
            for (int64_t q = 0; q < slots[a].n; ++q) {
                const int64_t d = slots[a].picks[q] - mine;
                if (d >= -1 && d <= 1) m |= 1u << a;
            }
        }
    } else if (lane < s && slots[lane].kind == SLOT_SMALL) {
        const SpecRes &o = res[lane];
        if (o.wlo <= res[s].rhi && res[s].rlo < o.whi) m |= 1u << lane;
    }
    m = wave_or(m);
    if (lane == 0) res[s].cmask = m;
}

template <int NB>
__device__ __forceinline__ void spec_commit(const chromo_replica *R, const SpecSlot &S, const SpecKeep &K)
{
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int64_t N = R->N;
    if (lane < S.n) {
        if (S.move == CHROMO_CHANGE_BINDING_STATE) {
            R->states[(int64_t)S.binder * N + K.b] = (int8_t)S.newst[lane];
        } else {
            const int c0 = (S.move == CHROMO_TANGENT_ROTATION) ? 3 : 0;
            const int c1 = (S.move == CHROMO_SLIDE) ? 3 : 9;
#pragma unroll
            for (int c = 0; c < 9; ++c)
                if (c >= c0 && c < c1) R->pos[c * N + K.b] = K.w[c];
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (K.own[j])
#pragma unroll
            for (int m = 0; m < Bp1; ++m) R->density[(int64_t)K.vox[j] * Bp1 + m] = K.nr[j][m];
}

__device__ __forceinline__ void slot_copy(SpecSlot &dst, const SpecSlot &src)
{
    const uint32_t *s = (const uint32_t *)&src;
    uint32_t *d = (uint32_t *)&dst;
    for (int i = lane_id(); i < (int)(sizeof(SpecSlot) / 4); i += WAVE) d[i] = s[i];
}


// Wavefront 0: copies the unresolved slots first..nq-1 of cs to the front of ns
// (only slot `first`, marked SLOT_SEQ, when rewind is set).  Returns their count.
__device__ __forceinline__ int spec_carry(const SpecSlot *cs, SpecSlot *ns, int first, int nq, int rewind)
{
    int j = 0;
    if (rewind) {
        slot_copy(ns[0], cs[first]);
        wave_sync();
        if (lane_id() == 0) ns[0].kind = SLOT_SEQ;
        j = 1;
    } else {
        for (int s = first; s < nq; ++s) slot_copy(ns[j++], cs[s]);
    }
    wave_sync();
    return j;
}

// Wavefront 0: tops the queue cs[0..nq) up with new proposals -- at most W, at
// most `remaining`, none after a SLOT_SEQ one -- and publishes the control words.
template <int NB, int W>
__device__ __forceinline__ void spec_fill(const chromo_replica *R, WaveRng &rng, const chromo_ctrl &C,
                                          const SpecLds &L, SpecSlot *cs, int nq, int remaining)
{
    int j = nq;
    bool stop = j > 0 && cs[j - 1].kind == SLOT_SEQ;
    while (!stop && j < W && j < remaining) {
        const int kind = gen_slot<NB>(R, rng, C, cs[j]);
        ++j;
        stop = kind == SLOT_SEQ;
    }
    int cnt = 0;
    while (cnt < j && cs[cnt].kind != SLOT_SEQ) ++cnt;
    if (lane_id() == 0) {
        L.ctl->nq = j;
        L.ctl->cnt = cnt;
        L.ctl->remaining = remaining;
        L.ctl->used = 0;
    }
}

// Wavefront 0: AcceptanceTracker updates (mc_stat.py:190-208), controller counts,
// replica counters and log records of the resolved slots 0..F-1, one slot per
// lane; the same totals as record() called once per slot in order.
__device__ __forceinline__ void spec_book(const chromo_replica *R, chromo_ctrl &C, Counters &K,
                                          const SpecSlot *cs, const SpecRes *res, int F)
{
    const int lane = lane_id();
    const bool live = lane < F && cs[lane < F ? lane : 0].kind != SLOT_EMPTY;
    int64_t n = 0, nb = 0, bytes = 0;
    int acc = 0, tie = 0;
    if (live) {
        const SpecSlot &S = cs[lane];
        const SpecRes &q = res[lane];
        n = S.n;
        nb = q.nbins;
        acc = q.accepted;
        tie = fabs(q.u - q.expdE) < 1E-12;
        bytes = move_bytes(S.move, n, nb, R->B, acc);
    }
    const uint64_t live_m = __ballot(live);
    const uint64_t acc_m = __ballot(live && acc);
    const uint64_t tie_m = __ballot(live && tie);
    static_assert(SPEC_WMAX <= 8, "bookkeeping sums over the first eight lanes");
    const int64_t sn = sum_lanes8_i64(n), sb = sum_lanes8_i64(nb), sy = sum_lanes8_i64(bytes);
    const int64_t logged = K.logged;
    const int nlive = __popcll(live_m);
    if (R->log != nullptr && live) {
        const int64_t at = logged + __popcll(live_m & ((1ull << lane) - 1));
        if (at < R->log_cap) {
            const SpecSlot &S = cs[lane];
            const SpecRes &q = res[lane];
            chromo_move_log &Lg = R->log[at];
            Lg.move = S.move;
            Lg.accepted = q.accepted;
            Lg.n = S.n;
            Lg.ind0 = S.move == CHROMO_TANGENT_ROTATION ? S.picks[0] : S.ind0;
            Lg.indf = S.move == CHROMO_TANGENT_ROTATION ? S.picks[0] + 1 : S.indf;
            Lg.dE_poly = q.dEp;
            Lg.dE_field = q.dEf;
            Lg.u = q.u;
            Lg.exp_dE = q.expdE;
            Lg.n_bins = q.nbins;
        }
    }
    wave_sync();
    if (lane == 0) {
        double rate = C.rate;
        for (int s = 0; s < F; ++s)
            if ((live_m >> s) & 1) rate = (C.alpha * (double)((acc_m >> s) & 1)) + (1 - C.alpha) * rate;
        C.rate = rate;
        C.num_attempt += F;
        C.num_success += __popcll(acc_m);
        K.prop += nlive;
        K.acc += __popcll(acc_m);
        K.sumn += sn;
        K.sumb += sb;
        K.ties += __popcll(tie_m);
        K.bytes += sy;
        if (R->log != nullptr && logged < R->log_cap)
            K.logged = logged + nlive < R->log_cap ? logged + nlive : R->log_cap;
    }
    wave_sync();
}

// One controller block.  C and K live in LDS: only wavefront 0 touches them, and
// keeping them out of registers leaves those to the evaluation.
template <int NB, int W>
__device__ __forceinline__ void spec_block(const chromo_replica *R, WaveRng &rng, chromo_ctrl &C, Counters &K,
                                           const SpecLds &L)
{
    const int wv = wave_id(), lane = lane_id();
    int cur = 0;
    int nq = 0;                        // wavefront 0: slots queued in the current buffer
    int remaining = C.num_per_cycle;   // wavefront 0: proposals of the block not yet resolved
    int pend = 0;                      // wavefront 0: error to publish before the next barrier
    int64_t rounds = 0;
    // every round resolves a proposal or turns its first slot into a sequential one
    const int64_t max_rounds = 2 * (int64_t)C.num_per_cycle + 64;
    for (;;) {
        SpecSlot *cs = cur ? L.slots[1] : L.slots[0];
        SpecSlot *ns = cur ? L.slots[0] : L.slots[1];
        if (wv == 0) {
            // L.ctl->err is written only here, before the barrier after which every
            // wavefront reads it
            if (++rounds > max_rounds) pend = ERR_SPIN;
            if (pend && lane == 0) L.ctl->err = pend;
            if (!pend && !L.ctl->err) spec_fill<NB, W>(R, rng, C, L, cs, nq, remaining);
        }
        __syncthreads();
        if (L.ctl->err || L.ctl->remaining <= 0) break;
        const int cnt = L.ctl->cnt, nqa = L.ctl->nq;
        if (cs[0].kind == SLOT_SEQ) {
            if (wv == 0) {
                SeqIO *io = L.seq;
                for (int i = lane; i < CK_WORDS; i += WAVE) io->ck[i] = cs[0].ck[i];
                if (lane == 0) {
                    io->mti = rng.mti;
                    io->C = C;
                }
                wave_sync();
                seq_step<NB, W>(R, io, L.mt);
                rng_restore(rng, io->ck);
                rng.mti = io->mti;
                if (lane == 0) C = io->C;
                wave_sync();
                const MoveOut o = io->out;
                if (o.err) {
                    pend = o.err;
                } else if (o.n > 0) {
                    record(R, C, o, K);
                }
                remaining -= 1;
                nq = 0;
            }
        } else {
            SpecKeep keep;
            const bool mine = wv < cnt && cs[wv].kind == SLOT_SMALL;
            if (mine) eval_small<NB>(R, cs[wv], L.res[wv], keep, L, wv);
            __syncthreads();
            if (mine && !L.res[wv].resample && !L.res[wv].overflow) spec_conflicts(cs, L.res, keep, L, wv);
            __syncthreads();
            // resolution, identical in every wavefront
            int F = cnt, rew = 0;
            uint32_t accm = 0;
            for (int s = 0; s < cnt; ++s) {
                if (cs[s].kind == SLOT_EMPTY) continue;
                const SpecRes &q = L.res[s];
                if ((q.cmask & accm) || (q.overflow && s > 0)) {
                    F = s;
                    break;
                }
                // the first slot always fits the conflict table; if it did not, the
                // sequential path redoes it
                if (q.resample || q.overflow) {
                    F = s;
                    rew = 1;
                    break;
                }
                if (q.accepted) accm |= 1u << s;
            }
            if (mine && wv < F && L.res[wv].accepted) spec_commit<NB>(R, cs[wv], keep);
            if (mine) {
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (keep.ins[j]) {
                        L.hkey[keep.hs[j]] = HKEY_EMPTY;
                        L.hmask[keep.hs[j]] = 0;
                    }
            }
            if (wv == 0) {
                spec_book(R, C, K, cs, L.res, F);
                nq = spec_carry(cs, ns, F, nqa, rew);
                remaining -= F;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

template <int NB, int W>
__device__ __forceinline__ void spec_run(const chromo_replica *R, int64_t steps)
{
    constexpr int Bp1 = NB + 1;
    const SpecLayout SL = spec_layout(Bp1, W);
    SpecLds L;
    L.slots[0] = (SpecSlot *)(lds_dyn + SL.slots);
    L.slots[1] = L.slots[0] + W;
    L.res = (SpecRes *)(lds_dyn + SL.res);
    L.hkey = (int32_t *)(lds_dyn + SL.hkey);
    L.hmask = (uint32_t *)(lds_dyn + SL.hmask);
    L.ctl = (SpecCtl *)(lds_dyn + SL.ctl);
    L.seq = (SeqIO *)(lds_dyn + SL.seq);
    L.mt = (uint32_t *)(lds_dyn + SL.mt);
    __shared__ chromo_ctrl sctrl[MAX_CTRL];
    __shared__ Counters sK;
    const int wv = wave_id(), lane = lane_id();
    for (int i = (int)threadIdx.x; i < SPEC_HASH; i += WAVE * W) {
        L.hkey[i] = HKEY_EMPTY;
        L.hmask[i] = 0;
    }
    const int nc = R->n_ctrl < MAX_CTRL ? R->n_ctrl : MAX_CTRL;
    WaveRng rng;
    bool need_mt = false;
    if (wv == 0) {
        lds_hash_clear(Bp1);
        need_mt = any_binding(R);
        rng_load(R, rng, L.mt, need_mt);
        if (lane == 0) {
            counters_load(R, sK);
            for (int i = 0; i < nc; ++i) sctrl[i] = R->ctrl[i];
            L.ctl->err = 0;
        }
    }
    __syncthreads();
    for (int64_t s = 0; s < steps; ++s) {
        for (int ci = 0; ci < nc; ++ci) {
            if (sctrl[ci].on == 1) spec_block<NB, W>(R, rng, sctrl[ci], sK, L);
            if (wv == 0 && lane == 0 && !L.ctl->err) update_amplitudes_dev(sctrl[ci]);
            __syncthreads();
            if (L.ctl->err) break;
        }
        if (L.ctl->err) break;
    }
    if (wv == 0) {
        rng_store(R, rng, need_mt);
        if (lane == 0)
            for (int i = 0; i < nc; ++i) R->ctrl[i] = sctrl[i];
        counters_store(R, sK, L.ctl->err);
    }
}

}  // namespace
