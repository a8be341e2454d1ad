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
// against the current state with seg_eval, in its own LDS block.  Resolution
// walks the proposals in order: proposal s keeps its evaluated dE unless an
// accepted proposal s' < s of the same round wrote a bead or a voxel that s
// reads; then s and all later proposals go to the next round and are evaluated
// again.  Each chain therefore is the reference's sequential chain; only the
// number of rounds depends on conflicts.  Moves of more than SPEC_NMAX beads run
// one at a time through mc_step_dev.
// ===========================================================================

constexpr int SPEC_NMAX = SEG_MAX;    // beads (tangent picks) one wavefront evaluates
constexpr int SPEC_WMAX = 2;          // wavefronts per replica
constexpr int SPEC_HASH = 1024;       // conflict-table entries
constexpr int SPEC_HASH_SHIFT = 22;   // 32 - log2(SPEC_HASH)
constexpr int SPEC_HASH_SOFT = 512;   // insertions by wavefronts 1.. (wavefront 0 adds <= SEG_UMAX)
static_assert(SPEC_HASH_SOFT + SEG_UMAX <= SPEC_HASH, "the conflict table must never fill");
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
    int32_t head;                             // ring position of the round's first slot
    int32_t tag;                              // the round's number in the launch (never repeats)
    int32_t ready;                            // = tag once the later slots' conflict masks are set
    int32_t ins;                              // = tag once slot 0's voxels are in the conflict table
    int32_t stall;                            // = tag when a later wavefront gave up waiting for ins
};

// LDS exchange area of the one-at-a-time path (seq_step, seq_step_coop).
struct SeqIO {
    uint32_t ck[CK_WORDS];
    int32_t mti, pad0;
    chromo_ctrl C;
    MoveOut out;
    int32_t coop, pad1;      // seq_step_coop: 0 wavefront 0 alone, 1 both in seg_eval, 2 both in big_eval
    double u;                // its Metropolis draw
    int32_t move, binder;    // its wave-uniform fields
    int64_t n, ind0, indf;
    double T[12];
};

struct SpecLayout {
    size_t seg, seg_bytes, mt, slots, res, hkey, hmask, ctl, seq, total;
};

// LDS: one evaluator block per wavefront, the MT19937 block, two slot buffers,
// per-wavefront results, the conflict table and the control words.
__host__ __device__ inline SpecLayout spec_layout(int Bp1, int W)
{
    SpecLayout S;
    S.seg = 0;
    S.seg_bytes = seg_layout(Bp1).total;
    size_t o = S.seg_bytes * (size_t)W;
    S.mt = o;
    o = align_up(o + MT_LDS_BYTES, 16);
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
    SegLds mine;    // this wavefront's evaluator block
    SegLds first;   // wavefront 0's (the one-at-a-time path)
    char *wide;     // the W blocks as one wide block (wide_lds_bytes), or nullptr if they are too small
};
// A stream checkpoint in LDS: the glibc words and indices, and the Philox position
// (counter, index in block; words 35-38 are not used: rng_restore needs the
// position only).
__device__ __forceinline__ void rng_save(const WaveRng &g, uint32_t *ck)
{
    const int lane = lane_id();
    uint64_t ctr = 0;
    int bi = 4;
    if (g.mode == CHROMO_RNG_PHILOX) philox_pos(g, ctr, bi);
    if (lane < 31) ck[lane] = g.g;
    if (lane == 0) {
        ck[31] = (uint32_t)g.gf;
        ck[32] = (uint32_t)g.gr;
        ck[33] = (uint32_t)ctr;
        ck[34] = (uint32_t)(ctr >> 32);
        ck[39] = (uint32_t)bi;
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
    g.gf = uni_i32((int)ck[31]);
    g.gr = uni_i32((int)ck[32]);
    if (g.mode == CHROMO_RNG_PHILOX) philox_set(g, (uint64_t)ck[33] | ((uint64_t)ck[34] << 32), (int)ck[39]);
}

// capped_exponential (bead_selection.pyx:19-66) with the one draw x: ok is
// cleared when the reference would draw again (r > cap) or loop for ever.
__device__ __forceinline__ int64_t cexp_once(int32_t x, int64_t window, int64_t cap, bool &ok)
{
    if (cap < 1) {
        ok = false;
        return 0;
    }
    const int64_t r = cexp_trunc(x, window);
    if (r > cap) ok = false;
    return r;
}

// from_point (bead_selection.pyx) with the side draw x_side and one exponential
// draw x_exp.
__device__ __forceinline__ int64_t from_point_once(int32_t x_side, int32_t x_exp, int64_t w, int64_t nb,
                                                   int64_t ind0, bool &ok)
{
    if (w > nb || w < 1) {
        ok = false;
        return 0;
    }
    if ((x_side & 1) == 0) {
        const int64_t ws = imax64(imin64(w, ind0), 1), ub = imax64(ind0, 1);
        if (ws > ub) ok = false;
        return ub - cexp_once(x_exp, ws, ws, ok);
    }
    const int64_t ws = imax64(imin64(w, nb - ind0), 1), ub = imax64(nb - ind0, 1);
    if (ws > ub) ok = false;
    return cexp_once(x_exp, ws, ws, ok) + ind0;
}

// Philox, crank-shaft / end-pivot / slide: the usual proposal -- every capped
// exponential accepted at its first draw, a window of 1..SPEC_NMAX beads, no extra
// axis draws -- from the next eight words of the batch, read in one LDS round
// trip instead of one per draw.  The same words, in the same order and with the
// same arithmetic, as gen_slot's draw-at-a-time code.  Consumes nothing and returns
// false for any other case, which gen_slot then draws one word at a time.
template <int NB>
__device__ __forceinline__ bool gen_fast(const chromo_replica *R, WaveRng &rng, const chromo_ctrl &C, SpecSlot &S)
{
    const int mv = C.move;
    if (rng.mode != CHROMO_RNG_PHILOX ||
        (mv != CHROMO_CRANK_SHAFT && mv != CHROMO_END_PIVOT && mv != CHROMO_SLIDE))
        return false;
    const int p = uni_i32(rng.pidx);
    if (p + 8 > 4 * WAVE) return false;
    int32_t d[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) d[t] = (int32_t)(uni_u32(rng.mt[p + t]) >> 1);
    const int64_t N = R->N;
    const double amp_move = C.amp_move;
    const int64_t amp_bead = C.amp_bead;
    bool ok = true;
    int used = 0;
    int64_t ind0 = 0, indf = 0;
    double ang = 0;
    int32_t s1 = 0, s2 = 0, lhs = 0, ud = 0;
    if (mv == CHROMO_CRANK_SHAFT) {
        ang = amp_move * (((double)d[0] / RMAX) - 0.5);
        const int64_t b0 = (int64_t)(int32_t)(((double)d[1] / RMAX) * (double)N);
        const int64_t b1 = imax64(from_point_once(d[2], d[3], amp_bead, N, b0, ok), 1);
        check_bead_bounds(b0, b1, N, ind0, indf);
        ud = d[4];
        used = 5;
    } else if (mv == CHROMO_END_PIVOT) {
        ang = amp_move * (((double)d[0] / RMAX) - 0.5);
        lhs = (int32_t)(d[1] & 1);
        if (amp_bead > N) ok = false;
        const int64_t r = cexp_once(d[2], amp_bead, amp_bead, ok);
        if (lhs == 1) {
            ind0 = 0;
            indf = r + 1;
        } else {
            ind0 = N - r;
            indf = N;
        }
        if (indf > N) ok = false;
        s1 = d[3];
        s2 = d[4];
        ud = d[5];
        used = 6;
    } else {
        ang = amp_move * ((double)d[0] / RMAX);
        s1 = d[1];
        s2 = d[2];
        const int64_t b0 = draw_mod(d[3], N);
        const int64_t b1 = from_point_once(d[4], d[5], amp_bead, N, b0, ok);
        check_bead_bounds(b0, b1, N, ind0, indf);
        ud = d[6];
        used = 7;
    }
    const int64_t n = indf - ind0;
    if (!ok || n < 1 || n > SPEC_NMAX) return false;
    rng.pidx = p + used;
    if (lane_id() == 0) {
        S.kind = SLOT_SMALL;
        S.move = mv;
        S.binder = 0;
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
    return true;
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
    if (gen_fast<NB>(R, rng, C, S)) return SLOT_SMALL;
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
        lhs = (int32_t)(rnd31(rng) & 1);
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
        int64_t b0 = draw_mod(rnd31(rng), N);
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (!err) {
            check_bead_bounds(b0, b1, N, ind0, indf);
            n = indf - ind0;
        }
        seq = err || n > SPEC_NMAX;
    } else if (mv == CHROMO_TANGENT_ROTATION) {
        ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t k = draw_mod(rnd31(rng), amp_bead) + 1;
        n = k;
        seq = (k > N) || (k > SPEC_NMAX);
        int64_t fpick = 0;
        int32_t fd1 = 0, fd2 = 0;
        if (!seq && tangent_draws_at_once(rng, k, N, fpick, fd1, fd2)) {
            if (lane < k) {
                S.picks[lane] = fpick;
                S.ps1[lane] = fd1;
                S.ps2[lane] = fd2;
            }
        } else if (!seq) {
            int64_t mine = 0;
            for (int64_t i = 0; i < k && !seq; ++i) {
                for (int it = 0;; ++it) {
                    // a pick that does not end is left to mc_step_dev, which reports it
                    if (it >= SPIN_LIMIT) {
                        seq = true;
                        break;
                    }
                    int64_t c = draw_mod(rnd31(rng), N);
                    if (!wave_any(lane < i && mine == c)) {
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
        int64_t b0 = draw_mod(rnd31(rng), N);
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

// One proposal of a sequential round (the RNG, the controller and the result
// pass through LDS, io).  Inlined: a call boundary here makes the callee save
// most of the register file on every call and the caller spill around it.
// Wavefront 0 draws the proposal (mc_propose); a
// crank-shaft, slide or end-pivot move of more than SEG_MAX beads with the field
// on is then evaluated by both (up to SEG_WIDE beads seg_eval COOP, in the wide
// block, exchange area after it; longer ones big_eval COOP), everything else by
// wavefront 0 alone (mc_evaluate).  The numbers are those of mc_step_dev.  Both
// wavefronts call it.
template <int NB, int W>
__device__ __forceinline__ void seq_step_coop(const chromo_replica *R, SeqIO *io, uint32_t *mt, const SegLds &G,
                                              char *wide)
{
    static_assert(W == 2, "two wavefronts");
    const int wv = wave_id(), lane = lane_id();
    WaveRng rng;
    chromo_ctrl C;
    MoveOut o;
    Proposal P;
    double u = 0;
    bool go = false;
    const SegLds GW = seg_lds<SEG_WIDE>(wide, NB + 1);
    const BigLds GB = big_lds(wide, NB + 1, wide_lds_bytes(NB + 1));
    if (wv == 0) {
        rng.mode = uni_i32(R->rng_mode);
        rng.k0 = uni_u32(R->rng[CHROMO_RNG_PHILOX_KEY]);
        rng.k1 = uni_u32(R->rng[CHROMO_RNG_PHILOX_KEY + 1]);
        rng.pbase = PHILOX_NONE;
        rng.pidx = 0;
        rng.mt = mt;
        rng_restore(rng, io->ck);
        rng.mti = uni_i32(io->mti);
        C = io->C;
        o.err = 0;
        o.n = 0;
        go = mc_propose<NB>(R, rng, C, o, P, u);
        // the RNG and the controller go back to LDS now, so that they are not
        // held in registers through the evaluation
        rng_save(rng, io->ck);
        const bool both = go && R->field_on && wide != nullptr && P.n > SEG_MAX && big_move(P.move);
        const int coop = !both ? 0 : P.n <= SEG_WIDE ? 1 : GB.ucap > 0 ? 2 : 0;
        if (coop == 1) seg_clear<SEG_WIDE>(GW);
        if (lane == 0) {
            io->mti = rng.mti;
            io->C = C;
            io->coop = coop;
            io->u = u;
            io->move = P.move;
            io->binder = P.binder;
            io->n = P.n;
            io->ind0 = P.ind0;
            io->indf = P.indf;
            for (int q = 0; q < 12; ++q) io->T[q] = P.T[q];
        }
    }
    __syncthreads();
    const int coop = io->coop;
    if (coop) {
        if (wv != 0) {
            P.move = io->move;
            P.binder = io->binder;
            P.n = io->n;
            P.ind0 = io->ind0;
            P.indf = io->indf;
            for (int q = 0; q < 12; ++q) P.T[q] = io->T[q];
            P.mypick = 0;
            P.mynew = 0;
            for (int q = 0; q < 9; ++q) P.myT[q] = 0;
            u = io->u;
        }
        SegRes sr;
        if (coop == 1) {
            SegKeep<NB> K;
            seg_eval<NB, SEG_WIDE, false, true>(R, P, u, GW, sr, K, wide + seg_layout<SEG_WIDE>(NB + 1).total);
            if (wv == 0 && !sr.overflow) {
                if (sr.accepted) seg_commit<NB>(R, P.move, P.n, P.binder, P.mynew, K, GW, sr.U);
                wave_sync_global();
            }
        } else {
            big_eval<NB, false, true>(R, P, u, GB, sr);
            if (wv == 0 && !sr.overflow) {
                if (sr.accepted) big_commit<NB>(R, P, GB, sr.U);
                wave_sync_global();
            }
        }
        if (wv == 0) {
            // as mc_evaluate after its SEG_WIDE or long-move evaluation
            for (int b = 0; b < W; ++b)
                seg_clear<SEG_MAX>(seg_lds<SEG_MAX>(wide + (size_t)b * seg_layout<SEG_MAX>(NB + 1).total, NB + 1));
            if (!sr.overflow) {
                o.dEp = sr.dEp;
                o.dEf = sr.dEf;
                o.u = u;
                o.expdE = sr.expdE;
                o.accepted = sr.accepted;
                o.nbins = sr.U;
            } else if (P.n > R->bead_cap || P.n > R->seg_cap) {
                o.err = ERR_SEGCAP;
                o.n = 0;
            } else {
                move_body<1, NB>(R, u, P, o);
            }
            if (o.err) o.n = 0;
            if (lane == 0) {
                if (!o.err) track_acceptance(io->C, o.accepted);
                io->out = o;
            }
            wave_sync();
        }
    } else if (wv == 0) {
        if (go) mc_evaluate<NB>(R, C, o, P, u, G, wide, W);
        if (lane == 0) {
            io->C = C;
            io->out = o;
        }
        wave_sync();
    }
}

// The proposal of slot S from its stored draws: the arithmetic of mc_step_dev's
// proposal part (move_funcs.pyx).  The axis and fulcrum beads lie in the window or
// next to it: their coordinates come from B.  False when a crank-shaft's axis
// beads coincide: the reference then draws a random axis, which only the
// one-at-a-time path can do in stream order.
template <int NB>
__device__ __forceinline__ bool slot_proposal(const chromo_replica *R, const SpecSlot &S, Proposal &P,
                                              const SegBeads<NB> &B)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    const int mv = S.move;
    const int64_t n = S.n, ind0 = S.ind0, indf = S.indf;
    P.move = mv;
    P.n = n;
    P.ind0 = ind0;
    P.indf = indf;
    P.binder = S.binder;
    P.mypick = 0;
    P.mynew = 0;
#pragma unroll
    for (int q = 0; q < 12; ++q) P.T[q] = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) P.myT[q] = 0;
    if (mv == CHROMO_CRANK_SHAFT) {
        // get_crank_shaft_axis / _fulcrum, move_funcs.pyx:157-280
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
        const double sa = sin(S.ang), ca = cos(S.ang);
        double ra[3], rb[3], pt[3], dir[3];
        seg_r3(B, ia, ind0, indf, n, ra);
        seg_r3(B, ib, ind0, indf, n, rb);
        seg_r3(B, ifc, ind0, indf, n, pt);
        for (int i = 0; i < 3; ++i) dir[i] = ra[i] - rb[i];
        double mag = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        if (mag < 1E-5) return false;
        double sc = 1.0 / mag;
        for (int i = 0; i < 3; ++i) dir[i] = dir[i] * sc;
        axis_rotation_cs(dir, pt, ca, sa, P.T);
    } else if (mv == CHROMO_END_PIVOT) {
        // get_end_pivot_fulcrum, move_funcs.pyx:349-398
        int64_t ifc;
        if (ind0 == 0 && indf != N) ifc = indf;
        else if (ind0 != 0 && indf == N) ifc = ind0 - 1;
        else if (ind0 == 0 && indf == N && S.lhs == 1) ifc = indf - 1;
        else ifc = ind0;
        double pt[3], axis[3], sa, ca;
        sphere_from_uni(S.s1, S.s2, S.ang, axis, sa, ca);
        seg_r3(B, ifc, ind0, indf, n, pt);
        axis_rotation_cs(axis, pt, ca, sa, P.T);
    } else if (mv == CHROMO_SLIDE) {
        double dir[3], sa, ca;
        sphere_from_uni(S.s1, S.s2, 0.0, dir, sa, ca);
        for (int i = 0; i < 3; ++i) dir[i] *= S.ang;
        P.T[3] = dir[0];
        P.T[7] = dir[1];
        P.T[11] = dir[2];
    } else if (mv == CHROMO_TANGENT_ROTATION) {
        const double sa = sin(S.ang), ca = cos(S.ang);
        if (lane < n) {
            double dir[3], T[12];
            const double origin[3] = {0., 0., 0.};
            P.mypick = S.picks[lane];
            sphere_from(S.ps1[lane], S.ps2[lane], dir);
            axis_rotation_cs(dir, origin, ca, sa, T);
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < 3; ++q) P.myT[3 * j + q] = T[4 * j + q];
        }
    } else {
        if (lane < n) P.mynew = S.newst[lane];
    }
    return true;
}

// Wavefront `slot` evaluates slot S into res, keeps the trial state in K and in
// its evaluator block, and publishes the voxels it reads to the conflict table
// (their table slots replace the first-position entries of G.uaux).  Returns
// whether it inserted into the table.
template <int NB>
__device__ __forceinline__ bool eval_slot(const chromo_replica *R, const SpecSlot &S, SpecRes &res, SegKeep<NB> &K,
                                          const SpecLds &L, const SegLds &G, int slot)
{
    const int lane = lane_id();
    const int64_t ind0 = S.ind0, indf = S.indf;
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
    // every load of the evaluation but the densities', issued before its arithmetic
    SegBeads<NB> B;
    const int64_t pick = (S.move == CHROMO_TANGENT_ROTATION && lane < S.n) ? S.picks[lane] : 0;
    seg_load<NB>(R, S.move, S.n, ind0, indf, pick, B);
    Proposal P;
    if (!slot_proposal<NB>(R, S, P, B)) {
        if (lane == 0) res.resample = 1;
        wave_sync();
        return false;
    }
    SegRes sr;
    const double u = (double)S.ud / RMAX;
    seg_eval_loaded<NB>(R, P, u, G, sr, K, B);
    if (sr.overflow) {
        // more voxels than seg_eval holds: the one-at-a-time path evaluates it
        if (lane == 0) res.overflow = 1;
        wave_sync();
        return false;
    }
    const int U = sr.U;
    bool inserted = false;
    int full = 0;
    if (U > 0) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&L.ctl->used, U);
        base = bcast0(base);
        if (slot > 0 && base + U > SPEC_HASH_SOFT) {
            full = 1;
        } else {
            inserted = true;
            for (int q = lane; q < U; q += WAVE) {
                const int32_t v = G.uvox[q];
                uint32_t h = ((uint32_t)v * 2654435761u) >> SPEC_HASH_SHIFT;
                int probes = 0;
                for (;;) {
                    int old = atomicCAS(&L.hkey[h], HKEY_EMPTY, v);
                    if (old == HKEY_EMPTY || old == v) break;
                    h = (h + 1) & (uint32_t)(SPEC_HASH - 1);
                    if (++probes >= SPEC_HASH) {
                        full = 1;
                        break;
                    }
                }
                if (probes >= SPEC_HASH) {
                    G.uaux[q] = -1;
                    continue;
                }
                atomicOr(&L.hmask[h], 1u << slot);
                G.uaux[q] = (int32_t)h;
            }
        }
    }
    full = wave_any(full);
    if (lane == 0) {
        res.dEp = sr.dEp;
        res.dEf = sr.dEf;
        res.u = u;
        res.expdE = sr.expdE;
        res.accepted = sr.accepted;
        res.nbins = U;
        res.overflow = full;
    }
    wave_sync();
    return inserted;
}

// Earlier slots of this round whose writes slot s reads: shared voxels through
// the conflict table, shared beads through the bead ranges (tangent rotation:
// picks within one bead of each other).
template <int RING>
__device__ __forceinline__ const SpecSlot &ring_at(const SpecSlot *ring, int head, int s)
{
    return ring[(head + s) & (RING - 1)];
}

template <int RING>
__device__ __forceinline__ void spec_conflicts(const SpecSlot *ring, int head, SpecRes *res, const SpecLds &L,
                                               const SegLds &G, int s)
{
    static_assert((SPEC_WMAX - 1) * SPEC_NMAX <= WAVE, "one lane per (earlier slot, pick)");
    const int lane = lane_id();
    const uint32_t low = (1u << s) - 1u;
    uint32_t m = 0;
    const int U = res[s].nbins;
    for (int q = lane; q < U; q += WAVE) {
        const int h = G.uaux[q];
        if (h >= 0) m |= L.hmask[h] & low;
    }
    const SpecSlot &S = ring_at<RING>(ring, head, s);
    if (S.move == CHROMO_TANGENT_ROTATION) {
        const int a = lane / SPEC_NMAX, c = lane % SPEC_NMAX;
        if (a < s && ring_at<RING>(ring, head, a).kind == SLOT_SMALL && c < S.n) {
            const SpecSlot &A = ring_at<RING>(ring, head, a);
            const int64_t mine = S.picks[c];
            for (int64_t q = 0; q < A.n; ++q) {
                const int64_t d = A.picks[q] - mine;
                if (d >= -1 && d <= 1) m |= 1u << a;
            }
        }
    } else if (lane < s && ring_at<RING>(ring, head, lane).kind == SLOT_SMALL) {
        const SpecRes &o = res[lane];
        if (o.wlo <= res[s].rhi && res[s].rlo < o.whi) m |= 1u << lane;
    }
    m = wave_or(m);
    if (lane == 0) res[s].cmask = m;
}

// Clears this wavefront's conflict-table entries (after every wavefront has read
// the table this round).
__device__ __forceinline__ void spec_release(const SpecLds &L, const SegLds &G, int U)
{
    for (int q = lane_id(); q < U; q += WAVE) {
        const int h = G.uaux[q];
        if (h >= 0) {
            L.hkey[h] = HKEY_EMPTY;
            L.hmask[h] = 0;
        }
    }
}

// Wavefront 0: appends proposals to the queue of `q` slots starting at ring
// position head until it holds `limit`, or `remaining` (the block's unresolved
// proposals), or ends with a SLOT_SEQ one.  Returns the new queue length.
template <int NB, int RING>
__device__ __forceinline__ int spec_gen(const chromo_replica *R, WaveRng &rng, const chromo_ctrl &C, SpecSlot *ring,
                                        int head, int q, int limit, int remaining)
{
    if (q >= limit || q >= remaining) return q;
    bool stop = q > 0 && ring_at<RING>(ring, head, q - 1).kind == SLOT_SEQ;
    while (!stop && q < limit && q < remaining) {
        const int kind = gen_slot<NB>(R, rng, C, ring[(head + q) & (RING - 1)]);
        ++q;
        stop = kind == SLOT_SEQ;
    }
    return q;
}

// One wavefront: AcceptanceTracker updates (mc_stat.py:190-208), controller counts,
// replica counters and log records of the resolved slots 0..F-1, one slot per
// lane; the same totals as record() called once per slot in order.
template <int RING>
__device__ __forceinline__ void spec_book(const chromo_replica *R, chromo_ctrl &C, Counters &K,
                                          const SpecSlot *ring, int head, const SpecRes *res, int F)
{
    const int lane = lane_id();
    const bool live = lane < F && ring_at<RING>(ring, head, lane < F ? lane : 0).kind != SLOT_EMPTY;
    int64_t n = 0, nb = 0, bytes = 0;
    int acc = 0, tie = 0;
    if (live) {
        const SpecSlot &S = ring_at<RING>(ring, head, lane);
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
            const SpecSlot &S = ring_at<RING>(ring, head, lane);
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
        // every field is read before any is written: the compiler cannot tell C from
        // K, and interleaved reads and writes would wait on LDS one at a time
        const double alpha = C.alpha;
        double rate = C.rate;
        const auto na = C.num_attempt, nsu = C.num_success;
        const int64_t kp = K.prop, ka = K.acc, kn = K.sumn, kb = K.sumb, kt = K.ties, ky = K.bytes;
        for (int s = 0; s < F; ++s)
            if ((live_m >> s) & 1) rate = (alpha * (double)((acc_m >> s) & 1)) + (1 - alpha) * rate;
        const int nacc = __popcll(acc_m);
        C.rate = rate;
        C.num_attempt = na + F;
        C.num_success = nsu + nacc;
        K.prop = kp + nlive;
        K.acc = ka + nacc;
        K.sumn = kn + sn;
        K.sumb = kb + sb;
        K.ties = kt + __popcll(tie_m);
        K.bytes = ky + sy;
        if (R->log != nullptr && logged < R->log_cap)
            K.logged = logged + nlive < R->log_cap ? logged + nlive : R->log_cap;
    }
    wave_sync();
}

// One controller block.  C and K live in LDS: only wavefront 0 touches them, and
// keeping them out of registers leaves those to the evaluation.
//
// The slots wait in a ring of 2W.  A round evaluates the first min(W, queued)
// of them.  Slot 0 of a round stands whatever the others do (nothing precedes
// it), so wavefront 0 commits it as soon as it is evaluated -- a later slot that
// reads what it writes is redone anyway -- and then draws the following
// proposals into the ring while the other wavefronts finish.
template <int NB, int W>
__device__ __forceinline__ void spec_block(const chromo_replica *R, WaveRng &rng, chromo_ctrl &C, Counters &K,
                                           const SpecLds &L)
{
    constexpr int RING = 2 * W;
    static_assert((RING & (RING - 1)) == 0, "ring positions wrap by mask");
    const int wv = wave_id(), lane = lane_id();
    SpecSlot *ring = L.slots[0];
    int head = 0;                      // wavefront 0: ring position of the first queued slot
    int nq = 0;                        // wavefront 0: slots queued
    int remaining = C.num_per_cycle;   // wavefront 0: proposals of the block not yet resolved
    int pend = 0;                      // wavefront 0: error to publish before the next barrier
    int32_t tag = uni_i32(L.ctl->tag); // wavefront 0: the round's number in the launch
    int64_t rounds = 0;
    // every round resolves a proposal or turns its first slot into a sequential one
    const int64_t max_rounds = 2 * (int64_t)C.num_per_cycle + 64;
    // drawing ahead is off where a redone proposal would have to rewind MT19937
    const bool ahead = !(rng.mode == CHROMO_RNG_GLIBC && C.move == CHROMO_CHANGE_BINDING_STATE);
    for (;;) {
        if (wv == 0) {
            // L.ctl->err is written only here, before the barrier after which every
            // wavefront reads it
            if (++rounds > max_rounds) pend = ERR_SPIN;
            if (pend && lane == 0) L.ctl->err = pend;
            // only wavefront 0 writes L.ctl->err and L.ctl->tag: pend and tag are their values
            if (!pend) {
                tag += 1;
                nq = spec_gen<NB, RING>(R, rng, C, ring, head, nq, W, remaining);
                const int nr = nq < W ? nq : W;
                // leading slots that are not SLOT_SEQ, all kinds read at once
                const int kd = lane < nr ? ring_at<RING>(ring, head, lane).kind : SLOT_SEQ;
                const uint64_t seqm = __ballot(kd == SLOT_SEQ);
                const int cnt = seqm ? (int)__builtin_ctzll(seqm) : nr;
                if (lane == 0) {
                    L.ctl->nq = nr;
                    L.ctl->cnt = cnt;
                    L.ctl->remaining = remaining;
                    L.ctl->used = 0;
                    L.ctl->head = head;
                    L.ctl->tag = tag;
                }
            }
        }
        __syncthreads();
        if (L.ctl->err || L.ctl->remaining <= 0) break;
        const int cnt = L.ctl->cnt, h = L.ctl->head;
        if (ring_at<RING>(ring, h, 0).kind == SLOT_SEQ) {
            if (wv == 0) {
                SeqIO *io = L.seq;
                const SpecSlot &S0 = ring_at<RING>(ring, h, 0);
                for (int i = lane; i < CK_WORDS; i += WAVE) io->ck[i] = S0.ck[i];
                if (lane == 0) {
                    io->mti = rng.mti;
                    io->C = C;
                }
                wave_sync();
            }
            // the W evaluator blocks serve as one wide block; both wavefronts
            // enter, and evaluate a long crank-shaft, slide or end-pivot together
            seq_step_coop<NB, W>(R, L.seq, L.mt, L.first, L.wide);
            if (wv == 0) {
                SeqIO *io = L.seq;
                rng_restore(rng, io->ck);
                rng.mti = uni_i32(io->mti);
                if (lane == 0) C = io->C;
                wave_sync();
                const MoveOut o = io->out;
                if (o.err) {
                    pend = o.err;
                } else if (o.n > 0) {
                    record(R, C, o, K);
                }
                remaining -= 1;
                // nothing is drawn past a SLOT_SEQ slot
                head = (head + 1) & (RING - 1);
                nq = 0;
            }
        } else {
            SegKeep<NB> keep;
            const SpecSlot &mys = ring_at<RING>(ring, h, wv < W ? wv : 0);
            const bool mine = wv < cnt && mys.kind == SLOT_SMALL;
            bool inserted = false;
            if (mine) inserted = eval_slot<NB>(R, mys, L.res[wv], keep, L, L.mine, wv);
            bool committed = false;
            static_assert(W <= 2, "one wavefront on each side of the round's flags");
            if (wv == 0) {
                // slot 0's voxels are in the conflict table: the later slot may check
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0) __hip_atomic_store(&L.ctl->ins, L.ctl->tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (mine && L.res[0].accepted && !L.res[0].resample && !L.res[0].overflow) {
                    const int mynew = lane < mys.n ? mys.newst[lane] : 0;
                    seg_commit<NB>(R, mys.move, mys.n, mys.binder, mynew, keep, L.mine, L.res[0].nbins);
                    committed = true;
                }
            }
            if (wv > 0) {
                // wait for slot 0's table entries, set the later slot's conflict mask,
                // then the flag wavefront 0 waits on
                const int32_t tag = uni_i32(L.ctl->tag);
                int spins = 0;
                while (uni_i32(__hip_atomic_load(&L.ctl->ins, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != tag) {
                    if (++spins >= SPIN_LIMIT) break;
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                // a table that did not arrive: the slot waits for the next round, and
                // since this wavefront may then read slot 0's result before it is
                // written, wavefront 0 stops the replica (ERR_SPIN) rather than let
                // the two resolve the round differently
                if (spins >= SPIN_LIMIT && lane == 0) {
                    if (mine) L.res[wv].overflow = 1;
                    L.ctl->stall = tag;
                }
                wave_sync();
                if (mine && !L.res[wv].resample && !L.res[wv].overflow)
                    spec_conflicts<RING>(ring, h, L.res, L, L.mine, wv);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0) __hip_atomic_store(&L.ctl->ready, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                // draw the following proposals while the other wavefronts check
                // conflicts, commit and book
                if (ahead && !pend) nq = spec_gen<NB, RING>(R, rng, C, ring, head, nq, RING, remaining);
                if (W > 1) {
                    int spins = 0;
                    while (uni_i32(__hip_atomic_load(&L.ctl->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) !=
                           tag) {
                        if (++spins >= SPIN_LIMIT) {
                            pend = ERR_SPIN;
                            break;
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    if (uni_i32(L.ctl->stall) == tag) pend = ERR_SPIN;
                }
            }
            // resolution, identical in every wavefront; lane s reads slot s's result,
            // all slots in one LDS round trip
            uint32_t kd = SLOT_EMPTY, cmk = 0, ov = 0, rs = 0, ac = 0;
            if (lane < cnt) {
                const SpecRes &q = L.res[lane];
                kd = (uint32_t)ring_at<RING>(ring, h, lane).kind;
                cmk = q.cmask;
                ov = (uint32_t)q.overflow;
                rs = (uint32_t)q.resample;
                ac = (uint32_t)q.accepted;
            }
            int F = cnt, rew = 0;
            uint32_t accm = 0;
            for (int s = 0; s < cnt; ++s) {
                if ((int)readlane_u32(kd, s) == SLOT_EMPTY) continue;
                const uint32_t q_cmask = readlane_u32(cmk, s), q_overflow = readlane_u32(ov, s);
                if ((q_cmask & accm) || (q_overflow && s > 0)) {
                    F = s;
                    break;
                }
                // the first slot always fits the conflict table; if it did not, the
                // sequential path redoes it
                if (readlane_u32(rs, s) || q_overflow) {
                    F = s;
                    rew = 1;
                    break;
                }
                if (readlane_u32(ac, s)) accm |= 1u << s;
            }
            if (mine && !committed && wv < F && L.res[wv].accepted) {
                const int mynew = lane < mys.n ? mys.newst[lane] : 0;
                seg_commit<NB>(R, mys.move, mys.n, mys.binder, mynew, keep, L.mine, L.res[wv].nbins);
            }
            if (inserted) spec_release(L, L.mine, L.res[wv].nbins);
            // the last wavefront books the round; only the barrier below orders C
            // and K for the next reader
            if (wv == W - 1) spec_book<RING>(R, C, K, ring, h, L.res, F);
            if (wv == 0) {
                if (rew) {
                    // slot F is redone one at a time; what was drawn after it is
                    // dropped (the sequential path rewinds the streams to it)
                    if (lane == 0) ring[(head + F) & (RING - 1)].kind = SLOT_SEQ;
                    nq = F + 1;
                }
                head = (head + F) & (RING - 1);
                nq -= F;
                remaining -= F;
            }
        }
        __syncthreads();
    }
}

template <int NB, int W>
__device__ __forceinline__ void spec_run(const chromo_replica *R, int64_t steps)
{
    constexpr int Bp1 = NB + 1;
    const SpecLayout SL = spec_layout(Bp1, W);
    const int wv = wave_id(), lane = lane_id();
    SpecLds L;
    L.mine = seg_lds(lds_dyn + SL.seg + SL.seg_bytes * wv, Bp1);
    L.first = seg_lds(lds_dyn + SL.seg, Bp1);
    L.wide = SL.seg_bytes * W >= wide_lds_bytes(Bp1) ? lds_dyn + SL.seg : nullptr;
    seg_clear(L.mine);
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
    for (int i = (int)threadIdx.x; i < SPEC_HASH; i += WAVE * W) {
        L.hkey[i] = HKEY_EMPTY;
        L.hmask[i] = 0;
    }
    const int nc = R->n_ctrl < MAX_CTRL ? R->n_ctrl : MAX_CTRL;
    WaveRng rng;
    bool need_mt = false;
    if (wv == 0) {
        need_mt = any_binding(R);
        rng_load(R, rng, L.mt, need_mt);
        if (lane == 0) {
            counters_load(R, sK);
            for (int i = 0; i < nc; ++i) sctrl[i] = R->ctrl[i];
            L.ctl->err = 0;
            L.ctl->tag = 0;
            L.ctl->ready = -1;
            L.ctl->ins = -1;
            L.ctl->stall = -1;
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
