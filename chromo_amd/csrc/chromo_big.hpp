// chromo_big.hpp -- evaluation of a long crank-shaft, end-pivot or slide move
// (more beads than a wavefront has lanes) by one wavefront, in LDS.
//
// The beads are taken BIG_CHUNK at a time, in order: chunk c holds the move's
// voxel positions 2 BIG_CHUNK c .. 2 BIG_CHUNK (c + 1) - 1 (position p = 2i + k,
// k = 0 current, k = 1 trial, as fields.pyx:1434-1520 walks them) and their
// contributions e = 8p + l.
//
//  * An LDS hash keyed by voxel lives across the chunks.  A voxel is numbered at
//    its first contribution, chunk after chunk and in contribution order within a
//    chunk: the reference's unique voxels in first-occurrence order (U = n_bins).
//  * Per chunk, each slot of the hash gathers the 64-bit mask of the chunk's
//    positions that touch the voxel.  The lane that takes the voxel visits those
//    positions in increasing order and adds their contributions, with the 1e-18
//    rules of fields.pyx:1491-1520, to the voxel's delta densities, which stay in
//    LDS from chunk to chunk.  So every voxel sees its contributions in the
//    reference's order: the delta densities are its left fold bit for bit.
//  * Energy terms per voxel, wave reduction, Metropolis, as in seg_eval.
//
// A move with more unique voxels than the block holds is reported as an overflow
// (the caller evaluates it through the global-arena path).  The commit recomputes
// the trial coordinates chunk by chunk from the unchanged current ones.
#pragma once

#include "chromo_seg.hpp"

namespace {

using namespace chromo;

constexpr int BIG_CHUNK = 32;               // beads per chunk
constexpr int BIG_POS = 2 * BIG_CHUNK;      // positions per chunk: one 64-bit mask
constexpr int BIG_E = 8 * BIG_POS;          // contributions per chunk
constexpr int BIG_HASH = 2048;              // hash slots
constexpr int BIG_HASH_SHIFT = 21;          // 32 - log2(BIG_HASH)
static_assert(BIG_POS == 64, "one mask word per chunk");

// Exchange words of the two-wavefront evaluation: per hash round the number of
// first contributions and of newly numbered voxels, and the hash-full flag.
constexpr int BIG_ROUNDS = BIG_E / WAVE;
constexpr int BIG_XCH_WORDS = 2 * BIG_ROUNDS + 2;

struct BigLayout {
    size_t c0, pw, prec, eslot, tlist, hkey, hq, hpos, xch, uvox, uacc, total;
    int ucap;   // unique voxels the block holds (0: too small to use)
};

// LDS block of `bytes` bytes for B + 1 = Bp1 species.  Before a chunk is hashed
// at most ucap keys are present and the chunk adds at most BIG_E: every probe
// finds a free slot.
__host__ __device__ inline BigLayout big_layout(int Bp1, size_t bytes)
{
    BigLayout L;
    size_t o = 0;
    L.c0 = o;    o += sizeof(double) * BIG_E;            // corner weight / vol_bin per contribution
    L.pw = o;    o += sizeof(double) * 4 * BIG_POS;      // lower-corner weights per position
    L.prec = o;  o += 16 * (size_t)BIG_POS;              // position records (int4)
    L.eslot = o; o += sizeof(int32_t) * BIG_E;           // hash slot per contribution
    L.tlist = o; o += sizeof(int32_t) * BIG_E;           // hash slots the chunk touches
    L.hpos = o;  o += sizeof(uint64_t) * BIG_HASH;       // chunk positions touching the voxel
    L.hkey = o;  o += sizeof(int32_t) * BIG_HASH;        // voxel
    L.hq = o;    o += sizeof(int32_t) * BIG_HASH;        // its number, -1 before it has one
    L.xch = o;   o += sizeof(int32_t) * BIG_XCH_WORDS;   // two-wavefront exchange
    const size_t per = sizeof(int32_t) + sizeof(double) * (size_t)Bp1;
    long avail = (long)bytes - (long)o - 16;
    long cap = avail > 0 ? avail / (long)per : 0;
    if (cap > BIG_HASH - BIG_E - 1) cap = BIG_HASH - BIG_E - 1;
    L.ucap = (int)cap;
    L.uacc = o;  o += sizeof(double) * (size_t)Bp1 * (size_t)(cap > 0 ? cap : 0);   // delta, then new densities
    L.uvox = o;  o += sizeof(int32_t) * (size_t)(cap > 0 ? cap : 0);
    L.total = o;
    return L;
}

struct BigLds {
    double *c0, *pw, *uacc;
    int4 *prec;
    int32_t *eslot, *tlist, *hkey, *hq, *uvox, *xch;
    uint64_t *hpos;
    int ucap;
};

__device__ __forceinline__ BigLds big_lds(char *base, int Bp1, size_t bytes)
{
    const BigLayout L = big_layout(Bp1, bytes);
    BigLds S;
    S.c0 = (double *)(base + L.c0);
    S.pw = (double *)(base + L.pw);
    S.prec = (int4 *)(base + L.prec);
    S.eslot = (int32_t *)(base + L.eslot);
    S.tlist = (int32_t *)(base + L.tlist);
    S.hpos = (uint64_t *)(base + L.hpos);
    S.hkey = (int32_t *)(base + L.hkey);
    S.hq = (int32_t *)(base + L.hq);
    S.xch = (int32_t *)(base + L.xch);
    S.uacc = (double *)(base + L.uacc);
    S.uvox = (int32_t *)(base + L.uvox);
    S.ucap = L.ucap;
    return S;
}

__device__ __forceinline__ bool big_move(int mv)
{
    return mv == CHROMO_CRANK_SHAFT || mv == CHROMO_END_PIVOT || mv == CHROMO_SLIDE;
}

// Voxel of corner l of position record pr.
__device__ __forceinline__ int big_corner(const int4 &pr, int l, int nx, int ny)
{
    const int x = (l & 1) ? (pr.x >> 16) : (pr.x & 0xffff);
    const int y = (l & 2) ? (pr.y >> 16) : (pr.y & 0xffff);
    const int z = (l & 4) ? (pr.z >> 16) : (pr.z & 0xffff);
    return x + nx * (y + ny * z);
}

// The first-occurrence numbering of one chunk by both wavefronts (wavefront wv
// takes the hash rounds r = wv mod 2): pass 1 finds each round's first
// contributions and newly seen voxels and publishes their counts in S.xch; pass 2
// places them after the earlier rounds' ones.  T (0 on entry) and U come out as
// the one-wavefront loop leaves them.
template <int NB>
__device__ __forceinline__ void big_number_coop(const BigLds &S, int rounds, int E, bool degen, int nx, int ny,
                                                int wv, int &T, int &U)
{
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    // contribution e's slot, whether it is the slot's first in the chunk, and
    // whether that voxel is new; recomputed in pass 2 rather than held across
    // the barrier (hq[h] is written only by h's first contribution, after its read)
    auto classify = [&](int e, int &h, bool &first, bool &fresh) {
        first = false;
        fresh = false;
        h = 0;
        if (e >= E) return;
        const int p = e >> 3, l = e & 7;
        h = S.eslot[e];
        bool f = (S.hpos[h] & ((1ull << p) - 1)) == 0;   // no earlier position touches it
        if (f && degen) {
            // corners of one position can coincide on a one-voxel axis
            const int4 pr = S.prec[p];
            const int v = big_corner(pr, l, nx, ny);
            for (int l2 = 0; l2 < l; ++l2)
                if (big_corner(pr, l2, nx, ny) == v) f = false;
        }
        first = f;
        fresh = f && S.hq[h] < 0;
    };
    for (int r = wv; r < rounds; r += 2) {
        int h;
        bool first, fresh;
        classify(lane + WAVE * r, h, first, fresh);
        const int ct = __popcll(__ballot(first)), cn = __popcll(__ballot(fresh));
        if (lane == 0) {
            S.xch[r] = ct;
            S.xch[BIG_ROUNDS + r] = cn;
        }
    }
    __syncthreads();
    int toff = T, uoff = U;
    for (int r = 0; r < rounds; ++r) {
        if ((r & 1) == wv) {
            int h;
            bool first, fresh;
            classify(lane + WAVE * r, h, first, fresh);
            int tcnt, ncnt;
            const int t = toff + wave_excl_count(first, tcnt);
            if (first) S.tlist[t] = h;
            const int q = uoff + wave_excl_count(fresh, ncnt);
            if (fresh && q < S.ucap) {
                S.hq[h] = q;
                S.uvox[q] = S.hkey[h];
#pragma unroll
                for (int m = 0; m < Bp1; ++m) S.uacc[q * Bp1 + m] = 0;
            }
        }
        toff += S.xch[r];
        uoff += S.xch[BIG_ROUNDS + r];
    }
    T = toff;
    U = uoff;
}

// Evaluates the crank-shaft, end-pivot or slide proposal P (n >= 1 beads, any
// length) against the committed state, like seg_eval: polymer dE
// (SSWLC.compute_dE, polymers.pyx:1024-1082), field dE
// (UniformDensityField.compute_dE, fields.pyx:1149-1233), Metropolis with u
// (mc_sim.pyx:152-182).  Writes nothing outside S.  Called by one whole wavefront.
//
// COOP: both wavefronts of the workgroup call it with the same proposal.  Per
// chunk, wavefront k places the positions 2i + k, the hash rounds alternate
// between them, the first-occurrence numbering is a two-pass prefix over the
// rounds (counts through S.xch), and the touched voxels' folds alternate by
// blocks of 64; the ordering by fold length and the energy terms stay on
// wavefront 0, which alone holds the result (wavefront 1 returns before the
// energy terms).  Every number is the one-wavefront evaluation's.
template <int NB, bool TW = false, bool COOP = false>
__device__ __forceinline__ void big_eval(const chromo_replica *R, const Proposal &P, double u, const BigLds &S,
                                         SegRes &res)
{
    static_assert(!COOP || !TW, "the cooperative evaluation has no twist terms");
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int wv = COOP ? wave_id() : 0;
    constexpr int NW = COOP ? 2 : 1;
    const int64_t N = R->N;
    const GLOBAL_AS double *Pp = gmem(R->pos);
    const GLOBAL_AS int8_t *states = gmem(R->states);
    const int64_t n = P.n, ind0 = P.ind0, indf = P.indf;
    const bool check_field = R->field_on;
    const int nx = (int)R->nx, ny = (int)R->ny, nz = (int)R->nz;
    const bool degen = nx == 1 || ny == 1 || nz == 1;
    const double vol_bin = R->vol_bin;
    const double *avol = R->access_vol;
    const bool thr = R->fast_field == 0;
    const int ct = R->confine_type;
    const double cl = R->confine_length;
    res.overflow = 0;
    const VoxGrid VG = vox_grid(R);

    for (int i = lane + WAVE * wv; i < BIG_HASH; i += WAVE * NW) {
        S.hkey[i] = HKEY_EMPTY;
        S.hq[i] = -1;
        S.hpos[i] = 0;
    }
    if (COOP && wv == 0 && lane == 0) S.xch[2 * BIG_ROUNDS] = 0;
    seg_sync<COOP>();

    double ef = 0, er = 0;
    int dbl[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) dbl[j] = 0;
    int out_t = 0, out_c = 0;
    int U = 0;
    int over = 0;
    const int64_t nchunks = (n + BIG_CHUNK - 1) / BIG_CHUNK;
    for (int64_t c = 0; c < nchunks && !over; ++c) {
        const int64_t i = c * BIG_CHUNK + lane;
        const bool act = lane < BIG_CHUNK && i < n;
        const int nb = (int)(n - c * BIG_CHUNK < BIG_CHUNK ? n - c * BIG_CHUNK : BIG_CHUNK);
        const int64_t b = act ? ind0 + i : ind0;
        double o[9], w[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) o[q] = 0;
        int8_t so[NB];
#pragma unroll
        for (int m = 0; m < NB; ++m) so[m] = 0;
        if (act) {
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = Pp[q * N + b];
#pragma unroll
            for (int m = 0; m < NB; ++m) so[m] = states[(int64_t)m * N + b];
        }
        seg_transform(P, o, w);

        // polymer dE: the two bonds that leave the segment
        if (act && i == 0 && ind0 != 0) {
            double r0[3], t30[3];
            for (int j = 0; j < 3; ++j) {
                r0[j] = Pp[j * N + ind0 - 1];
                t30[j] = Pp[(3 + j) * N + ind0 - 1];
            }
            Bond bd = load_bond(R->bonds, N, ind0 - 1);
            double twt = 0, twc = 0;
            if constexpr (TW) twist_forward(R->bonds, R->pos, N, ind0 - 1, o, w, twt, twc);
            ef = pair_dE_forward(bd, r0, &o[0], &w[0], t30, &o[3], &w[3], twt, twc);
        }
        if (act && i == n - 1 && indf != N) {
            double r1[3], t31[3];
            for (int j = 0; j < 3; ++j) {
                r1[j] = Pp[j * N + indf];
                t31[j] = Pp[(3 + j) * N + indf];
            }
            Bond bd = load_bond(R->bonds, N, indf - 1);
            double twt = 0, twc = 0;
            if constexpr (TW) twist_reverse(R->bonds, R->pos, N, indf, o, w, twt, twc);
            er = pair_dE_reverse(bd, &o[0], &w[0], r1, &o[3], &w[3], t31, twt, twc);
        }
        if (c == 0 && ind0 != 0) ef = readlane_f64(ef, 0);
        if (c == nchunks - 1 && indf != N) er = readlane_f64(er, (int)((n - 1) % BIG_CHUNK));
        if (!check_field) continue;

        // count_doubly_bound and confinement (no state change: trial states = current)
#pragma unroll
        for (int j = 0; j < NB; ++j) dbl[j] += __popcll(__ballot(act && so[j] == 2));
        if (ct == 1) {
            out_t += __popcll(__ballot(act && sqrt(dot3(&w[0], &w[0])) > cl));
            out_c += __popcll(__ballot(act && sqrt(dot3(&o[0], &o[0])) > cl));
        } else if (ct != 0) {
#pragma unroll
            for (int j = 0; j < 3; ++j) out_t += __popcll(__ballot(act && fabs(w[j]) > cl / 2));
        }

        // positions 2i (current) and 2i + 1 (trial), chunk-local
        int ties = 0;
        if (act) {
            int32_t sp = 0;
#pragma unroll
            for (int m = 0; m < NB; ++m) sp |= ((int32_t)(uint8_t)so[m]) << (8 * m);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (COOP && k != wv) continue;
                double pt[3];
                for (int j = 0; j < 3; ++j) pt[j] = k == 0 ? o[j] : w[j];
                int64_t ix[3];
                double wt[3];
                voxel_lower(VG, pt, ix, wt, ties);
                const int p = 2 * lane + k;
                const int x = (int)ix[0], y = (int)ix[1], z = (int)ix[2];
                S.prec[p] = make_int4(x | (wrap1(x, nx) << 16), y | (wrap1(y, ny) << 16),
                                      z | (wrap1(z, nz) << 16), sp);
#pragma unroll
                for (int j = 0; j < 3; ++j) S.pw[4 * p + j] = wt[j];
            }
        }
        count_ties(R, ties);
        seg_sync<COOP>();

        // hash the chunk's contributions; each slot collects the positions touching it
        const int E = 16 * nb;
        const int rounds = (E + WAVE - 1) / WAVE;
        int full = 0;
        for (int r = wv; r < rounds; r += NW) {
            const int e = lane + WAVE * r;
            if (e < E) {
                const int p = e >> 3, l = e & 7;
                const int4 pr = S.prec[p];
                const double *wq = &S.pw[4 * p];
                // weight / access volume of this contribution (fields.pyx:1472-1488)
                const int v = big_corner(pr, l, nx, ny);
                S.c0[e] = corner_weight(l, wq[0], wq[1], wq[2]) / (avol ? avol[v] : vol_bin);
                uint32_t h = ((uint32_t)v * 2654435761u) >> BIG_HASH_SHIFT;
                int probes = 0;
                for (; probes < BIG_HASH; ++probes) {
                    const int old = atomicCAS(&S.hkey[h], HKEY_EMPTY, v);
                    if (old == HKEY_EMPTY || old == v) break;
                    h = (h + 1) & (BIG_HASH - 1);
                }
                if (probes >= BIG_HASH) {
                    full = 1;
                    S.eslot[e] = -1;
                } else {
                    atomicOr((unsigned long long *)&S.hpos[h], 1ull << p);
                    S.eslot[e] = (int)h;
                }
            }
        }
        if constexpr (COOP) {
            if (wave_any(full) && lane == 0) S.xch[2 * BIG_ROUNDS] = 1;
            __syncthreads();
            full = S.xch[2 * BIG_ROUNDS];
        }
        if (wave_any(full)) {
            over = 1;
            break;
        }
        if constexpr (!COOP) wave_sync();

        // first contribution of each touched voxel in this chunk: the list of touched
        // slots, and numbers for the voxels not seen in earlier chunks
        int T = 0;
        if constexpr (COOP) {
            big_number_coop<NB>(S, rounds, E, degen, nx, ny, wv, T, U);
        } else {
        for (int r = 0; r < rounds; ++r) {
            const int e = lane + WAVE * r;
            bool first = false;
            int h = 0, v = 0;
            if (e < E) {
                const int p = e >> 3, l = e & 7;
                h = S.eslot[e];
                const uint64_t hp = S.hpos[h];
                first = (hp & ((1ull << p) - 1)) == 0;   // no earlier position touches it
                if (first && degen) {
                    // corners of one position can coincide on a one-voxel axis
                    const int4 pr = S.prec[p];
                    v = big_corner(pr, l, nx, ny);
                    for (int l2 = 0; l2 < l; ++l2)
                        if (big_corner(pr, l2, nx, ny) == v) first = false;
                }
            }
            int tcnt;
            const int t = T + wave_excl_count(first, tcnt);
            if (first) S.tlist[t] = h;
            T += tcnt;
            const bool fresh = first && S.hq[h] < 0;
            int ncnt;
            const int q = U + wave_excl_count(fresh, ncnt);
            if (fresh && q < S.ucap) {
                S.hq[h] = q;
                S.uvox[q] = S.hkey[h];
#pragma unroll
                for (int m = 0; m < Bp1; ++m) S.uacc[q * Bp1 + m] = 0;
            }
            U += ncnt;
        }
        }
        if (U > S.ucap) {
            over = 1;
            break;
        }
        seg_sync<COOP>();
        if (T > WAVE && wv == 0) {
            // touched voxels longest fold first (fold_class): each round of the
            // fold loop below then holds voxels of similar length
            constexpr int J = BIG_E / WAVE;
            int hh[J], cls[J], rank[J];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int t = lane + WAVE * j;
                hh[j] = t < T ? S.tlist[t] : 0;
                cls[j] = t < T ? fold_class(__popcll(S.hpos[hh[j]])) : 8;
                rank[j] = 0;
            }
            wave_rank_by_class<J>(cls, rank);
            wave_sync();
#pragma unroll
            for (int j = 0; j < J; ++j)
                if (cls[j] < 8) S.tlist[rank[j]] = hh[j];
            wave_sync();
        }
        if (COOP && T > WAVE) __syncthreads();

        // fold: touched voxel t on lane t mod 64, its positions in increasing order
        for (int t0 = WAVE * wv; t0 < T; t0 += WAVE * NW) {
            const int t = t0 + lane;
            const bool own = t < T;
            int h = 0, q = 0, vx = -1, vy = -1, vz = -1;
            uint64_t hits = 0;
            double acc[Bp1];
#pragma unroll
            for (int m = 0; m < Bp1; ++m) acc[m] = 0;
            if (own) {
                h = S.tlist[t];
                q = S.hq[h];
                const int v = S.hkey[h];
                vx = v % nx;
                vy = (v / nx) % ny;
                vz = v / (nx * ny);
                hits = S.hpos[h];
#pragma unroll
                for (int m = 0; m < Bp1; ++m) acc[m] = S.uacc[q * Bp1 + m];
            }
            for (;;) {
                const bool more = hits != 0;
                if (!wave_any(more)) break;
                if (!more) continue;
                const int p = __builtin_ctzll(hits);
                hits &= hits - 1;
                const int4 pr = S.prec[p];
                const double pre = (p & 1) ? 1. : -1.;
                const int x0 = pr.x & 0xffff, x1 = pr.x >> 16, y0 = pr.y & 0xffff, y1 = pr.y >> 16,
                          z0 = pr.z & 0xffff, z1 = pr.z >> 16;
                if (!degen) {
                    const int l = (vx == x1 ? 1 : 0) | (vy == y1 ? 2 : 0) | (vz == z1 ? 4 : 0);
                    fold_c0<NB>(S.c0[p * 8 + l], pre, pr.w, acc, thr);
                } else {
                    int dx = vx - x0, dy = vy - y0, dz = vz - z0;
                    if (dx < 0) dx += nx;
                    if (dy < 0) dy += ny;
                    if (dz < 0) dz += nz;
                    for (int l = 0; l < 8; ++l)
                        if (axis_hit(l & 1, dx, nx) && axis_hit((l >> 1) & 1, dy, ny) &&
                            axis_hit((l >> 2) & 1, dz, nz))
                            fold_c0<NB>(S.c0[p * 8 + l], pre, pr.w, acc, thr);
                }
            }
            if (own) {
#pragma unroll
                for (int m = 0; m < Bp1; ++m) S.uacc[q * Bp1 + m] = acc[m];
                S.hpos[h] = 0;
            }
        }
        seg_sync<COOP>();
    }
    if (over) {
        res.overflow = 1;
        res.U = U;
        return;
    }
    if (COOP && wv != 0) return;

    double dEp = 0;
    if (ind0 != 0) dEp += ef;
    if (indf != N) dEp += er;

    double dEf = 0;
    if (check_field) {
        // energy terms; the densities are read four voxels per lane at a time
        const GLOBAL_AS double *rho_g = gmem(R->density);
        double S2[NB], SX[NB * NB], NS = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) S2[j] = 0;
#pragma unroll
        for (int j = 0; j < NB * NB; ++j) SX[j] = 0;
        for (int q0 = 0; q0 < U; q0 += 4 * WAVE) {
            double rho[4][Bp1];
            int vv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int q = q0 + a * WAVE + lane;
                vv[a] = q < U ? S.uvox[q] : 0;
#pragma unroll
                for (int m = 0; m < Bp1; ++m) rho[a][m] = q < U ? rho_g[(int64_t)vv[a] * Bp1 + m] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int q = q0 + a * WAVE + lane;
                if (q < U) {
                    double acc[Bp1];
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) acc[m] = S.uacc[q * Bp1 + m];
                    voxel_dE_terms<NB>(R, vv[a], rho[a], acc, S2, SX, NS);
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) S.uacc[q * Bp1 + m] = rho[a][m] + acc[m];
                }
            }
        }
        const double share = wave_sum(lane_field_share<NB>(R, S2, SX, NS));
        dEf = field_dE_total<NB>(R, share, dbl, dbl, out_t, out_c, false);
    }

    double dE = 0;
    dE += dEp;
    if (check_field) dE += dEf;
    res.dEp = dEp;
    res.dEf = dEf;
    res.expdE = exp(-dE);
    res.accepted = u < res.expdE;
    res.U = check_field ? U : 0;
}

// Writes an accepted big_eval proposal: the trial coordinates, recomputed from the
// current ones (MCAdapter.accept, moves.pyx:190-226), and the new densities of the
// touched voxels (update_affected_densities, fields.pyx:1968-1975).
template <int NB>
__device__ __forceinline__ void big_commit(const chromo_replica *R, const Proposal &P, const BigLds &S, int U)
{
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int64_t N = R->N;
    GLOBAL_AS double *Pp = gmem(R->pos);
    const int c1 = P.move == CHROMO_SLIDE ? 3 : 9;
    for (int64_t i = lane; i < P.n; i += WAVE) {
        const int64_t b = P.ind0 + i;
        double o[9], w[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) o[q] = Pp[q * N + b];
        seg_transform(P, o, w);
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if (q < c1) Pp[q * N + b] = w[q];
    }
    GLOBAL_AS double *rho = gmem(R->density);
    for (int q = lane; q < U; q += WAVE) {
        const int64_t v = S.uvox[q];
#pragma unroll
        for (int m = 0; m < Bp1; ++m) rho[v * Bp1 + m] = S.uacc[q * Bp1 + m];
    }
}

}  // namespace
