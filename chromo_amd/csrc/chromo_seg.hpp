// chromo_seg.hpp -- evaluation of one move of at most SEG_MAX beads by one
// wavefront.
//
// Lane i holds bead i (i < n): its current and trial coordinates, states and
// elastic partials.  A move has 2n voxel positions p = 2i + k (k = 0 current,
// k = 1 trial; chromo/fields.pyx:1434-1520 walks them in this order) and each
// position touches the 8 corners of its voxel cell, so contribution e = 8p + l.
//
//  1. The 16n contributions, 64 at a time in contribution order, enter a
//     per-wavefront LDS hash keyed by voxel that keeps each voxel's smallest
//     contribution index; a wave scan then numbers the first occurrences.  This
//     is the reference's list of unique voxels in first-occurrence order
//     (U = n_bins).
//  2. Voxel u goes to lane u mod 64 (at most SEG_VB per lane).  The hash also
//     keeps, per voxel, a 64-bit mask of the positions that touch it; the lane
//     visits those positions in increasing order and folds their contributions,
//     in contribution order, with the 1e-18 rules of fields.pyx:1491-1520: the
//     per-voxel delta density is the reference's left fold bit for bit.  The new
//     densities stay in registers until the commit.
//  3. Energy terms, wave reduction (fixed order), Metropolis.
//
// A move with more than SEG_VB * 64 unique voxels is reported as an overflow;
// the caller evaluates it through the global-arena path instead.
//
// The code is built for two capacities: SMAX = SEG_MAX (32 beads, the block every
// wavefront owns) and SMAX = SEG_WIDE (64 beads, a block that spans two of those).
#pragma once

#include "chromo_device.hpp"

namespace {

using namespace chromo;

constexpr int SEG_MAX = 32;                   // beads per move, per-wavefront block
constexpr int SEG_WIDE = 64;                  // beads per move, two-block block
constexpr int SEG_VB = 4;                     // unique voxels per lane
constexpr int SEG_UMAX = SEG_VB * 64;         // unique voxels one evaluation holds
constexpr int SEG_HASH = 1024;                // hash slots
constexpr int SEG_HASH_SHIFT = 22;            // 32 - log2(SEG_HASH)

// Sizes for a capacity of SMAX beads (one bead per lane, two positions per bead).
template <int SMAX>
struct SegCap {
    static_assert(SMAX <= 64, "one bead per lane");
    static constexpr int POS = 2 * SMAX;           // positions
    static constexpr int E = 8 * POS;              // contributions
    static constexpr int ROUNDS = E / 64;          // contribution rounds of 64 lanes
    static constexpr int MW = (POS + 63) / 64;     // 64-bit words of a position mask
    static_assert(E <= SEG_HASH, "at most one hash slot per contribution");
};

struct SegLayout {
    size_t c0, pw, prec, hpos, umask, hkey, hmin, eslot, uvox, ucrd, uaux, total;
};

// LDS block of the evaluator for moves of up to SMAX beads.
template <int SMAX = SEG_MAX>
__host__ __device__ inline SegLayout seg_layout(int)
{
    using C = SegCap<SMAX>;
    SegLayout S;
    size_t o = 0;
    S.c0 = o;    o += sizeof(double) * C::E;                   // corner weight / vol_bin per contribution
    S.pw = o;    o += sizeof(double) * 4 * C::POS;             // lower-corner weights per position
    S.prec = o;  o += 16 * (size_t)C::POS;                     // position records (int4)
    S.hpos = o;  o += sizeof(uint64_t) * C::MW * SEG_HASH;     // voxel -> positions touching it (bit mask)
    S.umask = o; o += sizeof(uint64_t) * C::MW * SEG_UMAX;     // the same per unique voxel
    S.hkey = o;  o += sizeof(int32_t) * SEG_HASH;              // voxel -> first contribution
    S.hmin = o;  o += sizeof(int32_t) * SEG_HASH;
    S.eslot = o; o += sizeof(int32_t) * C::E;                  // hash slot per contribution
    S.uvox = o;  o += sizeof(int32_t) * SEG_UMAX;              // unique voxel index
    S.ucrd = o;  o += 8 * (size_t)SEG_UMAX;                    // its cell (int2: x | y << 16, z)
    S.uaux = o;  o += sizeof(int32_t) * SEG_UMAX;              // hash slot; later a conflict-table slot
    S.total = (o + 255) / 256 * 256;
    return S;
}

struct SegLds {
    double *c0, *pw;
    int4 *prec;
    uint64_t *hpos, *umask;
    int32_t *hkey, *hmin, *eslot, *uvox, *uaux;
    int2 *ucrd;
};

template <int SMAX = SEG_MAX>
__device__ __forceinline__ SegLds seg_lds(char *base, int Bp1)
{
    const SegLayout L = seg_layout<SMAX>(Bp1);
    SegLds S;
    S.c0 = (double *)(base + L.c0);
    S.pw = (double *)(base + L.pw);
    S.prec = (int4 *)(base + L.prec);
    S.hpos = (uint64_t *)(base + L.hpos);
    S.umask = (uint64_t *)(base + L.umask);
    S.hkey = (int32_t *)(base + L.hkey);
    S.hmin = (int32_t *)(base + L.hmin);
    S.eslot = (int32_t *)(base + L.eslot);
    S.uvox = (int32_t *)(base + L.uvox);
    S.ucrd = (int2 *)(base + L.ucrd);
    S.uaux = (int32_t *)(base + L.uaux);
    return S;
}

// Empties the evaluator's hash (once per launch, by its wavefront).
template <int SMAX = SEG_MAX>
__device__ __forceinline__ void seg_clear(const SegLds &S)
{
    for (int i = lane_id(); i < SEG_HASH; i += WAVE) {
        S.hkey[i] = HKEY_EMPTY;
        S.hmin[i] = HMIN_EMPTY;
#pragma unroll
        for (int w = 0; w < SegCap<SMAX>::MW; ++w) S.hpos[i * SegCap<SMAX>::MW + w] = 0;
    }
    wave_sync();
}

// Wave-uniform results of one evaluation.
struct SegRes {
    double dEp, dEf, expdE;
    int U, accepted, overflow;
};

// Per-lane state kept from the evaluation to the commit: the trial coordinates of
// bead i == lane and the new densities of voxels lane + 64 j.
template <int NB>
struct SegKeep {
    double w[9];
    int64_t b;
    double nr[SEG_VB][NB + 1];
};

__device__ __forceinline__ int64_t pack_ix(const int64_t ix[3])
{
    return ix[0] | (ix[1] << 21) | (ix[2] << 42);
}

__device__ __forceinline__ bool axis_hit(int c, int d, int64_t nn)
{
    return nn == 1 ? true : (c == 0 ? d == 0 : d == 1);
}

// One contribution of the per-voxel fold (fields.pyx:1491-1520); c0 is the corner
// weight divided by vol_bin, sq the packed binder states of the position.
// thr: the 1e-18 rules apply (not under the fast field, fields.pyx:1346-1366).
template <int NB>
__device__ __forceinline__ void fold_c0(double c0, double pre, int32_t sq, double *acc, bool thr = true)
{
    const double t0 = pre * c0;
    if (!thr || fabs(t0) > 1E-18) acc[0] += t0;
#pragma unroll
    for (int m = 1; m <= NB; ++m) {
        const int st = (sq >> (8 * (m - 1))) & 0xff;
        const double t = pre * (c0 * (double)st);
        if (!thr || fabs(t) > 1E-18) acc[m] += t;
    }
}

// One entry of a voxel's fold list: the contribution's pre * c0 (pre = +-1, so
// exactly fold_c0's t0, and c * st exactly its pre * (c0 * st)) and the binder
// states of its position.
struct alignas(16) FoldEntry {
    double c;
    int32_t st, pad;
};

template <int NB>
__device__ __forceinline__ void fold_entry(const FoldEntry &f, double *acc, bool thr)
{
    if (!thr || fabs(f.c) > 1E-18) acc[0] += f.c;
#pragma unroll
    for (int m = 1; m <= NB; ++m) {
        const double t = f.c * (double)((f.st >> (8 * (m - 1))) & 0xff);
        if (!thr || fabs(t) > 1E-18) acc[m] += t;
    }
}

// The fold of one voxel's list (cnt entries from off) into acc, in the list's
// order (the reference's), two entries requested at a time.  Whole wave.
template <int NB>
__device__ __forceinline__ void fold_list(const FoldEntry *list, int off, int cnt, double *acc, bool thr)
{
    for (int k = 0; wave_any(k < cnt); k += 2) {
        FoldEntry f0 = {0., 0, 0}, f1 = {0., 0, 0};
        if (k < cnt) f0 = list[off + k];
        if (k + 1 < cnt) f1 = list[off + k + 1];
        if (k < cnt) fold_entry<NB>(f0, acc, thr);
        if (k + 1 < cnt) fold_entry<NB>(f1, acc, thr);
    }
}

// Contribution e of position p goes to entry off(q) + (positions of q's mask
// below p) of its voxel q's fold list (hmin maps its hash slot to q, ucrd.y's
// upper half holds off(q)).  The arrays are distinct LDS regions: with no
// aliasing to fear, the rounds' loads need not wait for earlier rounds' stores.
template <int MW, int ROUNDS>
__device__ __forceinline__ void seg_build_lists(int E, int rounds, const int32_t *__restrict__ eslot,
                                                const int32_t *__restrict__ hmin, const uint64_t *__restrict__ umask,
                                                const int2 *__restrict__ ucrd, const double *__restrict__ c0,
                                                const int4 *__restrict__ prec, FoldEntry *__restrict__ list,
                                                int r0 = 0, int rstep = 1)
{
    const int lane = lane_id();
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        if (r >= rounds) break;
        if ((r - r0) % rstep != 0) continue;
        const int e = lane + WAVE * r;
        if (e < E) {
            const int p = e >> 3;
            const int q = hmin[eslot[e]];
            int k = 0;
#pragma unroll
            for (int w = 0; w < MW; ++w) {
                const uint64_t mk = umask[q * MW + w];
                k += (w < (p >> 6)) ? __popcll(mk) : (w == (p >> 6) ? __popcll(mk & ((1ull << (p & 63)) - 1)) : 0);
            }
            FoldEntry f;
            f.c = ((p & 1) ? 1. : -1.) * c0[e];
            f.st = prec[p].w;
            f.pad = 0;
            list[(ucrd[q].y >> 16) + k] = f;
        }
    }
}

// Trial coordinates of bead (o) under proposal P, lane-local (move_funcs.pyx:139-154).
__device__ __forceinline__ void seg_transform(const Proposal &P, const double o[9], double w[9])
{
    const int mv = P.move;
    if (mv == CHROMO_CRANK_SHAFT || mv == CHROMO_END_PIVOT) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            w[j] = rowdot(&P.T[4 * j], o[0], o[1], o[2]) + P.T[4 * j + 3];
            w[3 + j] = rowdot(&P.T[4 * j], o[3], o[4], o[5]);
            w[6 + j] = rowdot(&P.T[4 * j], o[6], o[7], o[8]);
        }
    } else if (mv == CHROMO_SLIDE) {
#pragma unroll
        for (int j = 0; j < 3; ++j) w[j] = o[j] + P.T[4 * j + 3];
#pragma unroll
        for (int c = 3; c < 9; ++c) w[c] = o[c];
    } else if (mv == CHROMO_TANGENT_ROTATION) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            w[j] = o[j];
            w[3 + j] = rowdot(&P.myT[3 * j], o[3], o[4], o[5]);
            w[6 + j] = rowdot(&P.myT[3 * j], o[6], o[7], o[8]);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) w[c] = o[c];
    }
}

__device__ __forceinline__ int wrap1(int c, int nn) { return c + 1 == nn ? 0 : c + 1; }

// Neighbour block of a crank-shaft, end-pivot or slide move, one per wavefront
// (at most two per workgroup), filled straight from HBM into LDS: as doubles,
// 0-2 r and 3-5 t3 of bead ind0 - 1, 6-10 bond (ind0 - 1, ind0), 11-13 r and
// 14-16 t3 of bead indf, 17-21 bond (indf - 1, indf).
constexpr int SEG_NBR_WAVES = 2;
constexpr int SEG_NBR_DOUBLES = 22;

__device__ __forceinline__ double *seg_nbr()
{
    __shared__ double nbr[SEG_NBR_WAVES][32];
    return nbr[wave_id()];
}

// The neighbour block once its direct-to-LDS loads have landed.  The compiler's
// wait insertion does not tie these loads to later LDS reads of the block (it
// places its vmcnt wait after them), so the wait is explicit, and the asm's
// memory clobber keeps the reads below it.
__device__ __forceinline__ const double *seg_nbr_ready()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    return seg_nbr();
}

// What an evaluation reads of the chain, loaded in one batch before any of its
// arithmetic: lane i's bead (window bead ind0 + i, or tangent pick i), its
// coordinates, binder states and (binding moves) modifications, and the r, t3
// and bond parameters of the neighbours the polymer dE needs: beads ind0 - 1 and
// indf in the wavefront's neighbour block (seg_nbr), beads b - 1 and b + 1 of
// every tangent pick in registers.
template <int NB>
struct SegBeads {
    int64_t b;
    double o[9];
    int8_t so[NB], mods[NB];
    double lr[6], rr[6];   // r, t3 of beads b - 1 and b + 1
    Bond lb, rb;           // bonds (b - 1, b) and (b, b + 1)
    bool hasl, hasr;
};

template <int NB>
__device__ __forceinline__ void seg_load(const chromo_replica *R, int mv, int64_t n, int64_t ind0, int64_t indf,
                                         int64_t pick, SegBeads<NB> &B)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    const GLOBAL_AS double *Pp = gmem(R->pos);
    const bool tangent = mv == CHROMO_TANGENT_ROTATION;
    const bool binding = mv == CHROMO_CHANGE_BINDING_STATE;
    const bool act = lane < n;
    int64_t b = tangent ? pick : ind0 + lane;
    if (!act) b = tangent ? 0 : ind0;
    B.b = b;
#pragma unroll
    for (int c = 0; c < 9; ++c) B.o[c] = 0;
#pragma unroll
    for (int m = 0; m < NB; ++m) B.so[m] = B.mods[m] = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c) B.lr[c] = B.rr[c] = 0;
    B.lb = Bond{0, 0, 0, 0, 0};
    B.rb = Bond{0, 0, 0, 0, 0};
    if (act) {
#pragma unroll
        for (int c = 0; c < 9; ++c) B.o[c] = Pp[c * N + b];
#pragma unroll
        for (int m = 0; m < NB; ++m) B.so[m] = gmem(R->states)[(int64_t)m * N + b];
        if (binding) {
#pragma unroll
            for (int m = 0; m < NB; ++m) B.mods[m] = gmem(R->marks)[(int64_t)m * N + b];
        }
    }
    B.hasl = act && !binding && (tangent ? b != 0 : (lane == 0 && ind0 != 0));
    B.hasr = act && !binding && (tangent ? b + 1 != N : (lane == n - 1 && indf != N));
    if (tangent) {
        if (B.hasl) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                B.lr[j] = Pp[j * N + b - 1];
                B.lr[3 + j] = Pp[(3 + j) * N + b - 1];
            }
            B.lb = load_bond(R->bonds, N, b - 1);
        }
        if (B.hasr) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                B.rr[j] = Pp[j * N + b + 1];
                B.rr[3 + j] = Pp[(3 + j) * N + b + 1];
            }
            B.rb = load_bond(R->bonds, N, b);
        }
    } else if (!binding && N >= 2) {
        // dword j of the neighbour block on lane j, loaded into LDS directly: no
        // registers held while the proposal arithmetic runs
        const int d = lane >> 1;
        if (d < SEG_NBR_DOUBLES) {
            const bool left = d < 11;
            const int q = left ? d : d - 11;
            int64_t bead = left ? ind0 - 1 : indf, bond = left ? ind0 - 1 : indf - 1;
            bead = bead < 0 ? 0 : (bead > N - 1 ? N - 1 : bead);
            bond = bond < 0 ? 0 : (bond > N - 2 ? N - 2 : bond);
            const double *src = q < 6 ? R->pos + (int64_t)q * N + bead : R->bonds + (int64_t)(q - 6) * (N - 1) + bond;
            __builtin_amdgcn_global_load_lds((const void *)((const uint32_t *)src + (lane & 1)),
                                             (void __attribute__((address_space(3))) *)seg_nbr(), 4, 0, 0);
        }
    }
}

// r of bead x (wave-uniform, ind0 - 1 <= x <= indf, a bead seg_load fetched) of a
// non-tangent move of n beads.
template <int NB>
__device__ __forceinline__ void seg_r3(const SegBeads<NB> &B, int64_t x, int64_t ind0, int64_t indf, int64_t n,
                                       double v[3])
{
    if (x == ind0 - 1) {
        const double *nb = seg_nbr_ready();
        for (int j = 0; j < 3; ++j) v[j] = nb[j];
    } else if (x == indf) {
        const double *nb = seg_nbr_ready();
        for (int j = 0; j < 3; ++j) v[j] = nb[11 + j];
    } else {
        for (int j = 0; j < 3; ++j) v[j] = readlane_f64(B.o[j], (int)(x - ind0));
    }
}

// Evaluates proposal P (1 <= n <= SMAX) against the committed state: polymer
// dE (SSWLC.compute_dE, polymers.pyx:1024-1082), field dE
// (UniformDensityField.compute_dE, fields.pyx:1149-1233) and the Metropolis test
// with the uniform draw u (mc_sim.pyx:152-182).  Writes nothing outside S and K.
// Called by one whole wavefront; S's hash is empty on entry and on return.  B
// holds what seg_load fetched for this move.
// A barrier of the evaluator's phases: its wavefront, or with COOP both
// wavefronts of the workgroup.
template <bool COOP>
__device__ __forceinline__ void seg_sync()
{
    if constexpr (COOP) __syncthreads();
    else wave_sync();
}

// COOP: the two wavefronts of the workgroup evaluate one proposal together (a
// crank-shaft, slide or end-pivot move with the field on, its beads and the
// SEG_MAX-exceeding part of a sequential round).  Wavefront 1 takes the trial
// positions' voxel floors, every other round of contributions (hash, numbering,
// fold lists) and the odd voxel slots' folds, which it hands to wavefront 0
// through xbuf (LDS, COOP_XBUF_BYTES); everything whose order shapes a result
// (numbering, dealing, the energy terms, the reductions, Metropolis) runs on
// wavefront 0 in the order of the one-wavefront evaluation, so both give the same
// numbers.  res and K are wavefront 0's.
constexpr int COOP_XBUF_BYTES = 256 + 2 * 64 * 5 * 8;

template <int NB, int SMAX = SEG_MAX, bool TW = false, bool COOP = false>
__device__ __forceinline__ void seg_eval_loaded(const chromo_replica *R, const Proposal &P, double u,
                                                const SegLds &S, SegRes &res, SegKeep<NB> &K,
                                                const SegBeads<NB> &B, char *xbuf = nullptr)
{
    static_assert(!COOP || (SMAX == SEG_WIDE && !TW), "the cooperative evaluation is the 64-bead one");
    const int wv = COOP ? wave_id() : 0;
    using C = SegCap<SMAX>;
    constexpr int MW = C::MW;
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int64_t N = R->N;
    const int mv = P.move;
    const int64_t n = P.n, ind0 = P.ind0, indf = P.indf;
    const bool tangent = mv == CHROMO_TANGENT_ROTATION;
    const bool state_change = mv == CHROMO_CHANGE_BINDING_STATE;
    const bool check_field = R->field_on && !tangent;
    const bool act = lane < n;
    res.overflow = 0;

    // ---- bead i == lane: coordinates, states, trial
    K.b = B.b;
    double o[9];
    int8_t so[NB], sn[NB];
#pragma unroll
    for (int c = 0; c < 9; ++c) o[c] = B.o[c];
#pragma unroll
    for (int m = 0; m < NB; ++m) so[m] = sn[m] = B.so[m];
    if (act && state_change) {
#pragma unroll
        for (int m = 0; m < NB; ++m)
            if (m == P.binder) sn[m] = (int8_t)P.mynew;
    }
    double w[9];
    seg_transform(P, o, w);
#pragma unroll
    for (int c = 0; c < 9; ++c) K.w[c] = w[c];

    // ---- polymer dE
    double dEp = 0;
    if (tangent || state_change) {
        double e = 0;
        if (act && tangent) {
            double tlt = 0, tlc = 0, trt = 0, trc = 0;
            if constexpr (TW) {
                if (B.hasl) twist_forward(R->bonds, R->pos, N, B.b - 1, o, w, tlt, tlc);
                if (B.hasr) twist_reverse(R->bonds, R->pos, N, B.b + 1, o, w, trt, trc);
            }
            if (B.hasl) e += pair_dE_forward(B.lb, &B.lr[0], &o[0], &w[0], &B.lr[3], &o[3], &w[3], tlt, tlc);
            if (B.hasr) e += pair_dE_reverse(B.rb, &o[0], &w[0], &B.rr[0], &o[3], &w[3], &B.rr[3], trt, trc);
        } else if (act) {
            e = bead_binding_dE<NB>(R, B.mods, so, sn);
        }
        for (int i = 0; i < n; ++i) dEp += readlane_f64(e, i);
    } else {
        // the neighbours' r, t3 and bonds from the wavefront's neighbour block
        const double *nb = seg_nbr_ready();
        double ef = 0, er = 0;
        if (lane == 0 && ind0 != 0) {
            double r0[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) r0[j] = nb[j];
            const Bond bd{nb[6], nb[7], nb[8], nb[9], nb[10]};
            double twt = 0, twc = 0;
            if constexpr (TW) twist_forward(R->bonds, R->pos, N, ind0 - 1, o, w, twt, twc);
            ef = pair_dE_forward(bd, &r0[0], &o[0], &w[0], &r0[3], &o[3], &w[3], twt, twc);
        }
        if (lane == n - 1 && indf != N) {
            double r1[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) r1[j] = nb[11 + j];
            const Bond bd{nb[17], nb[18], nb[19], nb[20], nb[21]};
            double twt = 0, twc = 0;
            if constexpr (TW) twist_reverse(R->bonds, R->pos, N, indf, o, w, twt, twc);
            er = pair_dE_reverse(bd, &o[0], &w[0], &r1[0], &o[3], &w[3], &r1[3], twt, twc);
        }
        if (ind0 != 0) dEp += readlane_f64(ef, 0);
        if (indf != N) dEp += readlane_f64(er, (int)(n - 1));
    }

    // ---- field dE
    double dEf = 0;
    int U = 0;
    if (check_field) {
        const int nx = (int)R->nx, ny = (int)R->ny, nz = (int)R->nz;
        const double vol_bin = R->vol_bin;
        const double *avol = R->access_vol;
        const bool thr = R->fast_field == 0;
        // positions 2i (current) and 2i + 1 (trial; the current one for a state change)
        int ties = 0;
        const VoxGrid VG = vox_grid(R);
        if (act) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (COOP && k != wv) continue;
                double pt[3];
                for (int j = 0; j < 3; ++j) pt[j] = (k == 0 || state_change) ? o[j] : w[j];
                int64_t ix[3];
                double wt[3];
                voxel_lower(VG, pt, ix, wt, ties);
                const int p = 2 * lane + k;
                const int x = (int)ix[0], y = (int)ix[1], z = (int)ix[2];
                int32_t sp = 0;
#pragma unroll
                for (int m = 0; m < NB; ++m)
                    sp |= ((int32_t)(uint8_t)((k == 0 || !state_change) ? so[m] : sn[m])) << (8 * m);
                S.prec[p] = make_int4(x | (wrap1(x, nx) << 16), y | (wrap1(y, ny) << 16),
                                      z | (wrap1(z, nz) << 16), sp);
#pragma unroll
                for (int j = 0; j < 3; ++j) S.pw[4 * p + j] = wt[j];
            }
        }
        count_ties(R, ties);
        seg_sync<COOP>();

        // 1. unique voxels: hash the contributions in order, keep each voxel's first
        const int E = (int)(16 * n);
        const int rounds = (E + WAVE - 1) / WAVE;
#pragma unroll
        for (int r = 0; r < C::ROUNDS; ++r) {
            if (r >= rounds) break;
            if (COOP && (r & 1) != wv) continue;
            const int e = lane + WAVE * r;
            if (e < E) {
                const int4 pr = S.prec[e >> 3];
                const int l = e & 7;
                // weight / access volume of this contribution (fields.pyx:1472-1488)
                const double *wq = &S.pw[4 * (e >> 3)];
                const int x = (l & 1) ? (pr.x >> 16) : (pr.x & 0xffff);
                const int y = (l & 2) ? (pr.y >> 16) : (pr.y & 0xffff);
                const int z = (l & 4) ? (pr.z >> 16) : (pr.z & 0xffff);
                const int v = x + nx * (y + ny * z);
                S.c0[e] = corner_weight(l, wq[0], wq[1], wq[2]) / (avol ? avol[v] : vol_bin);
                uint32_t h = ((uint32_t)v * 2654435761u) >> SEG_HASH_SHIFT;
                // fewer distinct keys than slots before each insertion: the probe ends
                for (int probes = 0; probes < SEG_HASH; ++probes) {
                    const int old = atomicCAS(&S.hkey[h], HKEY_EMPTY, v);
                    if (old == HKEY_EMPTY || old == v) break;
                    h = (h + 1) & (SEG_HASH - 1);
                }
                atomicMin(&S.hmin[h], e);
                const int p = e >> 3;
                atomicOr((unsigned long long *)&S.hpos[h * MW + (p >> 6)], 1ull << (p & 63));
                S.eslot[e] = (int)h;
            }
        }
        seg_sync<COOP>();
        int32_t *rcnt = (int32_t *)xbuf;   // COOP: first occurrences per round
        if constexpr (COOP) {
#pragma unroll
            for (int r = 0; r < C::ROUNDS; ++r) {
                if (r >= rounds) break;
                if ((r & 1) != wv) continue;
                const int e = lane + WAVE * r;
                const int h = e < E ? S.eslot[e] : 0;
                const int c = __popcll(__ballot(e < E && S.hmin[h] == e));
                if (lane == 0) rcnt[r] = c;
            }
            __syncthreads();
        }
        int total = 0;
#pragma unroll
        for (int r = 0; r < C::ROUNDS; ++r) {
            if (r >= rounds) break;
            if (COOP && (r & 1) != wv) {
                total += rcnt[r];
                continue;
            }
            const int e = lane + WAVE * r;
            const int h = e < E ? S.eslot[e] : 0;
            const bool first = e < E && S.hmin[h] == e;
            int cnt;
            const int q = total + wave_excl_count(first, cnt);
            if (first && q < SEG_UMAX) {
                const int4 pr = S.prec[e >> 3];
                const int l = e & 7;
                const int x = (l & 1) ? (pr.x >> 16) : (pr.x & 0xffff);
                const int y = (l & 2) ? (pr.y >> 16) : (pr.y & 0xffff);
                const int z = (l & 4) ? (pr.z >> 16) : (pr.z & 0xffff);
                S.uvox[q] = x + nx * (y + ny * z);
                S.ucrd[q] = make_int2(x | (y << 16), z);
#pragma unroll
                for (int w = 0; w < MW; ++w) S.umask[q * MW + w] = S.hpos[h * MW + w];
                S.uaux[q] = h;
            }
            total += cnt;
        }
        U = total;
        seg_sync<COOP>();
        // Unless a grid axis has one voxel (a position's corners can then share a
        // voxel), each voxel's contributions are laid out below as one list, in
        // fold order, in the position-mask area of the hash; its hmin slots map
        // hash slot -> voxel meanwhile, and both are emptied after the fold.
        const bool degen = nx == 1 || ny == 1 || nz == 1;
        const bool listed = !degen && U <= SEG_UMAX;
        // release the hash (every key is some voxel's first occurrence)
        for (int q = lane; q < U && q < SEG_UMAX && wv == 0; q += WAVE) {
            const int h = S.uaux[q];
            S.hkey[h] = HKEY_EMPTY;
            if (!listed) {
                S.hmin[h] = HMIN_EMPTY;
#pragma unroll
                for (int w = 0; w < MW; ++w) S.hpos[h * MW + w] = 0;
            }
        }
        if (U > SEG_UMAX) {
            // too many voxels for the registers: the global-arena path evaluates it
            for (int e = lane; e < E && wv == 0; e += WAVE) {
                const int h = S.eslot[e];
                S.hkey[h] = HKEY_EMPTY;
                S.hmin[h] = HMIN_EMPTY;
#pragma unroll
                for (int w = 0; w < MW; ++w) S.hpos[h * MW + w] = 0;
            }
            wave_sync();
            res.overflow = 1;
            res.U = U;
            return;
        }
        wave_sync();
        if (U > WAVE && wv == 0) {
            // deal the voxels out longest fold first (fold_class), so that each
            // round of the fold loop below holds voxels of similar length
            int vv[SEG_VB], hh[SEG_VB], cls[SEG_VB], rank[SEG_VB];
            int2 cc[SEG_VB];
            uint64_t mm[SEG_VB][MW];
#pragma unroll
            for (int j = 0; j < SEG_VB; ++j) {
                const int q = lane + WAVE * j;
                const bool own = q < U;
                vv[j] = own ? S.uvox[q] : 0;
                hh[j] = own ? S.uaux[q] : 0;
                cc[j] = own ? S.ucrd[q] : make_int2(0, 0);
                int visits = 0;
#pragma unroll
                for (int w = 0; w < MW; ++w) {
                    mm[j][w] = own ? S.umask[q * MW + w] : 0;
                    visits += __popcll(mm[j][w]);
                }
                cls[j] = own ? fold_class(visits) : 8;
                rank[j] = 0;
            }
            wave_rank_by_class<SEG_VB>(cls, rank);
            wave_sync();
#pragma unroll
            for (int j = 0; j < SEG_VB; ++j) {
                if (cls[j] < 8) {
                    const int r = rank[j];
                    S.uvox[r] = vv[j];
                    S.uaux[r] = hh[j];
                    S.ucrd[r] = cc[j];
#pragma unroll
                    for (int w = 0; w < MW; ++w) S.umask[r * MW + w] = mm[j][w];
                }
            }
            wave_sync();
        }

        // 2. fold per voxel, voxel lane + 64 j on this lane
        const GLOBAL_AS double *rho_g = gmem(R->density);
        const int npos = (int)(2 * n);
        double S2[NB], SX[NB * NB], NS = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) S2[j] = 0;
#pragma unroll
        for (int j = 0; j < NB * NB; ++j) SX[j] = 0;
        FoldEntry *list = (FoldEntry *)S.hpos;
        if (listed && wv == 0) {
            // list offsets: voxels in (j, lane) order, each as long as its position mask;
            // the offset rides in the upper half of ucrd.y (z < 2^15)
            int base = 0;
#pragma unroll
            for (int j = 0; j < SEG_VB; ++j) {
                if (WAVE * j >= U) break;
                const int q = lane + WAVE * j;
                const bool own = q < U;
                int visits = 0;
#pragma unroll
                for (int w = 0; w < MW; ++w) visits += own ? __popcll(S.umask[q * MW + w]) : 0;
                int tot;
                const int off = base + wave_excl_scan(visits, tot);
                base += tot;
                if (own) {
                    S.hmin[S.uaux[q]] = q;
                    S.ucrd[q].y |= off << 16;
                }
            }
        }
        seg_sync<COOP>();
        if (listed) {
            seg_build_lists<MW, C::ROUNDS>(E, rounds, S.eslot, S.hmin, S.umask, S.ucrd, S.c0, S.prec, list,
                                           COOP ? wv : 0, COOP ? 2 : 1);
            seg_sync<COOP>();
        }
        double *xacc = (double *)(xbuf + 256);   // COOP: [slot j / 2][m][lane] of wavefront 1
        double accj[2][Bp1];                     // COOP: this wavefront's slots j = 2 jj + wv
        if constexpr (COOP) {
            if (listed) {
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = 2 * jj + wv;
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) accj[jj][m] = 0;
                    if (WAVE * j >= U) continue;
                    const int q = lane + WAVE * j;
                    int cnt = 0, off = 0;
                    if (q < U) {
#pragma unroll
                        for (int w = 0; w < MW; ++w) cnt += __popcll(S.umask[q * MW + w]);
                        off = S.ucrd[q].y >> 16;
                    }
                    fold_list<NB>(list, off, cnt, accj[jj], thr);
                    if (wv == 1) {
#pragma unroll
                        for (int m = 0; m < Bp1; ++m) xacc[(jj * Bp1 + m) * WAVE + lane] = accj[jj][m];
                    }
                }
            }
            __syncthreads();
            if (wv == 1) return;   // wavefront 0 goes on alone
        }
        if (listed) {
#pragma unroll
            for (int j = 0; j < SEG_VB; ++j) {
#pragma unroll
                for (int m = 0; m < Bp1; ++m) K.nr[j][m] = 0;
                if (WAVE * j >= U) continue;
                const int q = lane + WAVE * j;
                const bool own = q < U;
                double rho[Bp1], acc[Bp1];
#pragma unroll
                for (int m = 0; m < Bp1; ++m) {
                    rho[m] = 0;
                    acc[m] = 0;
                }
                int vown = 0, cnt = 0, off = 0;
                if (own) {
                    vown = S.uvox[q];
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) rho[m] = rho_g[(int64_t)vown * Bp1 + m];
#pragma unroll
                    for (int w = 0; w < MW; ++w) cnt += __popcll(S.umask[q * MW + w]);
                    off = S.ucrd[q].y >> 16;
                }
                if constexpr (COOP) {
                    // this voxel's fold, by wavefront 0 (even slots) or 1 (odd slots)
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) acc[m] = (j & 1) ? xacc[((j >> 1) * Bp1 + m) * WAVE + lane] : accj[j >> 1][m];
                } else {
                    fold_list<NB>(list, off, cnt, acc, thr);
                }
                if (own) {
                    voxel_dE_terms<NB>(R, vown, rho, acc, S2, SX, NS);
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) K.nr[j][m] = rho[m] + acc[m];
                }
            }
            // empty the hash again: the hmin map, the used slots' masks and the lists
            for (int q = lane; q < U; q += WAVE) {
                const int h = S.uaux[q];
                S.hmin[h] = HMIN_EMPTY;
#pragma unroll
                for (int w = 0; w < MW; ++w) S.hpos[h * MW + w] = 0;
            }
            for (int i = lane; i < E; i += WAVE) list[i] = FoldEntry{0., 0, 0};
            wave_sync();
        } else {
#pragma unroll
            for (int j = 0; j < SEG_VB; ++j) {
#pragma unroll
                for (int m = 0; m < Bp1; ++m) K.nr[j][m] = 0;
                if (WAVE * j >= U) continue;
                const int q = lane + WAVE * j;
                const bool own = q < U;
                double rho[Bp1], acc[Bp1];
#pragma unroll
                for (int m = 0; m < Bp1; ++m) {
                    rho[m] = 0;
                    acc[m] = 0;
                }
                int vx = -1, vy = -1, vz = -1;
                uint64_t hits[MW];
#pragma unroll
                for (int w = 0; w < MW; ++w) hits[w] = 0;
                int vown = 0;
                if (own) {
                    const int v = S.uvox[q];
                    vown = v;
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) rho[m] = rho_g[(int64_t)v * Bp1 + m];
                    const int2 c = S.ucrd[q];
                    vx = c.x & 0xffff;
                    vy = c.x >> 16;
                    vz = c.y;
#pragma unroll
                    for (int w = 0; w < MW; ++w) hits[w] = S.umask[q * MW + w];
                }
                // each lane visits the positions that touch its voxel, in order
                for (;;) {
                    bool more = false;
#pragma unroll
                    for (int w = 0; w < MW; ++w) more |= hits[w] != 0;
                    if (!wave_any(more)) break;
                    if (!more) continue;
                    int p = 0;
#pragma unroll
                    for (int w = MW - 1; w >= 0; --w)
                        if (hits[w] != 0) p = 64 * w + __builtin_ctzll(hits[w]);
#pragma unroll
                    for (int w = 0; w < MW; ++w)
                        if (w == (p >> 6)) hits[w] &= hits[w] - 1;
                    const int4 pr = S.prec[p];
                    const double pre = (p & 1) ? 1. : -1.;
                    const int x0 = pr.x & 0xffff, x1 = pr.x >> 16, y0 = pr.y & 0xffff, y1 = pr.y >> 16,
                              z0 = pr.z & 0xffff, z1 = pr.z >> 16;
                    if (!degen) {
                        const bool hx = vx == x0 || vx == x1, hy = vy == y0 || vy == y1, hz = vz == z0 || vz == z1;
                        if (hx && hy && hz) {
                            const int l = (vx == x1 ? 1 : 0) | (vy == y1 ? 2 : 0) | (vz == z1 ? 4 : 0);
                            fold_c0<NB>(S.c0[p * 8 + l], pre, pr.w, acc, thr);
                        }
                    } else {
                        int dx = vx - x0, dy = vy - y0, dz = vz - z0;
                        if (dx < 0) dx += nx;
                        if (dy < 0) dy += ny;
                        if (dz < 0) dz += nz;
                        if (own)
                            for (int l = 0; l < 8; ++l)
                                if (axis_hit(l & 1, dx, nx) && axis_hit((l >> 1) & 1, dy, ny) &&
                                    axis_hit((l >> 2) & 1, dz, nz))
                                    fold_c0<NB>(S.c0[p * 8 + l], pre, pr.w, acc, thr);
                    }
                }
                if (own) {
                    voxel_dE_terms<NB>(R, vown, rho, acc, S2, SX, NS);
#pragma unroll
                    for (int m = 0; m < Bp1; ++m) K.nr[j][m] = rho[m] + acc[m];
                }
            }
        }
        const double share = wave_sum(lane_field_share<NB>(R, S2, SX, NS));

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
    }

    // ---- Metropolis
    double dE = 0;
    dE += dEp;
    if (check_field) dE += dEf;
    res.dEp = dEp;
    res.dEf = dEf;
    res.expdE = exp(-dE);
    res.accepted = u < res.expdE;
    res.U = U;
}

// seg_eval_loaded with its own loads (the one-at-a-time path).
template <int NB, int SMAX = SEG_MAX, bool TW = false, bool COOP = false>
__device__ __forceinline__ void seg_eval(const chromo_replica *R, const Proposal &P, double u, const SegLds &S,
                                         SegRes &res, SegKeep<NB> &K, char *xbuf = nullptr)
{
    SegBeads<NB> B;
    seg_load<NB>(R, P.move, P.n, P.ind0, P.indf, P.mypick, B);
    seg_eval_loaded<NB, SMAX, TW, COOP>(R, P, u, S, res, K, B, xbuf);
}

// Writes an accepted proposal: bead coordinates or states (MCAdapter.accept,
// moves.pyx:190-226) and the touched voxels' densities
// (update_affected_densities, fields.pyx:1968-1975).
template <int NB>
__device__ __forceinline__ void seg_commit(const chromo_replica *R, int move, int64_t n, int binder, int mynew,
                                           const SegKeep<NB> &K, const SegLds &S, int U)
{
    constexpr int Bp1 = NB + 1;
    const int lane = lane_id();
    const int64_t N = R->N;
    if (lane < n) {
        if (move == CHROMO_CHANGE_BINDING_STATE) {
            gmem(R->states)[(int64_t)binder * N + K.b] = (int8_t)mynew;
        } else {
            const int c0 = (move == CHROMO_TANGENT_ROTATION) ? 3 : 0;
            const int c1 = (move == CHROMO_SLIDE) ? 3 : 9;
#pragma unroll
            for (int c = 0; c < 9; ++c)
                if (c >= c0 && c < c1) gmem(R->pos)[c * N + K.b] = K.w[c];
        }
    }
    GLOBAL_AS double *rho = gmem(R->density);
#pragma unroll
    for (int j = 0; j < SEG_VB; ++j) {
        const int q = lane + WAVE * j;
        if (q < U) {
            const int64_t v = S.uvox[q];
#pragma unroll
            for (int m = 0; m < Bp1; ++m) rho[v * Bp1 + m] = K.nr[j][m];
        }
    }
}

}  // namespace
