// chromo_mc.hpp -- device code of one Monte Carlo move, shared by the kernels.
//
// One wavefront runs a replica's sequential Markov chain
// (chromo/mc/mc_sim.pyx:26-182).  Moves of at most SEG_MAX beads are evaluated
// by seg_eval (chromo_seg.hpp).  Larger ones spread their 16*n voxel
// contributions (2 configurations x 8 corners per bead) over the 64 lanes and
// group them per voxel in the replica's global scratch arena through a hash
// table, ordered by contribution index and folded per voxel in the reference's
// order, so each voxel's delta density is the reference's left fold
// (chromo/fields.pyx:1491-1520) bit for bit.  Energy terms are reduced across the
// wavefront in a fixed order; no global atomics are used.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "chromo_device.hpp"

namespace {

using namespace chromo;

struct Proposal {
    int move;
    int64_t n, ind0, indf;
    double T[12];          // crank / end pivot rotation, slide translation in T[3,7,11]
    int binder;
    // per-lane registers for beads i == lane (< 64)
    int64_t mypick;        // tangent bead index
    double myT[9];         // tangent rotation
    int mynew;             // binding new state
};

}  // namespace

#include "chromo_seg.hpp"
#include "chromo_big.hpp"

namespace {

// ---------------------------------------------------------------------------
// Move arena
// ---------------------------------------------------------------------------

struct ArenaLayout {
    size_t bnew, tanT, epart, c0, udrho, urho;     // double arrays
    size_t inds;                                   // int64
    size_t evox, eslot, list, hkey, hmin, hu, uvox, uslot, ucnt, uoff, spk;   // int32
    size_t stold, stnew;                           // int8
    size_t total;
    int cap, ecap, H, hbits;
};

__host__ __device__ inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// cap = per-bead capacity (all moves), fcap = field-move capacity (<= cap): the
// per-contribution arrays hold 16 * fcap entries.
__host__ __device__ inline ArenaLayout arena_layout(int cap, int fcap, int Bp1, int H)
{
    ArenaLayout L;
    L.cap = cap;
    L.ecap = 16 * fcap;
    L.H = H;
    int hb = 0;
    while ((1 << hb) < H) ++hb;
    L.hbits = hb;
    size_t o = 0;
    L.bnew = o;  o += sizeof(double) * 9 * (size_t)cap;
    L.tanT = o;  o += sizeof(double) * 9 * (size_t)cap;
    L.epart = o; o += sizeof(double) * (size_t)cap;
    L.c0 = o;    o += sizeof(double) * (size_t)L.ecap;
    L.udrho = o; o += sizeof(double) * (size_t)Bp1 * L.ecap;
    L.urho = o;  o += sizeof(double) * (size_t)Bp1 * L.ecap;
    L.inds = o;  o += sizeof(int64_t) * (size_t)cap;
    L.evox = o;  o += sizeof(int32_t) * (size_t)L.ecap;
    L.eslot = o; o += sizeof(int32_t) * (size_t)L.ecap;
    L.list = o;  o += sizeof(int32_t) * (size_t)L.ecap;
    L.hkey = o;  o += sizeof(int32_t) * (size_t)H;
    L.hmin = o;  o += sizeof(int32_t) * (size_t)H;
    L.hu = o;    o += sizeof(int32_t) * (size_t)H;
    L.uvox = o;  o += sizeof(int32_t) * (size_t)L.ecap;
    L.uslot = o; o += sizeof(int32_t) * (size_t)L.ecap;
    L.ucnt = o;  o += sizeof(int32_t) * (size_t)L.ecap;
    L.uoff = o;  o += sizeof(int32_t) * (size_t)(L.ecap + 1);
    L.spk = o;   o += sizeof(int32_t) * 2 * (size_t)cap;
    L.stold = o; o += (size_t)CHROMO_MAX_BINDERS * cap;
    L.stnew = o; o += (size_t)CHROMO_MAX_BINDERS * cap;
    L.total = align_up(o, 256);
    return L;
}

__host__ __device__ inline int hash_size_for(int fcap)
{
    int h = 1;
    while (h < 2 * 16 * fcap) h <<= 1;
    return h;
}

__host__ __device__ inline ArenaLayout lds_layout(int Bp1)
{
    return arena_layout(SEG_LDS, SEG_LDS, Bp1, HASH_LDS);
}

__host__ __device__ inline ArenaLayout global_layout(int64_t bead_cap, int64_t seg_cap, int Bp1)
{
    int fc = seg_cap > 0 ? (int)seg_cap : 1;
    return arena_layout((int)bead_cap, fc, Bp1, hash_size_for(fc));
}

struct Arena {
    double *bnew, *tanT, *epart, *c0, *udrho, *urho;
    int64_t *inds;
    int32_t *evox, *eslot, *list, *hkey, *hmin, *hu, *uvox, *uslot, *ucnt, *uoff, *spk;
    int8_t *stold, *stnew;
    int cap, ecap, hmask, hshift;
};

__device__ inline Arena make_arena(char *base, const ArenaLayout &L)
{
    Arena A;
    A.bnew = (double *)(base + L.bnew);
    A.tanT = (double *)(base + L.tanT);
    A.epart = (double *)(base + L.epart);
    A.c0 = (double *)(base + L.c0);
    A.udrho = (double *)(base + L.udrho);
    A.urho = (double *)(base + L.urho);
    A.inds = (int64_t *)(base + L.inds);
    A.evox = (int32_t *)(base + L.evox);
    A.eslot = (int32_t *)(base + L.eslot);
    A.list = (int32_t *)(base + L.list);
    A.hkey = (int32_t *)(base + L.hkey);
    A.hmin = (int32_t *)(base + L.hmin);
    A.hu = (int32_t *)(base + L.hu);
    A.uvox = (int32_t *)(base + L.uvox);
    A.uslot = (int32_t *)(base + L.uslot);
    A.ucnt = (int32_t *)(base + L.ucnt);
    A.uoff = (int32_t *)(base + L.uoff);
    A.spk = (int32_t *)(base + L.spk);
    A.stold = (int8_t *)(base + L.stold);
    A.stnew = (int8_t *)(base + L.stnew);
    A.cap = L.cap;
    A.ecap = L.ecap;
    A.hmask = L.H - 1;
    A.hshift = 32 - L.hbits;
    return A;
}

constexpr size_t MT_LDS_BYTES = 624 * sizeof(uint32_t);

// Copies a replica record into LDS, all threads of the workgroup helping; the
// caller synchronises.  The kernels then read the record's fields from LDS: read
// from HBM through the table pointer they would be loaded again after every store
// the compiler cannot tell apart from the table.
__device__ __forceinline__ void load_record(chromo_replica *dst, const chromo_replica *src)
{
    static_assert(sizeof(chromo_replica) % 4 == 0, "word copy");
    const uint32_t *s = (const uint32_t *)src;
    uint32_t *d = (uint32_t *)dst;
    for (int i = (int)threadIdx.x; i < (int)(sizeof(chromo_replica) / 4); i += (int)blockDim.x) d[i] = s[i];
}

// Dynamic LDS of the field-dE tape kernel: the LDS move arena and an MT block.
inline size_t lds_bytes_for(int Bp1)
{
    ArenaLayout L = lds_layout(Bp1);
    return L.total + MT_LDS_BYTES;
}

// Dynamic LDS of the one-wavefront Monte Carlo kernels: one evaluator block for
// moves of up to SEG_WIDE beads (its start doubles as the SEG_MAX block), then the
// MT19937 state.
// The wide block of the one-at-a-time path: a SEG_WIDE evaluator block or two
// SEG_MAX ones, whichever is larger; the long-move evaluator uses all of it.
__host__ __device__ inline size_t wide_lds_bytes(int Bp1)
{
    const size_t a = seg_layout<SEG_WIDE>(Bp1).total, b = 2 * seg_layout<SEG_MAX>(Bp1).total;
    return a > b ? a : b;
}

inline size_t mc_lds_bytes(int Bp1) { return wide_lds_bytes(Bp1) + MT_LDS_BYTES; }

extern __shared__ char lds_dyn[];

// ---------------------------------------------------------------------------
// Per-move result
// ---------------------------------------------------------------------------

struct MoveOut {
    int64_t n, ind0, indf, nbins;
    double dEp, dEf, u, expdE;
    int accepted;
    int err;
};

// ---------------------------------------------------------------------------
// Move body: everything after the bead range is known.  KIND 0 = LDS arena,
// KIND 1 = global scratch arena.
// ---------------------------------------------------------------------------

// The arena is LDS (KIND 0) or global scratch (KIND 1); one wavefront uses it.
template <int KIND>
__device__ __forceinline__ void arena_sync()
{
    if (KIND == 0) wave_sync();
    else wave_sync_global();
}

template <int KIND, int NB, bool TW = false>
__device__ __forceinline__ void move_body(const chromo_replica *R, double u, const Proposal &P, MoveOut &out)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    constexpr int B = NB;
    constexpr int Bp1 = NB + 1;
    const int mv = P.move;
    const int64_t n = P.n;
    const int64_t ind0 = P.ind0, indf = P.indf;
    const bool check_field = R->field_on && mv != CHROMO_TANGENT_ROTATION;
    const bool state_change = (mv == CHROMO_CHANGE_BINDING_STATE);

    Arena A;
    if (KIND == 0) {
        A = make_arena(lds_dyn, lds_layout(Bp1));
    } else {
        A = make_arena((char *)R->scratch, global_layout(R->bead_cap, R->seg_cap, Bp1));
    }
    const int cap = A.cap;
    double *Pp = R->pos;

    // ---- per-bead: old and trial coordinates, states, elastic/binding partials,
    //      voxel corners of the current (k = 0) and trial (k = 1) positions.
    for (int64_t i = lane; i < n; i += WAVE) {
        int64_t b;
        if (mv == CHROMO_TANGENT_ROTATION) b = (i < WAVE) ? P.mypick : A.inds[i];
        else b = ind0 + i;
        double o[9], w[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) o[c] = Pp[c * N + b];
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
            double Tl[9];
            if (i < WAVE) {
#pragma unroll
                for (int q = 0; q < 9; ++q) Tl[q] = P.myT[q];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) Tl[q] = A.tanT[q * cap + i];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                w[j] = o[j];
                w[3 + j] = rowdot(&Tl[3 * j], o[3], o[4], o[5]);
                w[6 + j] = rowdot(&Tl[3 * j], o[6], o[7], o[8]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) w[c] = o[c];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) A.bnew[c * cap + i] = w[c];
        int8_t so[NB], sn[NB];
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            so[m] = R->states[(int64_t)m * N + b];
            sn[m] = so[m];
        }
        if (state_change) {
            int v = (i < WAVE) ? P.mynew : (int)A.stnew[P.binder * cap + i];
#pragma unroll
            for (int m = 0; m < NB; ++m)
                if (m == P.binder) sn[m] = (int8_t)v;
        }
        for (int m = 0; m < B; ++m) {
            A.stold[m * cap + i] = so[m];
            A.stnew[m * cap + i] = sn[m];
        }
        // elastic / binding partials
        if (mv == CHROMO_TANGENT_ROTATION) {
            double e = 0;
            if (b != 0) {
                double r0[3], t30[3];
                for (int j = 0; j < 3; ++j) {
                    r0[j] = Pp[j * N + b - 1];
                    t30[j] = Pp[(3 + j) * N + b - 1];
                }
                Bond bd = load_bond(R->bonds, N, b - 1);
                double twt = 0, twc = 0;
                if constexpr (TW) twist_forward(R->bonds, R->pos, N, b - 1, o, w, twt, twc);
                e += pair_dE_forward(bd, r0, &o[0], &w[0], t30, &o[3], &w[3], twt, twc);
            }
            if (b + 1 != N) {
                double r1[3], t31[3];
                for (int j = 0; j < 3; ++j) {
                    r1[j] = Pp[j * N + b + 1];
                    t31[j] = Pp[(3 + j) * N + b + 1];
                }
                Bond bd = load_bond(R->bonds, N, b);
                double twt = 0, twc = 0;
                if constexpr (TW) twist_reverse(R->bonds, R->pos, N, b + 1, o, w, twt, twc);
                e += pair_dE_reverse(bd, &o[0], &w[0], r1, &o[3], &w[3], t31, twt, twc);
            }
            A.epart[i] = e;
        } else if (mv == CHROMO_CHANGE_BINDING_STATE) {
            int8_t mods[CHROMO_MAX_BINDERS];
            for (int m = 0; m < B; ++m) mods[m] = R->marks[(int64_t)m * N + b];
            A.epart[i] = bead_binding_dE<NB>(R, mods, so, sn);
        } else {
            if (i == 0 && ind0 != 0) {
                double r0[3], t30[3];
                for (int j = 0; j < 3; ++j) {
                    r0[j] = Pp[j * N + ind0 - 1];
                    t30[j] = Pp[(3 + j) * N + ind0 - 1];
                }
                Bond bd = load_bond(R->bonds, N, ind0 - 1);
                double twt = 0, twc = 0;
                if constexpr (TW) twist_forward(R->bonds, R->pos, N, ind0 - 1, o, w, twt, twc);
                A.epart[0] = pair_dE_forward(bd, r0, &o[0], &w[0], t30, &o[3], &w[3], twt, twc);
            }
            if (i == n - 1 && indf != N) {
                double r1[3], t31[3];
                for (int j = 0; j < 3; ++j) {
                    r1[j] = Pp[j * N + indf];
                    t31[j] = Pp[(3 + j) * N + indf];
                }
                Bond bd = load_bond(R->bonds, N, indf - 1);
                double twt = 0, twc = 0;
                if constexpr (TW) twist_reverse(R->bonds, R->pos, N, indf, o, w, twt, twc);
                A.epart[1] = pair_dE_reverse(bd, &o[0], &w[0], r1, &o[3], &w[3], t31, twt, twc);
            }
        }
        if (check_field) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double pos[3];
                for (int j = 0; j < 3; ++j) pos[j] = (k == 0 || state_change) ? o[j] : w[j];
                int64_t ix[3];
                double wt[3];
                voxel_lower_1(R, pos, ix, wt);
                const int p = (int)(2 * i + k);
                const double vol_bin = R->vol_bin;
                // binder states of this configuration, one byte per binder
                int32_t sp = 0;
#pragma unroll
                for (int m = 0; m < NB; ++m)
                    sp |= ((int32_t)(uint8_t)((k == 0 || !state_change) ? so[m] : sn[m])) << (8 * m);
                A.spk[p] = sp;
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    const int64_t cv = corner_voxel(R, ix, l);
                    A.evox[p * 8 + l] = (int32_t)cv;
                    // weight / access volume of this contribution (fields.pyx:1472-1488)
                    A.c0[p * 8 + l] = corner_weight(l, wt[0], wt[1], wt[2]) / (R->access_vol ? R->access_vol[cv] : vol_bin);
                }
            }
        }
    }
    arena_sync<KIND>();

    // ---- polymer dE (SSWLC.compute_dE, polymers.pyx:1024-1082), uniform
    double dEp = 0;
    if (mv == CHROMO_TANGENT_ROTATION || mv == CHROMO_CHANGE_BINDING_STATE) {
        for (int64_t i = 0; i < n; ++i) dEp += A.epart[i];
    } else {
        if (ind0 != 0) dEp += A.epart[0];
        if (indf != N) dEp += A.epart[1];
    }

    // ---- field dE (UniformDensityField.compute_dE, fields.pyx:1149-1233)
    double dEf = 0;
    int U = 0;
    if (check_field) {
        const int E = (int)(16 * n);
        // unique voxels: hash insert, first-occurrence index per voxel.  The table
        // holds at least twice the contributions, so a probe that visits every
        // slot means it was not cleared: ERR_ARENA.
        int bad = 0;
        for (int e = lane; e < E; e += WAVE) {
            int v = A.evox[e];
            uint32_t h = ((uint32_t)v * 2654435761u) >> A.hshift;
            int probes = 0;
            for (;;) {
                int old = atomicCAS(&A.hkey[h], HKEY_EMPTY, v);
                if (old == HKEY_EMPTY || old == v) break;
                h = (h + 1) & (uint32_t)A.hmask;
                if (++probes > A.hmask) {
                    bad = 1;
                    break;
                }
            }
            if (bad) break;
            atomicMin(&A.hmin[h], e);
            A.eslot[e] = (int)h;
        }
        if (wave_any(bad)) {
            out.err = ERR_ARENA;
            return;
        }
        arena_sync<KIND>();
        // number unique voxels in first-occurrence order
        const int C = (E + WAVE - 1) / WAVE;
        const int e0 = lane * C;
        const int e1 = min(E, e0 + C);
        int cnt = 0;
        for (int e = e0; e < e1; ++e)
            if (A.hmin[A.eslot[e]] == e) ++cnt;
        int base = wave_excl_scan(cnt, U);
        if (U < 1 || U > E) {
            out.err = ERR_ARENA;
            return;
        }
        for (int e = e0; e < e1; ++e) {
            int s = A.eslot[e];
            if (A.hmin[s] == e) {
                A.uvox[base] = A.evox[e];
                A.uslot[base] = s;
                A.hu[s] = base;
                A.ucnt[base] = 0;
                ++base;
            }
        }
        arena_sync<KIND>();
        // gather current densities of the touched voxels
        const double *rho = R->density;
        for (int q = lane; q < U; q += WAVE) {
            int64_t v = A.uvox[q];
            for (int m = 0; m < Bp1; ++m) A.urho[m * A.ecap + q] = rho[v * Bp1 + m];
        }
        for (int e = lane; e < E; e += WAVE) atomicAdd(&A.ucnt[A.hu[A.eslot[e]]], 1);
        arena_sync<KIND>();
        {
            const int Cu = (U + WAVE - 1) / WAVE;
            const int q0 = lane * Cu, q1 = min(U, q0 + Cu);
            int s = 0;
            for (int q = q0; q < q1; ++q) s += A.ucnt[q];
            int tot;
            int off = wave_excl_scan(s, tot);
            for (int q = q0; q < q1; ++q) {
                A.uoff[q] = off;
                off += A.ucnt[q];
                A.ucnt[q] = 0;
            }
            if (lane == 0) A.uoff[U] = E;
        }
        arena_sync<KIND>();
        // counting-sort placement, filling each voxel's list from the front so that
        // it comes out (nearly) in contribution order; the insertion sort below
        // makes it exact
        int outside = 0;
        for (int e = lane; e < E; e += WAVE) {
            int q = A.hu[A.eslot[e]];
            int pos = (q >= 0 && q < U) ? A.uoff[q] + atomicAdd(&A.ucnt[q], 1) : -1;
            if (pos >= 0 && pos < E) A.list[pos] = e;
            else outside = 1;
        }
        if (wave_any(outside)) {
            out.err = ERR_ARENA;
            return;
        }
        arena_sync<KIND>();
        // per voxel: order contributions, fold (fields.pyx:1491-1520), energy terms.
        // Lane l handles the voxels whose first contribution e has e % 64 == l, in
        // increasing e: the same partial sums as the register path below.
        double S2[NB], SX[NB * NB], NS = 0;
        for (int j = 0; j < NB; ++j) S2[j] = 0;
        for (int j = 0; j < NB * NB; ++j) SX[j] = 0;
        const bool thr = R->fast_field == 0;
        for (int e = lane; e < E; e += WAVE) {
            const int hs = A.eslot[e];
            if (A.hmin[hs] != e) continue;
            const int q = A.hu[hs];
            const int s0 = A.uoff[q], s1 = A.uoff[q + 1];
            for (int a = s0 + 1; a < s1; ++a) {
                int x = A.list[a];
                int c = a - 1;
                while (c >= s0 && A.list[c] > x) {
                    A.list[c + 1] = A.list[c];
                    --c;
                }
                A.list[c + 1] = x;
            }
            double acc[NB + 1];
#pragma unroll
            for (int m = 0; m <= NB; ++m) acc[m] = 0;
            // items in groups of four: their loads are issued together, the adds
            // then run in list order
            for (int a = s0; a < s1; a += 4) {
                int ev[4];
                double cv[4];
                int32_t sv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) ev[u] = (a + u < s1) ? A.list[a + u] : 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cv[u] = (a + u < s1) ? A.c0[ev[u]] : 0.0;
                    sv[u] = (a + u < s1) ? A.spk[ev[u] >> 3] : 0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (a + u >= s1) break;
                    const double c0 = cv[u];
                    const double pre = ((ev[u] >> 3) & 1) ? 1. : -1.;
                    const double t0 = pre * c0;
                    if (!thr || fabs(t0) > 1E-18) acc[0] += t0;
#pragma unroll
                    for (int m = 1; m <= NB; ++m) {
                        const int st = (sv[u] >> (8 * (m - 1))) & 0xff;
                        const double t = pre * (c0 * (double)st);
                        if (!thr || fabs(t) > 1E-18) acc[m] += t;
                    }
                }
            }
            double r[CHROMO_MAX_BINDERS + 1];
            for (int m = 0; m < Bp1; ++m) {
                A.udrho[m * A.ecap + q] = acc[m];
                r[m] = A.urho[m * A.ecap + q];
            }
            voxel_dE_terms<NB>(R, A.uvox[q], r, acc, S2, SX, NS);
        }
        const double share = wave_sum(lane_field_share<NB>(R, S2, SX, NS));
        // count_doubly_bound (fields.pyx:1877-1937) and confinement (:121-202)
        int dbl_t[CHROMO_MAX_BINDERS], dbl[CHROMO_MAX_BINDERS];
        for (int j = 0; j < B; ++j) {
            dbl_t[j] = 0;
            dbl[j] = 0;
        }
        int out_t = 0, out_c = 0;
        for (int64_t i = lane; i < n; i += WAVE) {
            for (int j = 0; j < B; ++j) {
                int so = A.stold[j * cap + i];
                int stt = state_change ? A.stnew[j * cap + i] : so;
                dbl_t[j] += (stt == 2);
                dbl[j] += (so == 2);
            }
            if (R->confine_type != 0 && !state_change) {
                int64_t b = ind0 + i;
                double rc[3], rt[3];
                for (int j = 0; j < 3; ++j) {
                    rc[j] = Pp[j * N + b];
                    rt[j] = A.bnew[j * cap + i];
                }
                if (R->confine_type == 1) {
                    out_t += sqrt(dot3(rt, rt)) > R->confine_length;
                    out_c += sqrt(dot3(rc, rc)) > R->confine_length;
                } else {
                    for (int j = 0; j < 3; ++j) out_t += fabs(rt[j]) > R->confine_length / 2;
                }
            }
        }
        for (int j = 0; j < B; ++j) {
            dbl_t[j] = bcast0(wave_sum_int(dbl_t[j]));
            dbl[j] = bcast0(wave_sum_int(dbl[j]));
        }
        out_t = bcast0(wave_sum_int(out_t));
        out_c = bcast0(wave_sum_int(out_c));
        dEf = field_dE_total<NB>(R, share, dbl_t, dbl, out_t, out_c, state_change);
    }

    // ---- Metropolis (mc_sim.pyx:152-182)
    double dE = 0;
    dE += dEp;
    if (check_field) dE += dEf;
    double expdE = exp(-dE);
    int acc = u < expdE;

    // ---- commit (MCAdapter.accept, moves.pyx:190-226; update_affected_densities,
    //      fields.pyx:1968-1975)
    if (acc) {
        for (int64_t i = lane; i < n; i += WAVE) {
            int64_t b;
            if (mv == CHROMO_TANGENT_ROTATION) b = (i < WAVE) ? P.mypick : A.inds[i];
            else b = ind0 + i;
            if (mv == CHROMO_CHANGE_BINDING_STATE) {
                R->states[(int64_t)P.binder * N + b] = A.stnew[P.binder * cap + i];
            } else {
                int c0 = (mv == CHROMO_TANGENT_ROTATION) ? 3 : 0;
                int c1 = (mv == CHROMO_SLIDE) ? 3 : 9;
                for (int c = c0; c < c1; ++c) Pp[c * N + b] = A.bnew[c * cap + i];
            }
        }
        if (check_field) {
            double *rho = R->density;
            for (int q = lane; q < U; q += WAVE) {
                int64_t v = A.uvox[q];
                for (int m = 0; m < Bp1; ++m)
                    rho[v * Bp1 + m] = A.urho[m * A.ecap + q] + A.udrho[m * A.ecap + q];
            }
        }
    }
    // release the hash slots used by this move
    if (check_field) {
        for (int q = lane; q < U; q += WAVE) {
            int s = A.uslot[q];
            A.hkey[s] = HKEY_EMPTY;
            A.hmin[s] = HMIN_EMPTY;
        }
    }
    wave_sync_global();
    out.dEp = dEp;
    out.dEf = dEf;
    out.u = u;
    out.expdE = expdE;
    out.accepted = acc;
    out.nbins = U;
}

// ---------------------------------------------------------------------------
// mc_step: proposal (move_funcs.pyx), then move_body.  Wave-uniform until the
// per-bead phase.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void load_pos3(const double *Pp_, int64_t N, int64_t b, double v[3])
{
    const GLOBAL_AS double *Pp = gmem(Pp_);
    v[0] = Pp[b];
    v[1] = Pp[N + b];
    v[2] = Pp[2 * N + b];
}

// mc_step's proposal: the move's draws and their arithmetic, the capacity checks
// and the Metropolis draw u (mc_sim.pyx:106-171).  Returns true when the
// proposal goes on to its evaluation (n > 0, no error); out then holds n, ind0,
// indf, and out.err any error.
template <int NB>
__device__ __forceinline__ bool mc_propose(const chromo_replica *R, WaveRng &rng, chromo_ctrl &C, MoveOut &out,
                                           Proposal &P, double &u)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    const double *Pp = R->pos;
    P.move = C.move;
    P.binder = 0;
    P.mypick = 0;
    P.mynew = 0;
    for (int q = 0; q < 9; ++q) P.myT[q] = 0;
    int err = 0;
    out.err = 0;
    out.n = 0;
    C.num_attempt += 1;   // MCAdapter.propose, moves.pyx:151
    const double amp_move = C.amp_move;
    const int64_t amp_bead = C.amp_bead;
    int64_t n = 0, ind0 = 0, indf = 0;

    if (P.move == CHROMO_CRANK_SHAFT) {
        // crank_shaft, move_funcs.pyx:39-101
        double ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t b0 = (int64_t)(int32_t)(((double)rnd31(rng) / RMAX) * (double)N);
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (err) goto fail;
        b1 = imax64(b1, 1);
        check_bead_bounds(b0, b1, N, ind0, indf);
        n = indf - ind0;
        // get_crank_shaft_axis, move_funcs.pyx:157-234
        int64_t ia, ib;
        const int64_t indf_m1 = indf - 1, ind0_m1 = ind0 - 1;
        if (ind0 == indf_m1 && ind0 == 0) { ia = indf; ib = ind0; }
        else if (ind0 == indf_m1 && ind0 == N - 1) { ia = ind0; ib = ind0_m1; }
        else if (ind0 == 0 && indf == N) { ia = indf_m1; ib = ind0; }
        else if (ind0 == 0) { ia = indf; ib = ind0; }
        else if (indf == N) { ia = indf_m1; ib = ind0_m1; }
        else { ia = indf; ib = ind0_m1; }
        double ra[3], rb[3], dir[3];
        load_pos3(Pp, N, ia, ra);
        load_pos3(Pp, N, ib, rb);
        for (int i = 0; i < 3; ++i) dir[i] = ra[i] - rb[i];
        double mag = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        if (mag < 1E-5) {
            sample_unit_sphere(rng, dir);
        } else {
            double sc = 1.0 / mag;
            for (int i = 0; i < 3; ++i) dir[i] = dir[i] * sc;
        }
        // get_crank_shaft_fulcrum, move_funcs.pyx:237-280
        int64_t ifc;
        if (ind0 == 0 && indf != N) ifc = indf;
        else if (ind0 != 0 && indf == N) ifc = ind0_m1;
        else if (ind0 == 0 && indf == N) ifc = ind0;
        else ifc = ind0_m1;
        double pt[3];
        load_pos3(Pp, N, ifc, pt);
        axis_rotation(dir, pt, ang, P.T);
    } else if (P.move == CHROMO_END_PIVOT) {
        // end_pivot, move_funcs.pyx:285-346
        double ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t lhs = rnd31(rng) & 1;
        if (lhs == 1) {
            ind0 = 0;
            indf = from_left(rng, amp_bead, N, err) + 1;
        } else {
            ind0 = from_right(rng, amp_bead, N, err);
            indf = N;
        }
        // a window of the whole chain would select bead N (an IndexError in the reference)
        if (!err && indf > N) err = ERR_WINDOW;
        if (err) goto fail;
        n = indf - ind0;
        double axis[3], sa, ca;
        {
            const int32_t d1 = rnd31(rng);
            const int32_t d2 = rnd31(rng);
            sphere_from_uni(d1, d2, ang, axis, sa, ca);
        }
        // get_end_pivot_fulcrum, move_funcs.pyx:349-398
        int64_t ifc;
        if (ind0 == 0 && indf != N) ifc = indf;
        else if (ind0 != 0 && indf == N) ifc = ind0 - 1;
        else if (ind0 == 0 && indf == N && lhs == 1) ifc = indf - 1;
        else ifc = ind0;
        double pt[3];
        load_pos3(Pp, N, ifc, pt);
        axis_rotation_cs(axis, pt, ca, sa, P.T);
    } else if (P.move == CHROMO_SLIDE) {
        // slide, move_funcs.pyx:403-465
        double amp = amp_move * ((double)rnd31(rng) / RMAX);
        double dir[3];
        sample_unit_sphere(rng, dir);
        for (int i = 0; i < 3; ++i) dir[i] *= amp;
        P.T[3] = dir[0];
        P.T[7] = dir[1];
        P.T[11] = dir[2];
        int64_t b0 = draw_mod(rnd31(rng), N);
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (err) goto fail;
        check_bead_bounds(b0, b1, N, ind0, indf);
        n = indf - ind0;
    } else if (P.move == CHROMO_TANGENT_ROTATION) {
        // tangent_rotation + get_inds + rotate_select_beads, move_funcs.pyx:470-582
        double ang = amp_move * (((double)rnd31(rng) / RMAX) - 0.5);
        int64_t k = draw_mod(rnd31(rng), amp_bead) + 1;
        if (k > N) {
            err = ERR_TANGENT_K;
            goto fail;
        }
        if (k > SEG_MAX && k > R->bead_cap) {
            err = ERR_SEGCAP;
            goto fail;
        }
        n = k;
        const double origin[3] = {0., 0., 0.};
        int64_t fpick = 0;
        int32_t fd1 = 0, fd2 = 0;
        if (tangent_draws_at_once(rng, k, N, fpick, fd1, fd2)) {
            // every lane its own bead and rotation
            if (lane < k) {
                double dir[3], T[12];
                sphere_from(fd1, fd2, dir);
                axis_rotation(dir, origin, ang, T);
                P.mypick = fpick;
                for (int j = 0; j < 3; ++j)
                    for (int q = 0; q < 3; ++q) P.myT[3 * j + q] = T[4 * j + q];
            }
            ind0 = 0;
            indf = 0;
            goto picked;
        }
        {
        // beads past the first 64 keep their picks and rotations in the global arena
        Arena A = make_arena((char *)R->scratch, global_layout(R->bead_cap, R->seg_cap, NB + 1));
        for (int64_t i = 0; i < k && !err; ++i) {
            for (int it = 0;; ++it) {
                if (it >= SPIN_LIMIT) {
                    err = ERR_SPIN;
                    break;
                }
                int64_t c = draw_mod(rnd31(rng), N);
                int dup = (lane < i && P.mypick == c);
                for (int64_t j = WAVE + lane; j < i; j += WAVE) dup |= (A.inds[j] == c);
                if (!wave_any(dup)) {
                    if (i < WAVE) {
                        if (lane == i) P.mypick = c;
                    } else {
                        if (lane == 0) A.inds[i] = c;
                        wave_sync_global();
                    }
                    break;
                }
            }
        }
        if (err) goto fail;
        for (int64_t i = 0; i < k; ++i) {
            double dir[3], T[12], sa, ca;
            const int32_t d1 = rnd31(rng);
            const int32_t d2 = rnd31(rng);
            sphere_from_uni(d1, d2, ang, dir, sa, ca);
            axis_rotation_cs(dir, origin, ca, sa, T);
            if (i < WAVE) {
                if (lane == i) {
                    for (int j = 0; j < 3; ++j)
                        for (int q = 0; q < 3; ++q) P.myT[3 * j + q] = T[4 * j + q];
                }
            } else if (lane == 0) {
                for (int j = 0; j < 3; ++j)
                    for (int q = 0; q < 3; ++q) A.tanT[(3 * j + q) * A.cap + i] = T[4 * j + q];
            }
        }
        wave_sync_global();
        ind0 = 0;
        indf = 0;
        }
    picked:;
    } else {
        // change_binding_state, move_funcs.pyx:717-820
        if (R->B < 1) {
            err = ERR_NBINDERS;
            goto fail;
        }
        int64_t binder = rnd31(rng) % NB;
        int64_t ntails = R->sites[binder];
        int64_t b0 = draw_mod(rnd31(rng), N);
        int64_t b1 = from_point(rng, amp_bead, N, b0, err);
        if (err) goto fail;
        check_bead_bounds(b0, b1, N, ind0, indf);
        n = indf - ind0;
        P.binder = (int)binder;
        if (n > SEG_MAX && (n > R->seg_cap || n > R->bead_cap)) {
            err = ERR_SEGCAP;
            goto fail;
        }
        Arena A = make_arena((char *)R->scratch, global_layout(R->bead_cap, R->seg_cap, NB + 1));
        for (int64_t i = 0; i < n; ++i) {
            int st = (int)rnd_randint(rng, ntails + 1);
            if (i < WAVE) {
                if (lane == i) P.mynew = st;
            } else if (lane == 0) {
                A.stnew[binder * A.cap + i] = (int8_t)st;
            }
        }
        wave_sync_global();
    }
    P.n = n;
    P.ind0 = ind0;
    P.indf = indf;
    out.n = n;
    out.ind0 = ind0;
    out.indf = indf;
    if (n == 0) return false;   // mc_sim.pyx:149-151: no dE, no Metropolis draw
    if (P.move == CHROMO_TANGENT_ROTATION) {
        // report the first picked bead
        out.ind0 = readlane_i64(P.mypick, 0);
        out.indf = out.ind0 + 1;
    }
    if (n > SEG_MAX) {
        const bool uses_field = R->field_on && P.move != CHROMO_TANGENT_ROTATION;
        if (n > R->bead_cap || (uses_field && n > R->seg_cap)) {
            err = ERR_SEGCAP;
            goto fail;
        }
    }
    // the Metropolis draw follows the proposal's draws (mc_sim.pyx:171)
    u = rnd_unit(rng);
    return true;
fail:
    out.err = err;
    out.n = 0;
    return false;
}

// AcceptanceTracker.update_acceptance_rate (mc_stat.py:190-208)
__device__ __forceinline__ void track_acceptance(chromo_ctrl &C, bool accepted)
{
    if (accepted) {
        C.num_success += 1;
        C.rate = (C.alpha * 1.0) + (1 - C.alpha) * C.rate;
    } else {
        C.rate = (C.alpha * 0.0) + (1 - C.alpha) * C.rate;
    }
}

// mc_step's evaluation of proposal P with Metropolis draw u: the evaluator that
// holds it, the commit of an accepted move, the acceptance tracker.  G: the
// caller's SEG_MAX evaluator block.  wide: the start of an LDS region that holds a
// SEG_WIDE block (nullptr if none) and overlaps nstd SEG_MAX blocks, whose hashes
// are emptied again after it is used.
template <int NB, bool TW = false>
__device__ __forceinline__ void mc_evaluate(const chromo_replica *R, chromo_ctrl &C, MoveOut &out, const Proposal &P,
                                            double u, const SegLds &G, char *wide, int nstd)
{
    const int64_t n = P.n;
    {
        SegRes sr;
        sr.overflow = 1;
        if (n <= SEG_MAX) {
            SegKeep<NB> K;
            seg_eval<NB, SEG_MAX, TW>(R, P, u, G, sr, K);
            if (!sr.overflow) {
                if (sr.accepted) seg_commit<NB>(R, P.move, n, P.binder, P.mynew, K, G, sr.U);
                wave_sync_global();
            }
        } else if (n <= SEG_WIDE && wide != nullptr) {
            const SegLds GW = seg_lds<SEG_WIDE>(wide, NB + 1);
            seg_clear<SEG_WIDE>(GW);
            SegKeep<NB> K;
            seg_eval<NB, SEG_WIDE, TW>(R, P, u, GW, sr, K);
            if (!sr.overflow) {
                if (sr.accepted) seg_commit<NB>(R, P.move, n, P.binder, P.mynew, K, GW, sr.U);
                wave_sync_global();
            }
            // the wide block covered the SEG_MAX blocks: their hashes must be empty again
            for (int b = 0; b < nstd; ++b)
                seg_clear<SEG_MAX>(seg_lds<SEG_MAX>(wide + (size_t)b * seg_layout<SEG_MAX>(NB + 1).total, NB + 1));
        } else if (wide != nullptr && big_move(P.move)) {
            const BigLds GB = big_lds(wide, NB + 1, wide_lds_bytes(NB + 1));
            if (GB.ucap > 0) {
                big_eval<NB, TW>(R, P, u, GB, sr);
                if (!sr.overflow) {
                    if (sr.accepted) big_commit<NB>(R, P, GB, sr.U);
                    wave_sync_global();
                }
                for (int b = 0; b < nstd; ++b)
                    seg_clear<SEG_MAX>(seg_lds<SEG_MAX>(wide + (size_t)b * seg_layout<SEG_MAX>(NB + 1).total, NB + 1));
            }
        }
        if (!sr.overflow) {
            out.dEp = sr.dEp;
            out.dEf = sr.dEf;
            out.u = u;
            out.expdE = sr.expdE;
            out.accepted = sr.accepted;
            out.nbins = sr.U;
        }
        // larger moves, and the rare small one with too many voxels for seg_eval
        if (sr.overflow) {
            const bool uses_field = R->field_on && P.move != CHROMO_TANGENT_ROTATION;
            if (n > R->bead_cap || (uses_field && n > R->seg_cap)) {
                out.err = ERR_SEGCAP;
                out.n = 0;
                return;
            }
            move_body<1, NB, TW>(R, u, P, out);
        }
    }
    if (out.err) {
        out.n = 0;
        return;
    }
    track_acceptance(C, out.accepted);
}

// mc_step (mc_sim.pyx:106-182) on one wavefront: mc_propose, then mc_evaluate.
template <int NB, bool TW = false>
__device__ __forceinline__ void mc_step_dev(const chromo_replica *R, WaveRng &rng, chromo_ctrl &C, MoveOut &out,
                                            const SegLds &G, char *wide, int nstd)
{
    Proposal P;
    double u = 0;
    if (mc_propose<NB>(R, rng, C, out, P, u)) mc_evaluate<NB, TW>(R, C, out, P, u, G, wide, nstd);
}

// SimpleControl.update_amplitudes (mc_controller.py:126-213)
__device__ __forceinline__ void update_amplitudes_dev(chromo_ctrl &c)
{
    if (!c.controlled) return;
    double acc = c.rate;
    if (acc < 0.5) {
        double prop = c.amp_move * 0.95;
        if (prop > c.move_lo) {
            c.amp_move = prop;
        } else {
            c.amp_move = c.move_hi;
            double nb = (double)(c.amp_bead - 1);
            c.amp_bead = (int64_t)(c.bead_lo >= nb ? c.bead_lo : nb);
        }
    } else if (acc > 0.5) {
        double prop = c.amp_move / 0.95;
        if (prop < c.move_hi) {
            c.amp_move = prop;
        } else {
            c.amp_move = c.move_lo;
            double nb = (double)(c.amp_bead + 1);
            c.amp_bead = (int64_t)(c.bead_hi <= nb ? c.bead_hi : nb);
        }
    }
}

// Algorithmic HBM bytes of one evaluated proposal.
__device__ __forceinline__ int64_t move_bytes(int mv, int64_t n, int64_t U, int B, int acc)
{
    const int64_t Bp1 = B + 1;
    if (mv == CHROMO_TANGENT_ROTATION) return 48 * n + 48 * 2 * n + (acc ? 48 * n : 0);
    if (mv == CHROMO_CHANGE_BINDING_STATE)
        return 2 * B + 24 + 8 * Bp1 * 8 + (acc ? (B + 8 * Bp1 * 8) : 0);
    return 72 * n + 48 * 2 + B * n + 8 * Bp1 * U + (acc ? (72 * n + 8 * Bp1 * U) : 0);
}

// ---------------------------------------------------------------------------
// Replica-level state load / store
// ---------------------------------------------------------------------------

__device__ __forceinline__ void rng_load(const chromo_replica *R, WaveRng &rng, uint32_t *mt_lds, bool need_mt)
{
    const int lane = lane_id();
    const uint32_t *w = R->rng;
    // every field but the glibc words is wave-uniform: scalar registers
    rng.mode = uni_i32(R->rng_mode);
    rng.g = lane < 31 ? w[CHROMO_RNG_GLIBC_STATE + lane] : 0u;
    rng.gf = uni_i32((int)w[CHROMO_RNG_GLIBC_F]);
    rng.gr = uni_i32((int)w[CHROMO_RNG_GLIBC_R]);
    rng.mt = mt_lds;
    rng.mti = uni_i32((int)w[CHROMO_RNG_MT_INDEX]);
    if (need_mt && rng.mode == CHROMO_RNG_GLIBC) {
        for (int i = lane; i < 624; i += WAVE) mt_lds[i] = w[CHROMO_RNG_MT_STATE + i];
    }
    rng.k0 = uni_u32(w[CHROMO_RNG_PHILOX_KEY]);
    rng.k1 = uni_u32(w[CHROMO_RNG_PHILOX_KEY + 1]);
    rng.pbase = PHILOX_NONE;
    rng.pidx = 0;
    if (rng.mode == CHROMO_RNG_PHILOX) {
        philox_reset(rng);
        philox_set(rng, (uint64_t)w[CHROMO_RNG_PHILOX_CTR] | ((uint64_t)w[CHROMO_RNG_PHILOX_CTR + 1] << 32),
                   (int)w[CHROMO_RNG_PHILOX_IDX]);
    }
    wave_sync();
}

__device__ __forceinline__ void rng_store(const chromo_replica *R, const WaveRng &rng, bool need_mt)
{
    const int lane = lane_id();
    uint32_t *w = R->rng;
    if (lane < 31) w[CHROMO_RNG_GLIBC_STATE + lane] = rng.g;
    if (need_mt && rng.mode == CHROMO_RNG_GLIBC) {
        wave_sync();
        for (int i = lane; i < 624; i += WAVE) w[CHROMO_RNG_MT_STATE + i] = rng.mt[i];
    }
    uint64_t ctr = 0;
    int bi = 4;
    uint32_t b[4] = {0, 0, 0, 0};
    const bool philox = rng.mode == CHROMO_RNG_PHILOX;
    if (philox) philox_get(rng, ctr, bi, b);
    if (lane == 0) {
        w[CHROMO_RNG_GLIBC_F] = (uint32_t)rng.gf;
        w[CHROMO_RNG_GLIBC_R] = (uint32_t)rng.gr;
        w[CHROMO_RNG_MT_INDEX] = (uint32_t)rng.mti;
        if (philox) {
            w[CHROMO_RNG_PHILOX_CTR] = (uint32_t)ctr;
            w[CHROMO_RNG_PHILOX_CTR + 1] = (uint32_t)(ctr >> 32);
            for (int j = 0; j < 4; ++j) w[CHROMO_RNG_PHILOX_BUF + j] = b[j];
            w[CHROMO_RNG_PHILOX_IDX] = (uint32_t)bi;
        }
    }
}

__device__ __forceinline__ void lds_hash_clear(int Bp1)
{
    ArenaLayout L = lds_layout(Bp1);
    int32_t *hk = (int32_t *)(lds_dyn + L.hkey);
    int32_t *hm = (int32_t *)(lds_dyn + L.hmin);
    for (int i = lane_id(); i < HASH_LDS; i += WAVE) {
        hk[i] = HKEY_EMPTY;
        hm[i] = HMIN_EMPTY;
    }
    wave_sync();
}

struct Counters {
    int64_t prop, acc, sumn, sumb, ties, bytes, logged;
};

__device__ __forceinline__ void record(const chromo_replica *R, const chromo_ctrl &C, const MoveOut &o, Counters &K)
{
    K.prop += 1;
    K.acc += o.accepted;
    K.sumn += o.n;
    K.sumb += o.nbins;
    K.ties += fabs(o.u - o.expdE) < 1E-12;
    K.bytes += move_bytes(C.move, o.n, o.nbins, R->B, o.accepted);
    if (R->log != nullptr && K.logged < R->log_cap) {
        if (lane_id() == 0) {
            chromo_move_log &L = R->log[K.logged];
            L.move = C.move;
            L.accepted = o.accepted;
            L.n = o.n;
            L.ind0 = o.ind0;
            L.indf = o.indf;
            L.dE_poly = o.dEp;
            L.dE_field = o.dEf;
            L.u = o.u;
            L.exp_dE = o.expdE;
            L.n_bins = o.nbins;
        }
        K.logged += 1;
    }
}

__device__ __forceinline__ void counters_load(const chromo_replica *R, Counters &K)
{
    const int64_t *c = R->counters;
    K.prop = c[CHROMO_CNT_PROPOSALS];
    K.acc = c[CHROMO_CNT_ACCEPTED];
    K.sumn = c[CHROMO_CNT_SUM_N];
    K.sumb = c[CHROMO_CNT_SUM_BINS];
    K.ties = c[CHROMO_CNT_NEAR_TIES];
    K.bytes = c[CHROMO_CNT_BYTES];
    K.logged = c[CHROMO_CNT_LOGGED];
}

__device__ __forceinline__ void counters_store(const chromo_replica *R, const Counters &K, int err)
{
    if (lane_id() != 0) return;
    int64_t *c = R->counters;
    c[CHROMO_CNT_PROPOSALS] = K.prop;
    c[CHROMO_CNT_ACCEPTED] = K.acc;
    c[CHROMO_CNT_SUM_N] = K.sumn;
    c[CHROMO_CNT_SUM_BINS] = K.sumb;
    c[CHROMO_CNT_NEAR_TIES] = K.ties;
    c[CHROMO_CNT_BYTES] = K.bytes;
    c[CHROMO_CNT_LOGGED] = K.logged;
    if (err && c[CHROMO_CNT_ERROR] == 0) c[CHROMO_CNT_ERROR] = err;
}

__device__ __forceinline__ void counters_error(const chromo_replica *R, int err)
{
    if (lane_id() == 0 && R->counters[CHROMO_CNT_ERROR] == 0) R->counters[CHROMO_CNT_ERROR] = err;
}

__device__ __forceinline__ bool any_binding(const chromo_replica *R)
{
    for (int i = 0; i < R->n_ctrl; ++i)
        if (R->ctrl[i].move == CHROMO_CHANGE_BINDING_STATE && R->ctrl[i].on == 1) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <int NB, bool TW = false>
__device__ __forceinline__ void run_steps(const chromo_replica *R, WaveRng &rng, Counters &K, chromo_ctrl *sctrl,
                          int nc, int64_t steps, int &err, const SegLds &G, char *wide)
{
    for (int64_t s = 0; s < steps && !err; ++s) {
        for (int ci = 0; ci < nc && !err; ++ci) {
            chromo_ctrl c = sctrl[ci];
            if (c.on == 1) {
                for (int j = 0; j < c.num_per_cycle; ++j) {
                    MoveOut o;
                    mc_step_dev<NB, TW>(R, rng, c, o, G, wide, 1);
                    if (o.err) {
                        err = o.err;
                        break;
                    }
                    if (o.n > 0) record(R, c, o, K);
                }
            }
            if (!err) update_amplitudes_dev(c);
            __syncthreads();
            if (lane_id() == 0) sctrl[ci] = c;
            __syncthreads();
        }
    }
}

}  // namespace
