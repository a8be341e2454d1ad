// chromo_hip.hip -- MI355X (gfx950) kernels and C ABI for chromo's Monte Carlo inner loop:
// the one-wavefront persistent and one-move-per-launch kernels, RNG seeding,
// density initialisation, the field-dE tape and the host entry points.  The
// speculative multi-wavefront kernel is in chromo_spec.hip.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "chromo_density.hpp"
#include "chromo_mc.hpp"

using namespace chromo;

namespace chromo {
hipError_t spec_launch_w2(const chromo_replica *, int, int64_t, int, hipStream_t);
}  // namespace chromo

namespace {

// Persistent mc_sim: one workgroup (one wave) per replica.
__global__ void __launch_bounds__(64) k_mc_run(const chromo_replica *__restrict__ reps, int n_reps, int64_t steps,
                                               int max_bp1)
{
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    __shared__ chromo_replica Rs;
    load_record(&Rs, &reps[r]);
    wave_sync();
    const chromo_replica *R = &Rs;
    if (R->counters[CHROMO_CNT_ERROR] != 0) return;
    if (R->B < 1 || R->B + 1 > max_bp1) {
        counters_error(R, ERR_NBINDERS);
        return;
    }
    const int Bp1 = R->B + 1;
    const SegLds G = seg_lds(lds_dyn, Bp1);
    seg_clear(G);
    uint32_t *mt_lds = (uint32_t *)(lds_dyn + wide_lds_bytes(Bp1));
    const bool need_mt = any_binding(R);
    WaveRng rng;
    rng_load(R, rng, mt_lds, need_mt);
    Counters K;
    counters_load(R, K);
    const int nc = R->n_ctrl < MAX_CTRL ? R->n_ctrl : MAX_CTRL;
    __shared__ chromo_ctrl sctrl[MAX_CTRL];
    if (lane_id() == 0)
        for (int i = 0; i < nc; ++i) sctrl[i] = R->ctrl[i];
    __syncthreads();
    int err = 0;
    if (R->twist) {
        switch (R->B) {
        case 1: run_steps<1, true>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 2: run_steps<2, true>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 3: run_steps<3, true>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 4: run_steps<4, true>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        default: err = ERR_NBINDERS; break;
        }
    } else {
        switch (R->B) {
        case 1: run_steps<1>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 2: run_steps<2>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 3: run_steps<3>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        case 4: run_steps<4>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn); break;
        default: err = ERR_NBINDERS; break;
        }
    }
    rng_store(R, rng, need_mt);
    if (lane_id() == 0)
        for (int i = 0; i < nc; ++i) R->ctrl[i] = sctrl[i];
    counters_store(R, K, err);
}

// k_mc_run for the replicas with NB binder types and twist flag TW only; the others
// are left to the launch of their own kind.  One instantiation per kernel, so
// that its registers are allocated for that one alone (the all-kinds kernel above
// spills for the largest of its eight).
template <int NB, bool TW>
__global__ void __launch_bounds__(64) k_mc_run_nb(const chromo_replica *__restrict__ reps, int n_reps,
                                                  int64_t steps)
{
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    __shared__ chromo_replica Rs;
    load_record(&Rs, &reps[r]);
    wave_sync();
    const chromo_replica *R = &Rs;
    if (R->B != NB || (R->twist != 0) != TW) return;
    if (R->counters[CHROMO_CNT_ERROR] != 0) return;
    constexpr int Bp1 = NB + 1;
    const SegLds G = seg_lds(lds_dyn, Bp1);
    seg_clear(G);
    uint32_t *mt_lds = (uint32_t *)(lds_dyn + wide_lds_bytes(Bp1));
    const bool need_mt = any_binding(R);
    WaveRng rng;
    rng_load(R, rng, mt_lds, need_mt);
    Counters K;
    counters_load(R, K);
    const int nc = R->n_ctrl < MAX_CTRL ? R->n_ctrl : MAX_CTRL;
    __shared__ chromo_ctrl sctrl[MAX_CTRL];
    if (lane_id() == 0)
        for (int i = 0; i < nc; ++i) sctrl[i] = R->ctrl[i];
    __syncthreads();
    int err = 0;
    run_steps<NB, TW>(R, rng, K, sctrl, nc, steps, err, G, lds_dyn);
    rng_store(R, rng, need_mt);
    if (lane_id() == 0)
        for (int i = 0; i < nc; ++i) R->ctrl[i] = sctrl[i];
    counters_store(R, K, err);
}

template <int NB, bool TW>
hipError_t launch_run_nb(const chromo_replica *d_reps, int n_reps, int64_t steps, hipStream_t st)
{
    // the attribute is set once per process: setting it costs far more than a launch
    static bool attr_set = false;
    const size_t bytes = mc_lds_bytes(NB + 1);
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_mc_run_nb<NB, TW>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_mc_run_nb<NB, TW>), dim3(n_reps), dim3(WAVE), bytes, st, d_reps, n_reps, steps);
    return hipGetLastError();
}

// One mc_step per launch (the plain path / mc_step API).
__global__ void __launch_bounds__(64) k_mc_step(const chromo_replica *__restrict__ reps, int rep, int ci,
                                                int max_bp1)
{
    __shared__ chromo_replica Rs;
    load_record(&Rs, &reps[rep]);
    wave_sync();
    const chromo_replica *R = &Rs;
    if (R->counters[CHROMO_CNT_ERROR] != 0) return;
    if (R->B < 1 || R->B + 1 > max_bp1) {
        counters_error(R, ERR_NBINDERS);
        return;
    }
    const int Bp1 = R->B + 1;
    const SegLds G = seg_lds(lds_dyn, Bp1);
    seg_clear(G);
    uint32_t *mt_lds = (uint32_t *)(lds_dyn + wide_lds_bytes(Bp1));
    chromo_ctrl C = R->ctrl[ci];
    const bool need_mt = (C.move == CHROMO_CHANGE_BINDING_STATE);
    WaveRng rng;
    rng_load(R, rng, mt_lds, need_mt);
    Counters K;
    counters_load(R, K);
    MoveOut o;
    if (R->twist) {
        switch (R->B) {
        case 1: mc_step_dev<1, true>(R, rng, C, o, G, lds_dyn, 1); break;
        case 2: mc_step_dev<2, true>(R, rng, C, o, G, lds_dyn, 1); break;
        case 3: mc_step_dev<3, true>(R, rng, C, o, G, lds_dyn, 1); break;
        case 4: mc_step_dev<4, true>(R, rng, C, o, G, lds_dyn, 1); break;
        default: o.err = ERR_NBINDERS; o.n = 0; break;
        }
    } else {
        switch (R->B) {
        case 1: mc_step_dev<1>(R, rng, C, o, G, lds_dyn, 1); break;
        case 2: mc_step_dev<2>(R, rng, C, o, G, lds_dyn, 1); break;
        case 3: mc_step_dev<3>(R, rng, C, o, G, lds_dyn, 1); break;
        case 4: mc_step_dev<4>(R, rng, C, o, G, lds_dyn, 1); break;
        default: o.err = ERR_NBINDERS; o.n = 0; break;
        }
    }
    if (!o.err && o.n > 0) record(R, C, o, K);
    rng_store(R, rng, need_mt);
    if (lane_id() == 0) R->ctrl[ci] = C;
    counters_store(R, K, o.err);
}

// Flags replicas whose binder count is not in mask (bit B set for B present).
__global__ void k_check_binders(const chromo_replica *reps, int n_reps, int mask)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reps) return;
    const chromo_replica *R = &reps[r];
    const int B = R->B;
    // a twisting chain needs its twist bit: the one-wavefront launch runs it in the
    // twisting kernel of its binder count
    const bool bad = B < 1 || B > CHROMO_MAX_BINDERS || !(mask & (1 << B)) || (R->twist && !(mask & (1 << (8 + B))));
    if (bad && R->counters[CHROMO_CNT_ERROR] == 0) R->counters[CHROMO_CNT_ERROR] = ERR_NBINDERS;
}

__global__ void k_update_amplitudes(const chromo_replica *reps, int rep, int ci)
{
    if (threadIdx.x != 0) return;
    chromo_ctrl C = reps[rep].ctrl[ci];
    update_amplitudes_dev(C);
    reps[rep].ctrl[ci] = C;
}

// ---- RNG seeding ----------------------------------------------------------

__global__ void k_rng_seed(const chromo_replica *reps, int n_reps, const uint32_t *seeds,
                           const uint32_t *streams, int mode)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reps) return;
    uint32_t *w = reps[r].rng;
    uint32_t seed = seeds[r];
    if (mode == CHROMO_RNG_GLIBC) {
        // srandom_r (TYPE_3) followed by 310 discarded outputs
        int32_t st[31];
        if (seed == 0) seed = 1;
        int32_t word = (int32_t)seed;
        st[0] = word;
        for (int i = 1; i < 31; ++i) {
            int32_t hi = word / 127773;
            int32_t lo = word % 127773;
            word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            st[i] = word;
        }
        int f = 3, b = 0;
        for (int i = 0; i < 310; ++i) {
            uint32_t v = (uint32_t)st[f] + (uint32_t)st[b];
            st[f] = (int32_t)v;
            if (++f >= 31) {
                f = 0;
                ++b;
            } else if (++b >= 31) {
                b = 0;
            }
        }
        for (int i = 0; i < 31; ++i) w[CHROMO_RNG_GLIBC_STATE + i] = (uint32_t)st[i];
        w[CHROMO_RNG_GLIBC_F] = (uint32_t)f;
        w[CHROMO_RNG_GLIBC_R] = (uint32_t)b;
        // MT19937 seeded with 0 until np_seed is called
        uint32_t m = 0;
        w[CHROMO_RNG_MT_STATE] = m;
        for (int i = 1; i < 624; ++i) {
            m = 1812433253u * (m ^ (m >> 30)) + (uint32_t)i;
            w[CHROMO_RNG_MT_STATE + i] = m;
        }
        w[CHROMO_RNG_MT_INDEX] = 624;
    }
    w[CHROMO_RNG_PHILOX_KEY] = seed;
    w[CHROMO_RNG_PHILOX_KEY + 1] = streams ? streams[r] : (uint32_t)r;
    w[CHROMO_RNG_PHILOX_CTR] = 0;
    w[CHROMO_RNG_PHILOX_CTR + 1] = 0;
    w[CHROMO_RNG_PHILOX_IDX] = 4;
}

__global__ void k_np_seed(const chromo_replica *reps, int n_reps, const uint32_t *seeds)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reps) return;
    uint32_t *w = reps[r].rng;
    uint32_t m = seeds[r];
    w[CHROMO_RNG_MT_STATE] = m;
    for (int i = 1; i < 624; ++i) {
        m = 1812433253u * (m ^ (m >> 30)) + (uint32_t)i;
        w[CHROMO_RNG_MT_STATE + i] = m;
    }
    w[CHROMO_RNG_MT_INDEX] = 624;
}

// ---- scratch initialisation ----------------------------------------------

__global__ void k_scratch_init(const chromo_replica *reps, int n_reps)
{
    int r = blockIdx.y;
    if (r >= n_reps) return;
    const chromo_replica *R = &reps[r];
    if (R->scratch == nullptr || R->bead_cap <= 0) return;
    ArenaLayout L = global_layout(R->bead_cap, R->seg_cap, R->B + 1);
    int32_t *hk = (int32_t *)((char *)R->scratch + L.hkey);
    int32_t *hm = (int32_t *)((char *)R->scratch + L.hmin);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.H; i += gridDim.x * blockDim.x) {
        hk[i] = HKEY_EMPTY;
        hm[i] = HMIN_EMPTY;
    }
}

// ---- deterministic density initialisation: see chromo_density.hpp ---------

GridGeom grid_of(const chromo_replica &r)
{
    GridGeom G;
    G.nx = r.nx;
    G.ny = r.ny;
    G.nz = r.nz;
    for (int j = 0; j < 3; ++j) {
        G.width[j] = r.width[j];
        G.dxyz[j] = r.dxyz[j];
        G.half_width[j] = r.half_width[j];
        G.half_step[j] = r.half_step[j];
    }
    G.vol_bin = r.vol_bin;
    G.access_vol = r.access_vol;
    return G;
}

// ---- tape: field dE of a host-supplied trial configuration ----------------

template <int NB>
__device__ __forceinline__ void field_tape_body(const chromo_replica *R, const int64_t *inds, int64_t n,
                                const double *r_trial, const int8_t *st_trial, int state_change,
                                double *dE_out, int64_t *bins, double *delta, int64_t *nbins)
{
    const int lane = lane_id();
    const int64_t N = R->N;
    constexpr int B = NB, Bp1 = NB + 1;
    lds_hash_clear(Bp1);
    Arena A;
    if (n <= SEG_LDS) A = make_arena(lds_dyn, lds_layout(Bp1));
    else A = make_arena((char *)R->scratch, global_layout(R->bead_cap, R->seg_cap, Bp1));
    const int cap = A.cap;
    const double *Pp = R->pos;
    // get_confinement_dE (fields.pyx:121-202): beads (spherical) or coordinates
    // (cubical, trial only) outside the confinement
    int out_t = 0, out_c = 0;
    for (int64_t i = lane; i < n; i += WAVE) {
        int64_t b = inds[i];
        double o[3] = {Pp[b], Pp[N + b], Pp[2 * N + b]};
        double w[3] = {r_trial[3 * i], r_trial[3 * i + 1], r_trial[3 * i + 2]};
        if (R->confine_type == 1) {
            out_t += sqrt(dot3(w, w)) > R->confine_length;
            out_c += sqrt(dot3(o, o)) > R->confine_length;
        } else if (R->confine_type == 2) {
            for (int j = 0; j < 3; ++j) out_t += fabs(w[j]) > R->confine_length / 2;
        }
        for (int m = 0; m < B; ++m) {
            A.stold[m * cap + i] = R->states[(int64_t)m * N + b];
            A.stnew[m * cap + i] = st_trial[i * B + m];
        }
        for (int k = 0; k < 2; ++k) {
            const double *pos = (k == 0 || state_change) ? o : w;
            int64_t ix[3];
            double wt[3];
            voxel_lower_1(R, pos, ix, wt);
            const int p = (int)(2 * i + k);
            for (int l = 0; l < 8; ++l) {
                const int64_t cv = corner_voxel(R, ix, l);
                A.evox[p * 8 + l] = (int32_t)cv;
                A.c0[p * 8 + l] = corner_weight(l, wt[0], wt[1], wt[2]) / voxel_vol(R, cv);
            }
        }
    }
    __syncthreads();
    const int E = (int)(16 * n);
    for (int e = lane; e < E; e += WAVE) {
        int v = A.evox[e];
        uint32_t h = ((uint32_t)v * 2654435761u) >> A.hshift;
        for (;;) {
            int old = atomicCAS(&A.hkey[h], HKEY_EMPTY, v);
            if (old == HKEY_EMPTY || old == v) break;
            h = (h + 1) & (uint32_t)A.hmask;
        }
        atomicMin(&A.hmin[h], e);
        A.eslot[e] = (int)h;
    }
    __syncthreads();
    const int C = (E + WAVE - 1) / WAVE;
    const int e0 = lane * C, e1 = min(E, e0 + C);
    int cnt = 0;
    for (int e = e0; e < e1; ++e)
        if (A.hmin[A.eslot[e]] == e) ++cnt;
    int U;
    int base = wave_excl_scan(cnt, U);
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
    __syncthreads();
    for (int e = lane; e < E; e += WAVE) atomicAdd(&A.ucnt[A.hu[A.eslot[e]]], 1);
    __syncthreads();
    {
        const int Cu = (U + WAVE - 1) / WAVE;
        const int q0 = lane * Cu, q1 = min(U, q0 + Cu);
        int s = 0;
        for (int q = q0; q < q1; ++q) s += A.ucnt[q];
        int tot;
        int o2 = wave_excl_scan(s, tot);
        for (int q = q0; q < q1; ++q) {
            A.uoff[q] = o2;
            o2 += A.ucnt[q];
        }
        if (lane == 0) A.uoff[U] = E;
    }
    __syncthreads();
    for (int e = lane; e < E; e += WAVE) {
        int q = A.hu[A.eslot[e]];
        int pos = A.uoff[q] + atomicSub(&A.ucnt[q], 1) - 1;
        A.list[pos] = e;
    }
    __syncthreads();
    double S2[CHROMO_MAX_BINDERS], SX[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS], NS = 0;
    for (int j = 0; j < CHROMO_MAX_BINDERS; ++j) S2[j] = 0;
    for (int j = 0; j < CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS; ++j) SX[j] = 0;
    const double V = R->bead_vol, vf = R->vf_limit, chi = R->chi;
    const bool thr = R->fast_field == 0;
    for (int q = lane; q < U; q += WAVE) {
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
        double acc[CHROMO_MAX_BINDERS + 1];
        for (int m = 0; m < Bp1; ++m) acc[m] = 0;
        for (int a = s0; a < s1; ++a) {
            const int e = A.list[a];
            const int i = e >> 4, k = (e >> 3) & 1;
            const double c0 = A.c0[e];
            const double pre = k ? 1. : -1.;
            const double t0 = pre * c0;
            if (!thr || fabs(t0) > 1E-18) acc[0] += t0;
            for (int m = 1; m < Bp1; ++m) {
                int st = (k == 0 || !state_change) ? A.stold[(m - 1) * cap + i]
                                                   : A.stnew[(m - 1) * cap + i];
                double t = pre * (c0 * (double)st);
                if (!thr || fabs(t) > 1E-18) acc[m] += t;
            }
        }
        int64_t v = A.uvox[q];
        bins[q] = v;
        double r[CHROMO_MAX_BINDERS + 1];
        for (int m = 0; m < Bp1; ++m) {
            delta[(int64_t)q * Bp1 + m] = acc[m];
            r[m] = R->density[v * Bp1 + m];
        }
        for (int j = 0; j < B; ++j) {
            double aj = r[j + 1] + acc[j + 1];
            double d2 = aj * aj - r[j + 1] * r[j + 1];
            if (fabs(d2) < 1E-18) d2 = 0;
            S2[j] += d2;
            for (int kk = 0; kk < B; ++kk) {
                double dx = (aj * (r[kk + 1] + acc[kk + 1])) - (r[j + 1] * r[kk + 1]);
                if (fabs(dx) < 1E-18) dx = 0;
                SX[j * B + kk] += dx;
            }
        }
        double vf0 = r[0] * V;
        double vf1 = vf0 + (acc[0] * V);
        const double av = voxel_vol(R, v);
        if (vf1 > vf) NS += E_HUGE_FIELD * vf1;
        else NS += chi * (av / V) * (vf1 * vf1);
        if (vf0 > vf) NS -= E_HUGE_FIELD * vf0;
        else NS -= chi * (av / V) * (vf0 * vf0);
    }
    for (int j = 0; j < B; ++j) S2[j] = wave_sum(S2[j]);
    for (int j = 0; j < B * B; ++j) SX[j] = wave_sum(SX[j]);
    NS = wave_sum(NS);
    int dbl_t[CHROMO_MAX_BINDERS], dbl[CHROMO_MAX_BINDERS];
    for (int j = 0; j < B; ++j) {
        dbl_t[j] = 0;
        dbl[j] = 0;
    }
    for (int64_t i = lane; i < n; i += WAVE)
        for (int j = 0; j < B; ++j) {
            int so = A.stold[j * cap + i];
            int stt = state_change ? A.stnew[j * cap + i] : so;
            dbl_t[j] += (stt == 2);
            dbl[j] += (so == 2);
        }
    for (int j = 0; j < B; ++j) {
        dbl_t[j] = bcast0(wave_sum_int(dbl_t[j]));
        dbl[j] = bcast0(wave_sum_int(dbl[j]));
    }
    double dEb = 0.0;
    for (int i = 0; i < B; ++i) {
        dEb += R->pref[i] * S2[i];
        dEb += R->intra[i] * (double)(dbl_t[i] - dbl[i]);
    }
    for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j) dEb += R->cross[i * B + j] * SX[i * B + j];
    dEb += NS;
    for (int q = lane; q < U; q += WAVE) {
        int s = A.uslot[q];
        A.hkey[s] = HKEY_EMPTY;
        A.hmin[s] = HMIN_EMPTY;
    }
    out_t = bcast0(wave_sum_int(out_t));
    out_c = bcast0(wave_sum_int(out_c));
    double dE = 0.0;
    if (!state_change) {
        dE += (double)out_t * E_HUGE_FIELD;
        dE -= (double)out_c * E_HUGE_FIELD;
    }
    if (lane == 0) {
        *dE_out = dE + dEb;
        *nbins = U;
    }
}

__global__ void __launch_bounds__(64) k_field_tape(const chromo_replica *reps, int rep,
                                                   const int64_t *inds, int64_t n,
                                                   const double *r_trial, const int8_t *st_trial,
                                                   int state_change, double *dE_out,
                                                   int64_t *bins, double *delta, int64_t *nbins,
                                                   int max_bp1)
{
    const chromo_replica *R = &reps[rep];
    if (R->B < 1 || R->B + 1 > max_bp1) return;
    switch (R->B) {
    case 1: field_tape_body<1>(R, inds, n, r_trial, st_trial, state_change, dE_out, bins, delta, nbins); break;
    case 2: field_tape_body<2>(R, inds, n, r_trial, st_trial, state_change, dE_out, bins, delta, nbins); break;
    case 3: field_tape_body<3>(R, inds, n, r_trial, st_trial, state_change, dE_out, bins, delta, nbins); break;
    case 4: field_tape_body<4>(R, inds, n, r_trial, st_trial, state_change, dE_out, bins, delta, nbins); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

int set_err(int code, const std::string &m) { return chromo::set_last_error(code, m); }

#define HIPC(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return set_err(-2, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipError_t launch_err() { return hipGetLastError(); }

int check_reps(const chromo_replica *d_reps, int32_t n_reps)
{
    if (d_reps == nullptr) return set_err(-1, "replica table is NULL");
    if (n_reps < 1) return set_err(-1, "n_reps must be >= 1");
    return 0;
}

size_t g_lds_bytes = 0;   // field tape: LDS move arena + MT block
int g_lds_bp1 = 0;        // largest binder count + 1 the tape's arena fits
size_t g_mc_bytes = 0;    // Monte Carlo kernels: evaluator block + MT block, for B = 4

// Dynamic LDS per workgroup, sized for the largest binder count that fits the
// device's per-workgroup limit, set once per process.
int ensure_lds()
{
    if (g_lds_bytes) return 0;
    int dev = 0, lim = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    int bp1 = CHROMO_MAX_BINDERS + 1;
    while (bp1 > 1 && lds_bytes_for(bp1) + 1024 > (size_t)lim) --bp1;
    if (bp1 < 2) return set_err(-2, "device LDS per workgroup too small for the move arena");
    const size_t mc = mc_lds_bytes(CHROMO_MAX_BINDERS + 1);
    if (mc + 1024 > (size_t)lim) return set_err(-2, "device LDS per workgroup too small for the evaluator");
    HIPC(hipFuncSetAttribute((const void *)k_mc_run, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mc));
    HIPC(hipFuncSetAttribute((const void *)k_mc_step, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mc));
    HIPC(hipFuncSetAttribute((const void *)k_field_tape, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_bytes_for(bp1)));
    g_mc_bytes = mc;
    g_lds_bp1 = bp1;
    g_lds_bytes = lds_bytes_for(bp1);
    return 0;
}

}  // namespace

namespace chromo {
thread_local std::string g_last_error;
int set_last_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}
}  // namespace chromo

// chromo_capped_exponential_tape: cexp_trunc of n (draw, window) pairs.
__global__ void k_cexp_tape(const int32_t *draws, const int64_t *windows, int64_t n, int64_t *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cexp_trunc(draws[i], windows[i]);
}

// ---- checkpoint / restore -----------------------------------------------------
//
// File: "CHROMOCK" magic, u32 version, i32 replica count; per replica a header
// (N, B, n_ctrl, n_bins, rng_mode) and its state: positions (9 N doubles, as
// stored), binder states (B N bytes), voxel densities (n_bins (B+1) doubles),
// RNG block (CHROMO_RNG_WORDS words), controllers (n_ctrl records), counters.
// Marks, bond parameters and the field's parameters are inputs, not state: the
// table that restores must describe the same chains and fields.

namespace {

constexpr char CK_MAGIC[8] = {'C', 'H', 'R', 'O', 'M', 'O', 'C', 'K'};
constexpr uint32_t CK_VERSION = 1;

struct CkHeader {
    int64_t N, n_bins;
    int32_t B, n_ctrl, rng_mode, pad;
};

struct CkArrays {
    void *dev;
    size_t bytes;
};

CkHeader ck_header(const chromo_replica &r)
{
    CkHeader h{};
    h.N = r.N;
    h.n_bins = r.nx * r.ny * r.nz;
    h.B = r.B;
    h.n_ctrl = r.n_ctrl;
    h.rng_mode = r.rng_mode;
    return h;
}

void ck_arrays(const chromo_replica &r, CkArrays a[6])
{
    a[0] = {r.pos, sizeof(double) * 9 * (size_t)r.N};
    a[1] = {r.states, (size_t)r.B * (size_t)r.N};
    a[2] = {r.density, sizeof(double) * (size_t)(r.nx * r.ny * r.nz) * (size_t)(r.B + 1)};
    a[3] = {r.rng, sizeof(uint32_t) * CHROMO_RNG_WORDS};
    a[4] = {r.ctrl, sizeof(chromo_ctrl) * (size_t)r.n_ctrl};
    a[5] = {r.counters, sizeof(int64_t) * CHROMO_CNT_WORDS};
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

int read_table(const chromo_replica *d_reps, int32_t n_reps, hipStream_t st, std::vector<chromo_replica> &h)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    h.resize(n_reps);
    HIPC(hipMemcpyAsync(h.data(), d_reps, sizeof(chromo_replica) * n_reps, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

extern "C" {

int chromo_last_error(char *buf, size_t len)
{
    if (!buf || len == 0) return -1;
    const std::string &g_err = chromo::g_last_error;
    size_t n = g_err.size() < len - 1 ? g_err.size() : len - 1;
    memcpy(buf, g_err.data(), n);
    buf[n] = 0;
    return 0;
}

int chromo_abi_info(int64_t *version, int64_t *replica_bytes, int64_t *ctrl_bytes,
                    int64_t *log_bytes)
{
    if (version) *version = 1;
    if (replica_bytes) *replica_bytes = (int64_t)sizeof(chromo_replica);
    if (ctrl_bytes) *ctrl_bytes = (int64_t)sizeof(chromo_ctrl);
    if (log_bytes) *log_bytes = (int64_t)sizeof(chromo_move_log);
    return 0;
}

int64_t chromo_scratch_bytes(int64_t bead_cap, int64_t seg_cap, int32_t B)
{
    if (bead_cap <= 0) return 0;
    if (B < 0 || B > CHROMO_MAX_BINDERS || seg_cap > bead_cap || bead_cap > 0x3fffffff)
        return -1;
    ArenaLayout L = global_layout(bead_cap, seg_cap, B + 1);
    return (int64_t)L.total;
}

int chromo_rng_seed(const chromo_replica *d_reps, int32_t n_reps, const uint32_t *d_seeds,
                    const uint32_t *d_streams, int32_t mode, void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    if (mode != CHROMO_RNG_GLIBC && mode != CHROMO_RNG_PHILOX)
        return set_err(-1, "unknown rng mode");
    hipLaunchKernelGGL(k_rng_seed, dim3((n_reps + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                       d_reps, n_reps, d_seeds, d_streams, mode);
    HIPC(launch_err());
    return 0;
}

int chromo_np_seed(const chromo_replica *d_reps, int32_t n_reps, const uint32_t *d_seeds,
                   void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    hipLaunchKernelGGL(k_np_seed, dim3((n_reps + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                       d_reps, n_reps, d_seeds);
    HIPC(launch_err());
    return 0;
}

int chromo_density_init(const chromo_replica *d_reps, int32_t n_reps, void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<chromo_replica> h(n_reps);
    HIPC(hipMemcpyAsync(h.data(), d_reps, sizeof(chromo_replica) * n_reps, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    int64_t maxN = 0, maxBins = 0;
    for (auto &r : h) {
        if (r.field_on == 0) continue;
        if (r.B < 0 || r.B > CHROMO_MAX_BINDERS) return set_err(-1, "B out of range");
        int64_t nb = r.nx * r.ny * r.nz;
        if (nb > 2048LL * 2048LL || r.N > 0x7fffffff)
            return set_err(-1, "grid or chain too large for the density scatter");
        maxN = r.N > maxN ? r.N : maxN;
        maxBins = nb > maxBins ? nb : maxBins;
    }
    if (maxN == 0) return 0;
    DensityScratch S;
    HIPC(density_scratch_alloc(S, maxN, maxBins, st));
    for (int i = 0; i < n_reps; ++i) {
        const chromo_replica &r = h[i];
        if (r.field_on == 0) continue;
        PosView P{r.pos, r.N, 1};
        hipError_t e = density_scatter(grid_of(r), r.N, r.B + 1, P, r.states, r.density, 1, S, st);
        if (e != hipSuccess) {
            density_scratch_free(S, st);
            return set_err(-2, std::string("density scatter: ") + hipGetErrorString(e));
        }
    }
    density_scratch_free(S, st);
    return 0;
}

int chromo_scratch_init(const chromo_replica *d_reps, int32_t n_reps, void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scratch_init, dim3(64, n_reps), dim3(256), 0, (hipStream_t)stream, d_reps,
                       n_reps);
    HIPC(launch_err());
    return 0;
}

int chromo_mc_run(const chromo_replica *d_reps, int32_t n_reps, int64_t num_steps, void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    if (num_steps < 0) return set_err(-1, "num_steps < 0");
    rc = ensure_lds();
    if (rc) return rc;
    hipLaunchKernelGGL(k_mc_run, dim3(n_reps), dim3(64), g_mc_bytes, (hipStream_t)stream, d_reps,
                       n_reps, num_steps, CHROMO_MAX_BINDERS + 1);
    HIPC(launch_err());
    return 0;
}

int chromo_mc_run_tuned(const chromo_replica *d_reps, int32_t n_reps, int64_t num_steps,
                        int32_t waves_per_replica, int32_t binder_mask, void *stream)
{
    int rc = check_reps(d_reps, n_reps);
    if (rc) return rc;
    if (num_steps < 0) return set_err(-1, "num_steps < 0");
    if ((binder_mask & 0x1e) == 0 || (binder_mask & ~0x1e1e) != 0 || ((binder_mask >> 8) & ~binder_mask & 0x1e))
        return set_err(-1, "binder_mask must set only bits 1..4 (binder counts present) and 9..12 "
                           "(binder counts with twisting chains)");
    const int W = waves_per_replica;
    if (W == 4 || W == 8)
        return set_err(-1, "waves_per_replica = 4 and 8 are not built: the speculative round has "
                           "one wavefront on each side of its flags");
    if (W != 1 && W != 2) return set_err(-1, "waves_per_replica must be 1 or 2");
    rc = ensure_lds();
    if (rc) return rc;
    int bp1 = 0;
    for (int b = 1; b <= CHROMO_MAX_BINDERS; ++b)
        if (binder_mask & (1 << b)) bp1 = b + 1;
    hipStream_t st = (hipStream_t)stream;
    // replicas whose binder count is not in binder_mask fail with ERR_NBINDERS
    hipLaunchKernelGGL(k_check_binders, dim3((n_reps + 63) / 64), dim3(64), 0, st, d_reps, n_reps,
                       binder_mask);
    HIPC(launch_err());
    if (W == 1) {
        // one kernel per binder count (and twist flag) present
        hipError_t e = hipSuccess;
        if ((binder_mask & 2) && e == hipSuccess) e = launch_run_nb<1, false>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 4) && e == hipSuccess) e = launch_run_nb<2, false>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 8) && e == hipSuccess) e = launch_run_nb<3, false>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 16) && e == hipSuccess) e = launch_run_nb<4, false>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 0x200) && e == hipSuccess) e = launch_run_nb<1, true>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 0x400) && e == hipSuccess) e = launch_run_nb<2, true>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 0x800) && e == hipSuccess) e = launch_run_nb<3, true>(d_reps, n_reps, num_steps, st);
        if ((binder_mask & 0x1000) && e == hipSuccess) e = launch_run_nb<4, true>(d_reps, n_reps, num_steps, st);
        HIPC(e);
        return 0;
    }
    hipError_t e = chromo::spec_launch_w2(d_reps, n_reps, num_steps, binder_mask & 0x1e, st);
    if (e == hipErrorInvalidValue)
        return set_err(-1, "LDS of this launch shape exceeds the device limit");
    HIPC(e);
    return 0;
}

int chromo_mc_step(const chromo_replica *d_reps, int32_t rep, int32_t ctrl, void *stream)
{
    if (d_reps == nullptr || rep < 0) return set_err(-1, "bad replica");
    int rc = ensure_lds();
    if (rc) return rc;
    hipLaunchKernelGGL(k_mc_step, dim3(1), dim3(64), g_mc_bytes, (hipStream_t)stream, d_reps, rep,
                       ctrl, CHROMO_MAX_BINDERS + 1);
    HIPC(launch_err());
    return 0;
}

int chromo_update_amplitudes(const chromo_replica *d_reps, int32_t rep, int32_t ctrl, void *stream)
{
    if (d_reps == nullptr || rep < 0) return set_err(-1, "bad replica");
    hipLaunchKernelGGL(k_update_amplitudes, dim3(1), dim3(64), 0, (hipStream_t)stream, d_reps, rep,
                       ctrl);
    HIPC(launch_err());
    return 0;
}

int chromo_mc_run_plain(const chromo_replica *d_reps, int32_t rep, int64_t num_steps, void *stream)
{
    if (d_reps == nullptr || rep < 0) return set_err(-1, "bad replica");
    int rc = ensure_lds();
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    chromo_replica h;
    HIPC(hipMemcpyAsync(&h, d_reps + rep, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (h.n_ctrl < 0 || h.n_ctrl > MAX_CTRL) return set_err(-1, "n_ctrl out of range");
    std::vector<chromo_ctrl> c(h.n_ctrl);
    HIPC(hipMemcpyAsync(c.data(), h.ctrl, sizeof(chromo_ctrl) * h.n_ctrl, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int64_t s = 0; s < num_steps; ++s) {
        for (int ci = 0; ci < h.n_ctrl; ++ci) {
            if (c[ci].on == 1)
                for (int j = 0; j < c[ci].num_per_cycle; ++j)
                    hipLaunchKernelGGL(k_mc_step, dim3(1), dim3(64), g_mc_bytes, st, d_reps, rep, ci,
                                       CHROMO_MAX_BINDERS + 1);
            hipLaunchKernelGGL(k_update_amplitudes, dim3(1), dim3(64), 0, st, d_reps, rep, ci);
        }
        HIPC(launch_err());
    }
    return 0;
}

int chromo_field_dE_tape(const chromo_replica *d_reps, int32_t rep, const int64_t *d_inds,
                         int64_t n, const double *d_r_trial, const int8_t *d_states_trial,
                         int32_t state_change, double *d_dE, int64_t *d_bins, double *d_delta,
                         int64_t *d_n_bins, void *stream)
{
    if (d_reps == nullptr || rep < 0) return set_err(-1, "bad replica");
    if (n < 1) return set_err(-1, "n must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    chromo_replica h;
    HIPC(hipMemcpyAsync(&h, d_reps + rep, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (n > SEG_LDS && (n > h.seg_cap || n > h.bead_cap))
        return set_err(-1, "n exceeds the replica's scratch arena (seg_cap)");
    if (h.field_on == 0) return set_err(-1, "replica has no field");
    int rc = ensure_lds();
    if (rc) return rc;
    if (h.B < 1 || h.B + 1 > g_lds_bp1) return set_err(-1, "binder count exceeds the LDS arena");
    hipLaunchKernelGGL(k_field_tape, dim3(1), dim3(64), g_lds_bytes, st, d_reps, rep, d_inds, n,
                       d_r_trial, d_states_trial, state_change, d_dE, d_bins, d_delta, d_n_bins,
                       g_lds_bp1);
    HIPC(launch_err());
    return 0;
}

int chromo_checkpoint(const chromo_replica *d_reps, int32_t n_reps, const char *path, void *stream)
{
    if (!path) return set_err(-1, "path is NULL");
    hipStream_t st = (hipStream_t)stream;
    std::vector<chromo_replica> h;
    int rc = read_table(d_reps, n_reps, st, h);
    if (rc) return rc;
    File out;
    out.f = fopen(path, "wb");
    if (!out.f) return set_err(-1, std::string("cannot open ") + path + " for writing");
    bool ok = fwrite(CK_MAGIC, 1, 8, out.f) == 8 && fwrite(&CK_VERSION, 4, 1, out.f) == 1 &&
              fwrite(&n_reps, 4, 1, out.f) == 1;
    std::vector<char> buf;
    for (int32_t i = 0; i < n_reps && ok; ++i) {
        const CkHeader hd = ck_header(h[i]);
        ok = fwrite(&hd, sizeof(hd), 1, out.f) == 1;
        CkArrays a[6];
        ck_arrays(h[i], a);
        for (int k = 0; k < 6 && ok; ++k) {
            buf.resize(a[k].bytes);
            if (a[k].bytes) {
                HIPC(hipMemcpyAsync(buf.data(), a[k].dev, a[k].bytes, hipMemcpyDeviceToHost, st));
                HIPC(hipStreamSynchronize(st));
                ok = fwrite(buf.data(), 1, a[k].bytes, out.f) == a[k].bytes;
            }
        }
    }
    if (!ok) return set_err(-1, std::string("write to ") + path + " failed");
    if (fflush(out.f) != 0) return set_err(-1, std::string("write to ") + path + " failed");
    return 0;
}

int chromo_restore(const chromo_replica *d_reps, int32_t n_reps, const char *path, void *stream)
{
    if (!path) return set_err(-1, "path is NULL");
    hipStream_t st = (hipStream_t)stream;
    std::vector<chromo_replica> h;
    int rc = read_table(d_reps, n_reps, st, h);
    if (rc) return rc;
    File in;
    in.f = fopen(path, "rb");
    if (!in.f) return set_err(-1, std::string("cannot open ") + path);
    char magic[8];
    uint32_t ver = 0;
    int32_t n = 0;
    if (fread(magic, 1, 8, in.f) != 8 || memcmp(magic, CK_MAGIC, 8) != 0 || fread(&ver, 4, 1, in.f) != 1 ||
        fread(&n, 4, 1, in.f) != 1)
        return set_err(-1, std::string(path) + " is not a chromo checkpoint");
    if (ver != CK_VERSION) return set_err(-1, "checkpoint version " + std::to_string(ver) + " is not supported");
    if (n != n_reps)
        return set_err(-1, "checkpoint holds " + std::to_string(n) + " replicas, the table " + std::to_string(n_reps));
    // read and check everything before writing any device memory
    std::vector<std::vector<char>> data((size_t)n_reps * 6);
    for (int32_t i = 0; i < n_reps; ++i) {
        CkHeader hd{}, want = ck_header(h[i]);
        if (fread(&hd, sizeof(hd), 1, in.f) != 1) return set_err(-1, std::string(path) + " is truncated");
        if (hd.N != want.N || hd.B != want.B || hd.n_ctrl != want.n_ctrl || hd.n_bins != want.n_bins ||
            hd.rng_mode != want.rng_mode)
            return set_err(-1, "replica " + std::to_string(i) +
                                   ": checkpoint chain, grid, controllers or RNG mode differ from the table");
        CkArrays a[6];
        ck_arrays(h[i], a);
        for (int k = 0; k < 6; ++k) {
            auto &v = data[(size_t)i * 6 + k];
            v.resize(a[k].bytes);
            if (a[k].bytes && fread(v.data(), 1, a[k].bytes, in.f) != a[k].bytes)
                return set_err(-1, std::string(path) + " is truncated");
        }
    }
    for (int32_t i = 0; i < n_reps; ++i) {
        CkArrays a[6];
        ck_arrays(h[i], a);
        for (int k = 0; k < 6; ++k)
            if (a[k].bytes)
                HIPC(hipMemcpyAsync(a[k].dev, data[(size_t)i * 6 + k].data(), a[k].bytes, hipMemcpyHostToDevice, st));
    }
    HIPC(hipStreamSynchronize(st));
    return 0;
}

int chromo_capped_exponential_tape(const int32_t *d_draws, const int64_t *d_windows, int64_t n,
                                   int64_t *d_out, void *stream)
{
    if (n < 0 || (n > 0 && (d_draws == nullptr || d_windows == nullptr || d_out == nullptr)))
        return set_err(-1, "bad arguments");
    if (n == 0) return 0;
    const int threads = 256;
    const int64_t blocks = (n + threads - 1) / threads;
    if (blocks > 0x7fffffff) return set_err(-1, "n too large");
    hipLaunchKernelGGL(k_cexp_tape, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, d_draws,
                       d_windows, n, d_out);
    HIPC(launch_err());
    return 0;
}

}  // extern "C"
