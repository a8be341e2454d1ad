// chromo_device.hpp -- device building blocks of the MI355X Monte Carlo inner loop.
//
// Execution model: one wavefront (64 lanes, one workgroup) owns one replica's
// Markov chain and executes its moves one after another.  Scalar parts of a move
// (random draws, bead selection, the rotation matrix, the Metropolis test) run as
// wave-uniform code; per-bead and per-voxel parts spread over the lanes.
//
// All floating-point code is compiled with -ffp-contract=off and keeps the
// evaluation order of the reference (chromo/*.pyx), so that positions, voxel
// indices, delta densities and decisions reproduce the reference arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "chromo_hip.h"

namespace chromo {

// Records the message returned by chromo_last_error (this thread) and returns code.
int set_last_error(int code, const std::string &msg);

constexpr int WAVE = 64;
constexpr int SEG_LDS = 32;                 // moves with n_inds <= SEG_LDS use the LDS arena
constexpr int HASH_LDS = 1024;              // >= 2 * 16 * SEG_LDS
constexpr int MAX_CTRL = 8;
constexpr double RMAX = 2147483647.0;       // RAND_MAX as a double
constexpr double E_HUGE_FIELD = 1E99;       // chromo/fields.pyx:32
constexpr double E_HUGE_POLY = 1E25;        // chromo/polymers.pyx:35
constexpr double TIE_VOXEL_EPS = 1E-11;     // voxel units: a floor that an ulp could flip
constexpr int HKEY_EMPTY = -1;
constexpr int HMIN_EMPTY = 0x7fffffff;

enum : int {
    ERR_WINDOW = 1,       // bead selection window larger than the chain (ValueError)
    ERR_SEGCAP = 2,       // move larger than the scratch arena
    ERR_TANGENT_K = 3,    // more distinct beads requested than the chain has
    ERR_NBINDERS = 4,     // binder count outside 1..4 or not in the launch's binder_mask
    ERR_SPIN = 5,         // a rejection or probe loop ran past SPIN_LIMIT iterations
    ERR_ARENA = 6,        // a move-arena index check failed (internal error)
    ERR_TWIST_WAVES = 7,  // a twisting chain in the speculative (two-wavefront) kernel
};

// Iteration bound of every data-dependent loop (rejection sampling, distinct-bead
// picks, hash probes).  Valid inputs stay far below it; a loop that reaches it
// stops with ERR_SPIN, so every wavefront always reaches the end of the kernel.
constexpr int SPIN_LIMIT = 1 << 20;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
// Wavefront index in the workgroup, as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE)); }

// Orders this wavefront's LDS accesses before / after the call for all of its
// lanes (the LDS executes one wavefront's operations in order).  Used by code
// that one wavefront of a workgroup runs on its own.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// As wave_sync, and also waits for this wavefront's global-memory accesses.
__device__ __forceinline__ void wave_sync_global()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Wave-level helpers
// ---------------------------------------------------------------------------

// Lane l's value as a wave-uniform scalar (l must be wave-uniform).
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l)
{
    uint32_t lo = readlane_u32((uint32_t)v, l);
    uint32_t hi = readlane_u32((uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    return __longlong_as_double((long long)readlane_i64((int64_t)__double_as_longlong(v), l));
}

__device__ __forceinline__ double bcast0(double x) { return readlane_f64(x, 0); }

// The same pointer in the global address space: its loads and stores are global_*
// instructions, counted apart from LDS and scalar loads, so waiting for an LDS
// value does not also wait for outstanding HBM traffic (generic flat_*
// instructions count on both).
#define GLOBAL_AS __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GLOBAL_AS T *gmem(T *p)
{
    return (GLOBAL_AS T *)p;
}

// True when pred holds on any active lane.  A ballot, which the compiler knows is
// wave-uniform (a loop that exits on it keeps its other values in scalar
// registers); the runtime's __any is an opaque library call.
__device__ __forceinline__ bool wave_any(bool pred) { return __builtin_amdgcn_ballot_w64(pred) != 0; }

// A value every lane holds equally, marked as such for the compiler (scalar
// registers and scalar arithmetic).
__device__ __forceinline__ uint32_t uni_u32(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ int uni_i32(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni_u64(uint64_t x)
{
    return ((uint64_t)uni_u32((uint32_t)(x >> 32)) << 32) | uni_u32((uint32_t)x);
}
__device__ __forceinline__ int64_t bcast0(int64_t x) { return readlane_i64(x, 0); }
__device__ __forceinline__ int bcast0(int x) { return (int)readlane_u32((uint32_t)x, 0); }

// DPP lane permutations within rows of 16 lanes (VALU operations, no LDS trip).
constexpr int DPP_XOR1 = 0xB1;          // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;          // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <-> 7 - i within each 8
constexpr int DPP_MIRROR = 0x140;       // lane i <-> 15 - i within each 16
constexpr int DPP_SHR1 = 0x111, DPP_SHR2 = 0x112, DPP_SHR4 = 0x114, DPP_SHR8 = 0x118;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_i32<CTRL>((int)b);
    const int hi = dpp_i32<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <int CTRL>
__device__ __forceinline__ int64_t dpp_i64(int64_t x)
{
    const int lo = dpp_i32<CTRL>((int)x);
    const int hi = dpp_i32<CTRL>((int)(x >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over all 64 lanes in a fixed order: pairwise within rows of 16 (every lane
// of a row ends with the same row sum), then the four row sums left to right.
// Wave-uniform result.  Called with all lanes active.
__device__ __forceinline__ double wave_sum(double x)
{
    x += dpp_f64<DPP_XOR1>(x);
    x += dpp_f64<DPP_XOR2>(x);
    x += dpp_f64<DPP_HALF_MIRROR>(x);
    x += dpp_f64<DPP_MIRROR>(x);
    const double r0 = readlane_f64(x, 0), r1 = readlane_f64(x, 16), r2 = readlane_f64(x, 32),
                 r3 = readlane_f64(x, 48);
    return ((r0 + r1) + r2) + r3;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t x)
{
    x += dpp_i64<DPP_XOR1>(x);
    x += dpp_i64<DPP_XOR2>(x);
    x += dpp_i64<DPP_HALF_MIRROR>(x);
    x += dpp_i64<DPP_MIRROR>(x);
    return readlane_i64(x, 0) + readlane_i64(x, 16) + readlane_i64(x, 32) + readlane_i64(x, 48);
}

__device__ __forceinline__ int wave_sum_int(int x)
{
    x += dpp_i32<DPP_XOR1>(x);
    x += dpp_i32<DPP_XOR2>(x);
    x += dpp_i32<DPP_HALF_MIRROR>(x);
    x += dpp_i32<DPP_MIRROR>(x);
    return (int)(readlane_u32((uint32_t)x, 0) + readlane_u32((uint32_t)x, 16) + readlane_u32((uint32_t)x, 32) +
                 readlane_u32((uint32_t)x, 48));
}

// Sum over lanes 0..7 (wave-uniform), for values held by the first eight lanes.
__device__ __forceinline__ int64_t sum_lanes8_i64(int64_t x)
{
    x += dpp_i64<DPP_XOR1>(x);
    x += dpp_i64<DPP_XOR2>(x);
    x += dpp_i64<DPP_HALF_MIRROR>(x);
    return readlane_i64(x, 0);
}

__device__ __forceinline__ uint32_t wave_or(uint32_t m)
{
    int x = (int)m;
    x |= dpp_i32<DPP_XOR1>(x);
    x |= dpp_i32<DPP_XOR2>(x);
    x |= dpp_i32<DPP_HALF_MIRROR>(x);
    x |= dpp_i32<DPP_MIRROR>(x);
    return readlane_u32((uint32_t)x, 0) | readlane_u32((uint32_t)x, 16) | readlane_u32((uint32_t)x, 32) |
           readlane_u32((uint32_t)x, 48);
}

// Exclusive prefix sum over the wave; total receives the wave sum.
// Exclusive prefix count of a predicate over the wavefront, and its total:
// wave_excl_scan of 0 / 1 values from one ballot.  Whole wave.
__device__ __forceinline__ int wave_excl_count(bool pred, int &total)
{
    const uint64_t m = __builtin_amdgcn_ballot_w64(pred);
    total = __popcll(m);
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Fold-length classes, longest first: the positions a voxel's fold visits (1 or
// 2 -> 7, ..., 15 or more -> 0).
__device__ __forceinline__ int fold_class(int visits) { return visits >= 15 ? 0 : (16 - visits) >> 1; }

// Ranks of the lane's J items (item j of lane l is item l + 64 j) in the order
// (class, item) for classes 0..7; items of class 8 or more get no rank.  Items
// dealt out to lanes in that order give each round of a per-voxel fold loop
// voxels of similar length: the loop runs as long as its longest voxel.
// Deterministic.  Whole wave.
template <int J>
__device__ __forceinline__ void wave_rank_by_class(const int *cls, int *rank)
{
    int off = 0;
    for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const uint64_t m = __builtin_amdgcn_ballot_w64(cls[j] == c);
            if (cls[j] == c)
                rank[j] = off + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            off += __popcll(m);
        }
    }
}

__device__ __forceinline__ int wave_excl_scan(int x, int &total)
{
    // inclusive scan within rows of 16 (lanes shifted in from outside the row read 0)
    int v = x;
    v += __builtin_amdgcn_update_dpp(0, v, DPP_SHR1, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_SHR2, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_SHR4, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_SHR8, 0xF, 0xF, true);
    const int s0 = (int)readlane_u32((uint32_t)v, 15), s1 = (int)readlane_u32((uint32_t)v, 31),
              s2 = (int)readlane_u32((uint32_t)v, 47), s3 = (int)readlane_u32((uint32_t)v, 63);
    const int row = lane_id() >> 4;
    v += (row > 0 ? s0 : 0) + (row > 1 ? s1 : 0) + (row > 2 ? s2 : 0);
    total = s0 + s1 + s2 + s3;
    return v - x;
}

// ---------------------------------------------------------------------------
// Random streams (wave-uniform)
// ---------------------------------------------------------------------------

// glibc random_r TYPE_3 state: lane j (j < 31) holds word j; the front and rear
// indices are wave-uniform.  Replaces libc rand() as called throughout the
// reference (chromo/mc/mc_sim.pyx:171, chromo/mc/move_funcs.pyx,
// chromo/util/linalg.pyx, chromo/util/bead_selection.pyx).
struct WaveRng {
    int mode;
    uint32_t g;
    int gf, gr;
    uint32_t *mt;       // 624 words in LDS: numpy MT19937 state (glibc mode) or the Philox batch
    int mti;
    // Philox: a batch of WAVE consecutive blocks in LDS, block pbase + j at words
    // 4j .. 4j + 3 of mt, its base counter at words PHILOX_BASE.. (shared by every
    // WaveRng of the wavefront that uses this LDS); pidx is the next of its 4 * WAVE
    // words.  pbase == PHILOX_NONE: no batch.
    uint32_t k0, k1;
    uint64_t pbase;
    int pidx;
};

constexpr uint64_t PHILOX_NONE = ~0ull;
constexpr int PHILOX_BASE = 4 * 64;     // words of the batch, then its base counter

__device__ __forceinline__ int32_t glibc_draw(WaveRng &R)
{
    R.gf = uni_i32(R.gf);
    R.gr = uni_i32(R.gr);
    uint32_t a = readlane_u32(R.g, R.gf);
    uint32_t b = readlane_u32(R.g, R.gr);
    uint32_t v = a + b;
    if (lane_id() == R.gf) R.g = v;
    if (++R.gf >= 31) {
        R.gf = 0;
        ++R.gr;
    } else if (++R.gr >= 31) {
        R.gr = 0;
    }
    return (int32_t)(v >> 1);
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Blocks base .. base + WAVE - 1 of the stream, one per lane, into the LDS batch:
// the same words a block-at-a-time generator produces, WAVE blocks per Philox
// latency.
__device__ __forceinline__ void philox_fill(WaveRng &R, uint64_t base)
{
    R.pbase = uni_u64(base);
    const int lane = lane_id();
    const uint64_t c = R.pbase + (uint64_t)lane;
    uint32_t w[4] = {(uint32_t)c, (uint32_t)(c >> 32), 0u, 0u};
    philox10(w, uni_u32(R.k0), uni_u32(R.k1));
    *(uint4 *)&R.mt[4 * lane] = make_uint4(w[0], w[1], w[2], w[3]);
    if (lane == 0) {
        R.mt[PHILOX_BASE] = (uint32_t)R.pbase;
        R.mt[PHILOX_BASE + 1] = (uint32_t)(R.pbase >> 32);
    }
    wave_sync();
}

__device__ __forceinline__ uint32_t philox_draw(WaveRng &R)
{
    R.pidx = uni_i32(R.pidx);
    if (R.pidx >= 4 * WAVE) {
        philox_fill(R, R.pbase + WAVE);
        R.pidx = 0;
    }
    const uint32_t v = uni_u32(R.mt[R.pidx]);
    R.pidx++;
    return v;
}

// Marks the wavefront's LDS batch empty (before a WaveRng first uses it).
__device__ __forceinline__ void philox_reset(const WaveRng &R)
{
    if (lane_id() == 0) {
        R.mt[PHILOX_BASE] = (uint32_t)PHILOX_NONE;
        R.mt[PHILOX_BASE + 1] = (uint32_t)(PHILOX_NONE >> 32);
    }
    wave_sync();
}

// The block-at-a-time form of the Philox position, as stored between launches and
// in checkpoints: the next block to generate (ctr), the next word of the current
// block (bi, 4 when it is used up) and that block's words.
__device__ __forceinline__ void philox_get(const WaveRng &R, uint64_t &ctr, int &bi, uint32_t b[4])
{
    const uint64_t blk = R.pbase + (uint64_t)(R.pidx >> 2);   // block of the next draw
    if ((R.pidx & 3) == 0) {
        ctr = blk;
        bi = 4;
        b[0] = b[1] = b[2] = b[3] = 0;
    } else {
        const int l = R.pidx >> 2;
        ctr = blk + 1;
        bi = R.pidx & 3;
        for (int j = 0; j < 4; ++j) b[j] = uni_u32(R.mt[4 * l + j]);
    }
}

// The position (ctr, bi) of philox_get without the block words: all philox_set
// needs.
__device__ __forceinline__ void philox_pos(const WaveRng &R, uint64_t &ctr, int &bi)
{
    const uint64_t blk = R.pbase + (uint64_t)(R.pidx >> 2);
    const bool start = (R.pidx & 3) == 0;
    ctr = start ? blk : blk + 1;
    bi = start ? 4 : (R.pidx & 3);
}

// Moves to the position (ctr, bi) of philox_get; recomputes the LDS batch only
// when that position lies outside the batch it holds.
__device__ __forceinline__ void philox_set(WaveRng &R, uint64_t ctr, int bi)
{
    ctr = uni_u64(ctr);
    bi = uni_i32(bi);
    const uint64_t blk = bi >= 4 ? ctr : ctr - 1;
    const int w = bi >= 4 ? 0 : bi;
    wave_sync();
    const uint64_t have = uni_u64((uint64_t)R.mt[PHILOX_BASE] | ((uint64_t)R.mt[PHILOX_BASE + 1] << 32));
    if (have == PHILOX_NONE || blk < have || blk >= have + WAVE) {
        philox_fill(R, blk);
        R.pidx = w;
    } else {
        R.pbase = have;
        R.pidx = (int)(blk - have) * 4 + w;
    }
}

// MT19937 regeneration of the 624-word block (lane 0, LDS), then tempering.
__device__ void mt_twist(uint32_t *mt)
{
    wave_sync();
    if (lane_id() == 0) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (mt[i] & 0x80000000u) | (mt[i + 1 == 624 ? 0 : i + 1] & 0x7fffffffu);
            int j = i + 397;
            if (j >= 624) j -= 624;
            uint32_t v = mt[j] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            mt[i] = v;
        }
    }
    wave_sync();
}

__device__ __forceinline__ uint32_t mt_draw(WaveRng &R)
{
    R.mti = uni_i32(R.mti);
    if (R.mti >= 624) {
        mt_twist(R.mt);
        R.mti = 0;
    }
    uint32_t y = uni_u32(R.mt[R.mti++]);
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ int32_t rnd31(WaveRng &R)
{
    if (R.mode == CHROMO_RNG_GLIBC) return glibc_draw(R);
    return (int32_t)(philox_draw(R) >> 1);
}

__device__ __forceinline__ double rnd_unit(WaveRng &R) { return (double)rnd31(R) / RMAX; }

// Per lane, the Philox word t draws ahead of the next one (t = 0 is the next
// draw); valid while pidx + t < 4 * WAVE.  Advances nothing.
__device__ __forceinline__ uint32_t philox_peek(const WaveRng &R, int t)
{
    return R.mt[R.pidx + t];
}

// C's x % m for a draw x in [0, 2^31) and m >= 1, in 32-bit arithmetic.
__device__ __forceinline__ int64_t draw_mod(int32_t x, int64_t m)
{
    if (m > 0x7fffffff) return x;
    return (int64_t)((uint32_t)x % (uint32_t)m);
}

// The draws of a tangent rotation after its angle and count
// (move_funcs.pyx:470-582): k distinct beads, each rand() % N with a repeat drawn
// again, then two unit-sphere draws per bead.  When no bead repeats, these are
// the next 3k draws; in the Philox mode, with all of them in the current batch,
// lane i < k reads bead i and its two draws from the batch directly.  Returns
// false, having drawn nothing, when that does not apply: the caller then draws
// them one at a time.
__device__ __forceinline__ bool tangent_draws_at_once(WaveRng &R, int64_t k, int64_t N, int64_t &pick,
                                                      int32_t &d1, int32_t &d2)
{
    if (R.mode != CHROMO_RNG_PHILOX || k > WAVE || R.pidx + 3 * k > 4 * WAVE) return false;
    const int lane = lane_id();
    const bool act = lane < k;
    const int64_t c = draw_mod((int32_t)(philox_peek(R, act ? lane : 0) >> 1), N);
    bool rep = false;
    for (int j = 0; j + 1 < k; ++j) rep |= act && lane > j && c == readlane_i64(c, j);
    if (wave_any(rep)) return false;
    const int t = (int)k + 2 * lane;
    d1 = (int32_t)(philox_peek(R, act ? t : 0) >> 1);
    d2 = (int32_t)(philox_peek(R, act ? t + 1 : 0) >> 1);
    pick = act ? c : 0;
    R.pidx += (int)(3 * k);
    return true;
}

// numpy legacy RandomState.randint(0, high): masked rejection over 32-bit draws
// (chromo/mc/move_funcs.pyx:819).
__device__ __forceinline__ int64_t rnd_randint(WaveRng &R, int64_t high)
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
        uint32_t v = (R.mode == CHROMO_RNG_GLIBC ? mt_draw(R) : philox_draw(R)) & mask;
        if (v <= rng) return (int64_t)v;
    }
}

// ---------------------------------------------------------------------------
// Geometry (chromo/util/linalg.pyx, chromo/mc/move_funcs.pyx)
// ---------------------------------------------------------------------------

// uniform_sample_unit_sphere, linalg.pyx:23-59, from its two rand() draws
// sin and cos of up to three wave-uniform angles, one lane each: one sin and one
// cos latency for all three.  Called by the whole wavefront.
__device__ __forceinline__ void sincos_uni3(double a0, double a1, double a2, double s[3], double c[3])
{
    const int lane = lane_id();
    const double x = lane == 1 ? a1 : (lane == 2 ? a2 : a0);
    const double sv = sin(x), cv = cos(x);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        s[j] = readlane_f64(sv, j);
        c[j] = readlane_f64(cv, j);
    }
}

__device__ __forceinline__ double sphere_phi(int32_t d1) { return (double)d1 / RMAX * (2.0 * M_PI); }
__device__ __forceinline__ double sphere_theta(int32_t d2) { return acos(((double)d2 / RMAX) * 2 - 1); }

__device__ __forceinline__ void sphere_dir(double sp, double cp, double st, double ct, double v[3])
{
    v[0] = cp * st;
    v[1] = sp * st;
    v[2] = ct;
}

// Per lane: the unit vector of the draws d1, d2.
__device__ __forceinline__ void sphere_from(int32_t d1, int32_t d2, double v[3])
{
    const double phi = sphere_phi(d1), theta = sphere_theta(d2);
    sphere_dir(sin(phi), cos(phi), sin(theta), cos(theta), v);
}

// Wave-uniform draws and angle: the unit vector and the sine and cosine of ang,
// at once.  Called by the whole wavefront.
__device__ __forceinline__ void sphere_from_uni(int32_t d1, int32_t d2, double ang, double v[3], double &sa,
                                                double &ca)
{
    double s[3], c[3];
    sincos_uni3(sphere_phi(d1), sphere_theta(d2), ang, s, c);
    sphere_dir(s[0], c[0], s[1], c[1], v);
    sa = s[2];
    ca = c[2];
}

// Called by the whole wavefront.
__device__ __forceinline__ void sample_unit_sphere(WaveRng &R, double v[3])
{
    int32_t d1 = rnd31(R);
    int32_t d2 = rnd31(R);
    double sa, ca;
    sphere_from_uni(d1, d2, 0.0, v, sa, ca);
}

// arbitrary_axis_rotation, linalg.pyx:62-139 (rows 0..2 of the 4x4 matrix), with
// c = cos(ang) and s = sin(ang) given
__device__ __forceinline__ void axis_rotation_cs(const double a[3], const double p[3], double c, double s,
                                                 double T[12])
{
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

__device__ __forceinline__ void axis_rotation(const double a[3], const double p[3], double ang, double T[12])
{
    axis_rotation_cs(a, p, cos(ang), sin(ang), T);
}

// One row of transform_r_t3_t2 (move_funcs.pyx:139-154): left fold from zero.
__device__ __forceinline__ double rowdot(const double *Trow, double x0, double x1, double x2)
{
    double e = 0;
    e += Trow[0] * x0;
    e += Trow[1] * x1;
    e += Trow[2] * x2;
    return e;
}

// Cython float modulo with cdivision=False (fmod + sign fix).
// fmod(a, b) for b > 0 and |a / b| < 2^40, without branches.  The remainder
// a - q b, q the truncated quotient, is representable, so one fma gives it exactly
// once q is right; q = trunc(a / b) from the rounded division is off by at most one
// when the quotient is near an integer, which the sign / range of a first fma
// result shows.  Equal to fmod bit for bit, the sign of a zero included.
__device__ __forceinline__ double fmod_exact(double a, double b)
{
    double q = trunc(a / b);
    double r = fma(-q, b, a);
    const double dq = (a >= 0) ? ((r < 0) ? -1.0 : (r >= b ? 1.0 : 0.0))
                               : ((r > 0) ? 1.0 : (r <= -b ? -1.0 : 0.0));
    q += dq;
    r = fma(-q, b, a);
    return r == 0 ? copysign(0.0, a) : r;
}

// Python's float a % b for b > 0: fmod (exact), then b added to a nonzero result of
// the other sign.
__device__ __forceinline__ double pymod(double a, double b)
{
    double r;
    if (b > 0 && fabs(a) < 0x1p40 * b) r = fmod_exact(a, b);
    else r = fmod(a, b);
    r += (double)((r != 0) & ((r < 0) ^ (b < 0))) * b;
    return r;
}

// ---------------------------------------------------------------------------
// Bead selection (chromo/util/bead_selection.pyx), wave-uniform
// ---------------------------------------------------------------------------

__device__ __forceinline__ int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
__device__ __forceinline__ int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }

// The truncated exponential of one draw x (bead_selection.pyx:58-65), in the
// reference's operation order.  Every device path draws bead windows through this
// one function, and chromo_capped_exponential_tape exposes it for the exhaustive
// comparison with glibc's log10 near every integer crossing.
__device__ __forceinline__ int64_t cexp_trunc(int32_t x, int64_t window)
{
    return (int64_t)(-log10(((double)x / RMAX) + 0.00001) * (double)window * 0.45 + 1.0001);
}

// capped_exponential, bead_selection.pyx:19-66
// The draw accepts r <= cap with probability above 0.99 for cap >= 1; cap < 1 can
// never accept (the reference loops for ever) and is reported as ERR_SPIN.
__device__ __forceinline__ int64_t capped_exponential(WaveRng &R, int64_t window, int64_t cap, int &err)
{
    if (cap < 1) {
        err = ERR_SPIN;
        return 0;
    }
    int64_t r;
    int it = 0;
    do {
        r = cexp_trunc(rnd31(R), window);
        if (++it >= SPIN_LIMIT) {
            err = ERR_SPIN;
            return 0;
        }
    } while (r > cap);
    return r;
}

__device__ __forceinline__ int64_t from_left(WaveRng &R, int64_t w, int64_t nb, int &err)
{
    if (w > nb) {
        err = ERR_WINDOW;
        return 0;
    }
    return capped_exponential(R, w, w, err);
}

__device__ __forceinline__ int64_t from_right(WaveRng &R, int64_t w, int64_t nb, int &err)
{
    int64_t d = from_left(R, w, nb, err);
    return nb - d;
}

__device__ __forceinline__ int64_t from_point(WaveRng &R, int64_t w, int64_t nb, int64_t ind0, int &err)
{
    if (w > nb) {
        err = ERR_WINDOW;
        return 0;
    }
    if (w < 1) return ind0;
    int64_t side = rnd31(R) & 1;
    if (side == 0) {
        int64_t ws = imax64(imin64(w, ind0), 1);
        int64_t ub = imax64(ind0, 1);
        return from_right(R, ws, ub, err);
    }
    int64_t ws = imax64(imin64(w, nb - ind0), 1);
    int64_t ub = imax64(nb - ind0, 1);
    return from_left(R, ws, ub, err) + ind0;
}

__device__ __forceinline__ void check_bead_bounds(int64_t b0, int64_t b1, int64_t nb, int64_t &ind0,
                                                  int64_t &indf)
{
    b0 = imin64(b0, nb);
    b0 = imax64(b0, 0);
    if (b1 > nb) {
        ind0 = b0;
        indf = nb;
    } else if (b1 < 0) {
        ind0 = 0;
        indf = b0 + 1;
    } else if (b0 == b1) {
        ind0 = b0;
        indf = b0 + 1;
    } else {
        ind0 = imin64(b0, b1);
        indf = imax64(b0, b1);
    }
}

// ---------------------------------------------------------------------------
// SSWLC elastic energy (chromo/polymers.pyx:1148-1346)
// ---------------------------------------------------------------------------

struct Bond {
    double eb, ep, eperp, gam, eta;
};

__device__ __forceinline__ Bond load_bond(const double *bonds_, int64_t N, int64_t b)
{
    const GLOBAL_AS double *bonds = gmem(bonds_);
    const int64_t M = N - 1;
    Bond o;
    o.eb = bonds[b];
    o.ep = bonds[M + b];
    o.eperp = bonds[2 * M + b];
    o.gam = bonds[3 * M + b];
    o.eta = bonds[4 * M + b];
    return o;
}

__device__ __forceinline__ double dot3(const double *a, const double *b)
{
    double d = 0;
    d += a[0] * b[0];
    d += a[1] * b[1];
    d += a[2] * b[2];
    return d;
}

// E_pair, polymers.pyx:1148-1175 ((x)**2 is pow(x, 2.0) == x*x under glibc)
__device__ __forceinline__ double E_pair(const Bond &bd, const double *bend, double par,
                                         const double *perp)
{
    double q = par - bd.gam;
    return 0.5 * bd.eb * dot3(bend, bend) + 0.5 * bd.ep * (q * q) + 0.5 * bd.eperp * dot3(perp, perp);
}

// ---- twist (SSTWLC, chromo/polymers.pyx:1889-2320) ----------------------------
//
// The twist rows of a twisting chain's bond table: 5 eps_twist = lt / (delta lp),
// 6 the natural twist of the bond, bead_length * (2 pi / 10.5) / 0.332.

__device__ __forceinline__ void load_twist(const double *bonds_, int64_t N, int64_t b, double &et, double &tw0)
{
    const GLOBAL_AS double *bonds = gmem(bonds_);
    const int64_t M = N - 1;
    et = bonds[5 * M + b];
    tw0 = bonds[6 * M + b];
}

// compute_twist_angle_omega (chromo/polymers.pyx:3427-3461): with t1 = t2 x t3,
// omega = atan2(t2_0 . t1_1 - t1_0 . t2_1, t1_0 . t1_1 + t2_0 . t2_1).
__device__ __forceinline__ double twist_omega(const double *t2_0, const double *t3_0, const double *t2_1,
                                              const double *t3_1)
{
    const double t1_0[3] = {t2_0[1] * t3_0[2] - t2_0[2] * t3_0[1], t2_0[2] * t3_0[0] - t2_0[0] * t3_0[2],
                            t2_0[0] * t3_0[1] - t2_0[1] * t3_0[0]};
    const double t1_1[3] = {t2_1[1] * t3_1[2] - t2_1[2] * t3_1[1], t2_1[2] * t3_1[0] - t2_1[0] * t3_1[2],
                            t2_1[0] * t3_1[1] - t2_1[1] * t3_1[0]};
    return atan2(dot3(t2_0, t1_1) - dot3(t1_0, t2_1), dot3(t1_0, t1_1) + dot3(t2_0, t2_1));
}

// Python's float floor division a // b (b > 0): the exact remainder fmod(a, b),
// the quotient (a - mod) / b, one step down when the remainder has the other sign,
// then the nearest integer of that quotient.
__device__ __forceinline__ double py_floordiv(double a, double b)
{
    double mod = fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0 && ((b < 0) != (mod < 0))) div -= 1.0;
    if (div == 0) return copysign(0.0, a / b);
    double fl = floor(div);
    if (div - fl > 0.5) fl += 1.0;
    return fl;
}

// The twist term of E_pair_with_twist (polymers.pyx:2072-2082): the twist angle's
// departure from the natural twist, wrapped into [-pi, pi).
__device__ __forceinline__ double twist_energy(double et, double tw0, double omega)
{
    const double two_pi = 2 * M_PI;
    double d = omega - tw0;
    d -= two_pi * py_floordiv(d + M_PI, two_pi);
    return 0.5 * et * (d * d);
}

// bead_pair_dE_poly_forward, polymers.pyx:1177-1277; tw_t / tw_c: the twist terms
// of E_pair_with_twist for a twisting chain (polymers.pyx:2123-2175), else 0.
__device__ __forceinline__ double pair_dE_forward(const Bond &bd, const double *r0, const double *r1,
                                  const double *tr1, const double *t30, const double *t31,
                                  const double *tt31, double tw_t = 0, double tw_c = 0)
{
    double dr_t[3], dr[3], drp_t[3], drp[3], be_t[3], be[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = tr1[i] - r0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(t30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - t30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        be_t[i] = tt31[i] - t30[i] - drp_t[i] * bd.eta;
        be[i] = t31[i] - t30[i] - drp[i] * bd.eta;
    }
    return (E_pair(bd, be_t, par_t, drp_t) + tw_t) - (E_pair(bd, be, par, drp) + tw_c);
}

// bead_pair_dE_poly_reverse, polymers.pyx:1279-1346 (twist: polymers.pyx:2177-2248)
__device__ __forceinline__ double pair_dE_reverse(const Bond &bd, const double *r0, const double *tr0,
                                  const double *r1, const double *t30, const double *tt30,
                                  const double *t31, double tw_t = 0, double tw_c = 0)
{
    double dr_t[3], dr[3], drp_t[3], drp[3], be_t[3], be[3];
    for (int i = 0; i < 3; ++i) {
        dr_t[i] = r1[i] - tr0[i];
        dr[i] = r1[i] - r0[i];
    }
    double par_t = dot3(tt30, dr_t), par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp_t[i] = dr_t[i] - tt30[i] * par_t;
        drp[i] = dr[i] - t30[i] * par;
        be_t[i] = t31[i] - tt30[i] - drp_t[i] * bd.eta;
        be[i] = t31[i] - t30[i] - drp[i] * bd.eta;
    }
    return (E_pair(bd, be_t, par_t, drp_t) + tw_t) - (E_pair(bd, be, par, drp) + tw_c);
}

// Twist terms of the forward pair (bead a = b0 - 1 fixed, bead b0 moved: its
// current frame o[3..8], trial w[3..8]) and of the reverse pair (bead indf - 1
// moved, bead a = indf fixed), with bead a's t3 / t2 read from the chain.
__device__ __forceinline__ void twist_forward(const double *bonds, const double *pos, int64_t N, int64_t a,
                                              const double *o, const double *w, double &tw_t, double &tw_c)
{
    double et, tw0, t3a[3], t2a[3];
    load_twist(bonds, N, a, et, tw0);
    const GLOBAL_AS double *P = gmem(pos);
    for (int j = 0; j < 3; ++j) {
        t3a[j] = P[(3 + j) * N + a];
        t2a[j] = P[(6 + j) * N + a];
    }
    tw_c = twist_energy(et, tw0, twist_omega(t2a, t3a, &o[6], &o[3]));
    tw_t = twist_energy(et, tw0, twist_omega(t2a, t3a, &w[6], &w[3]));
}

__device__ __forceinline__ void twist_reverse(const double *bonds, const double *pos, int64_t N, int64_t a,
                                              const double *o, const double *w, double &tw_t, double &tw_c)
{
    double et, tw0, t3a[3], t2a[3];
    load_twist(bonds, N, a - 1, et, tw0);
    const GLOBAL_AS double *P = gmem(pos);
    for (int j = 0; j < 3; ++j) {
        t3a[j] = P[(3 + j) * N + a];
        t2a[j] = P[(6 + j) * N + a];
    }
    tw_c = twist_energy(et, tw0, twist_omega(&o[6], &o[3], t2a, t3a));
    tw_t = twist_energy(et, tw0, twist_omega(&w[6], &w[3], t2a, t3a));
}

// bead_binding_dE, polymers.pyx:1408-1538, with host-built -log(single-site Z)
// tables F[b][marks][state].
template <int NB>
__device__ __forceinline__ double bead_binding_dE(const chromo_replica *R, const int8_t *mods, const int8_t *st,
                                  const int8_t *stt)
{
    constexpr int B = NB;
    double dE = 0;
    if (R->max_binders != -1) {
        int64_t tot = 0;
        for (int h = 0; h < B; ++h) tot += stt[h];
        if (tot > R->max_binders) dE += E_HUGE_POLY * (double)(tot - R->max_binders);
        tot = 0;
        for (int h = 0; h < B; ++h) tot += st[h];
        if (tot > R->max_binders) dE -= E_HUGE_POLY * (double)(tot - R->max_binders);
    }
    const double f = R->mu_adjust;
    for (int b = 0; b < B; ++b) {
        dE += R->F[b][mods[b]][stt[b]];
        dE -= R->F[b][mods[b]][st[b]];
        const double mu = R->mu[b];
        if (mu > 0) {
            dE -= (double)stt[b] * (mu * (-f + 2));
            dE += (double)st[b] * (mu * (-f + 2));
        } else {
            dE -= (double)stt[b] * mu * f;
            dE += (double)st[b] * mu * f;
        }
    }
    return dE;
}

// Accessible volume of voxel v: the divisor of a bead's weights in it and the
// volume of its chi term (fields.pyx:1476, 1826); vol_bin unless the field has
// per-voxel volumes (spherical confinement without assume_fully_accessible).
__device__ __forceinline__ double voxel_vol(const chromo_replica *R, int64_t v)
{
    const double *av = R->access_vol;
    return av ? av[v] : R->vol_bin;
}

// ---------------------------------------------------------------------------
// Field dE terms (chromo/fields.pyx:1675-1842)
// ---------------------------------------------------------------------------

// One voxel's share of get_dE_binders_and_beads (fields.pyx:1723-1750) and of
// nonspecific_interact_dE (fields.pyx:1792-1842), added to a lane's running sums.
// r: current densities, a: delta densities (B+1 species each).
template <int NB>
__device__ __forceinline__ void voxel_dE_terms(const chromo_replica *R, int64_t v, const double *r, const double *a,
                                               double *S2, double *SX, double &NS)
{
    for (int j = 0; j < NB; ++j) {
        double aj = r[j + 1] + a[j + 1];
        double d2 = aj * aj - r[j + 1] * r[j + 1];
        if (fabs(d2) < 1E-18) d2 = 0;
        S2[j] += d2;
        for (int kk = 0; kk < NB; ++kk) {
            double dx = (aj * (r[kk + 1] + a[kk + 1])) - (r[j + 1] * r[kk + 1]);
            if (fabs(dx) < 1E-18) dx = 0;
            SX[j * NB + kk] += dx;
        }
    }
    const double V = R->bead_vol, vf = R->vf_limit, chi = R->chi, vol_bin = voxel_vol(R, v);
    double vf0 = r[0] * V;
    double vf1 = vf0 + (a[0] * V);
    // over vf_limit before and after: the reference's running sum adds +1e99 vf1
    // and -1e99 vf0 in CPython set order, which drops (vf1 == vf0) or rounds away
    // the small terms summed before this voxel; the result depends on that order.
    // Counted, one atomic per (rare) voxel event.
    if (vf0 > vf && vf1 > vf) atomicAdd((unsigned long long *)&R->counters[CHROMO_CNT_TIE_HUGE], 1ull);
    if (vf1 > vf) NS += E_HUGE_FIELD * vf1;
    else NS += chi * (vol_bin / V) * (vf1 * vf1);
    if (vf0 > vf) NS -= E_HUGE_FIELD * vf0;
    else NS -= chi * (vol_bin / V) * (vf0 * vf0);
}

// A lane's running sums weighted by the binder prefactors and cross-talk
// coefficients; summed over the wavefront this is the binder-pair, cross-talk and
// chi part of the field dE.
template <int NB>
__device__ __forceinline__ double lane_field_share(const chromo_replica *R, const double *S2,
                                                   const double *SX, double NS)
{
    double p = 0;
    for (int i = 0; i < NB; ++i) p += R->pref[i] * S2[i];
    for (int i = 0; i < NB; ++i)
        for (int j = 0; j < NB; ++j) p += R->cross[i * NB + j] * SX[i * NB + j];
    p += NS;
    return p;
}

// UniformDensityField.compute_dE's total (fields.pyx:1209-1233): confinement
// (trial minus current), the wave-summed share, the doubly-bound terms.
template <int NB>
__device__ __forceinline__ double field_dE_total(const chromo_replica *R, double share, const int *dbl_t,
                                                 const int *dbl, int out_t, int out_c, bool state_change)
{
    double dE = 0;
    if (!state_change) {
        dE += (double)out_t * E_HUGE_FIELD;
        dE -= (double)out_c * E_HUGE_FIELD;
    }
    double dEb = share;
    for (int i = 0; i < NB; ++i) dEb += R->intra[i] * (double)(dbl_t[i] - dbl[i]);
    dE += dEb;
    return dE;
}

// ---------------------------------------------------------------------------
// Field geometry (chromo/fields.pyx:1434-1673)
// ---------------------------------------------------------------------------

// The grid fields voxel_lower reads, taken out of the replica record once per
// evaluation: the record lives in LDS, and a field read there after the
// evaluation's own LDS stores would be read again at every use.
struct VoxGrid {
    double width[3], half_width[3], half_step[3], dxyz[3];
    int64_t n[3];
    int fast;
    double half_points, sub_width[3];
    int64_t n_sub[3];
};

__device__ __forceinline__ VoxGrid vox_grid(const chromo_replica *R)
{
    VoxGrid G;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        G.width[j] = R->width[j];
        G.half_width[j] = R->half_width[j];
        G.half_step[j] = R->half_step[j];
        G.dxyz[j] = R->dxyz[j];
        G.sub_width[j] = R->sub_width[j];
        G.n_sub[j] = R->n_sub[j];
    }
    G.n[0] = R->nx;
    G.n[1] = R->ny;
    G.n[2] = R->nz;
    G.fast = R->fast_field;
    G.half_points = R->half_points;
    return G;
}

// Lower voxel indices and lower-corner weights of one position.
//
// `ties` gains one per coordinate within TIE_VOXEL_EPS voxel of a voxel plane,
// where an ulp of difference in the position (device sin / cos / acos / sqrt
// against glibc's) could change the floor.  The caller adds the count to
// CHROMO_CNT_TIE_VOXEL; keeping the atomic out of here lets the three axes (and
// a caller's two positions) run as independent chains.
__device__ __forceinline__ void voxel_lower(const VoxGrid &G, const double *pos, int64_t ix[3], double w[3],
                                            int &ties)
{
    if (G.fast) {
        // the position's sub-bin floor((x + W/2) / sbw) mod (n_points n) and that
        // sub-bin's lower voxel and weight, from init_fast_field's expressions
        // (fields.pyx:614-671, 1255-1300)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double sw = G.sub_width[j];
            const double qs = (pos[j] + G.half_width[j]) / sw;
            const double sb = isfinite(qs) ? floor(qs) : 0.0;
            ties += fabs(qs - rint(qs)) < TIE_VOXEL_EPS;
            int64_t i = (fabs(sb) < 0x1p62) ? (int64_t)sb % G.n_sub[j] : 0;
            if (i < 0) i += G.n_sub[j];
            const double u = (((double)i - G.half_points) * sw) / G.dxyz[j];
            const double ind = floor(u);
            ix[j] = (ind == -1.0) ? G.n[j] - 1 : (int64_t)ind;
            w[j] = 1 - (u - ind);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double x = pymod(pos[j] + G.half_width[j], G.width[j]) - G.half_step[j];
        double q = x / G.dxyz[j];
        // finite coordinates give q in [-0.5, n - 0.5); a non-finite one is pinned to
        // voxel 0 so that no index can leave the grid
        if (!(q >= -1.0 && q < (double)G.n[j])) q = 0.0;
        const double fq = floor(q);
        ix[j] = (fq == -1.0) ? G.n[j] - 1 : (int64_t)(int32_t)fq;
        w[j] = 1 - (q - fq);
        ties += fabs(q - rint(q)) < TIE_VOXEL_EPS;
    }
}

__device__ __forceinline__ void voxel_lower(const chromo_replica *R, const double *pos, int64_t ix[3],
                                            double w[3], int &ties)
{
    voxel_lower(vox_grid(R), pos, ix, w, ties);
}

// voxel_lower for one lane on its own: the tie count goes straight to the counter.
__device__ __forceinline__ void voxel_lower_1(const chromo_replica *R, const double *pos, int64_t ix[3],
                                              double w[3])
{
    int ties = 0;
    voxel_lower(R, pos, ix, w, ties);
    if (ties) atomicAdd((unsigned long long *)&R->counters[CHROMO_CNT_TIE_VOXEL], (unsigned long long)ties);
}

// Adds the lanes' tie counts to CHROMO_CNT_TIE_VOXEL with one atomic.  Whole wave.
__device__ __forceinline__ void count_ties(const chromo_replica *R, int ties)
{
    if (!wave_any(ties != 0)) return;
    int tot;
    wave_excl_scan(ties, tot);
    if (lane_id() == 0) atomicAdd((unsigned long long *)&R->counters[CHROMO_CNT_TIE_VOXEL], (unsigned long long)tot);
}

// Voxel super-index of corner l (x fastest) with the wrap of
// precompute_ind_xyz_to_super (fields.pyx:673-685).
__device__ __forceinline__ int64_t corner_voxel(const chromo_replica *R, const int64_t ix[3], int l)
{
    int64_t x = ix[0] + (l & 1), y = ix[1] + ((l >> 1) & 1), z = ix[2] + ((l >> 2) & 1);
    if (x >= R->nx) x -= R->nx;
    if (y >= R->ny) y -= R->ny;
    if (z >= R->nz) z -= R->nz;
    return x + y * R->nx + z * R->nx * R->ny;
}

// Corner weight products of _generate_weight_vector_with_trial (fields.pyx:1524-1598).
__device__ __forceinline__ double corner_weight(int l, double wx, double wy, double wz)
{
    double a = (l & 1) ? (1 - wx) : wx;
    double b = (l & 2) ? (1 - wy) : wy;
    double c = (l & 4) ? (1 - wz) : wz;
    return a * b * c;
}

}  // namespace chromo
