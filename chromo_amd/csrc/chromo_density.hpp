// chromo_density.hpp -- deterministic voxel-density scatter shared by the MC
// initialisation and the energy / adjoint kernels.
//
// UniformDensityField.update_all_densities_for_all_polymers
// (chromo/fields.pyx:2041-2106) adds every bead's eight trilinear weights to
// the grid in bead order.  To reproduce those per-voxel sums bit for bit in
// parallel, beads are bucketed by their lower-corner voxel (counting sort with
// sorted buckets) and every voxel then gathers from the <= 8 buckets that reach
// it, merging them in ascending bead index.  No floating-point atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "chromo_device.hpp"

namespace chromo {
namespace {

// Grid of a UniformDensityField (fields.pyx:536-575).
struct GridGeom {
    int64_t nx, ny, nz;
    double width[3], dxyz[3], half_width[3], half_step[3];
    double vol_bin;
    const double *access_vol = nullptr;   // per-voxel accessible volumes, or nullptr: vol_bin
};

// Bead coordinate c of bead b is p[c * sc + b * sb].
struct PosView {
    const double *p;
    int64_t sc, sb;
};

__device__ __forceinline__ void grid_lower(const GridGeom &G, double x0, double x1, double x2,
                                           int64_t ix[3], double w[3])
{
    const int64_t nn[3] = {G.nx, G.ny, G.nz};
    const double xs[3] = {x0, x1, x2};
    for (int j = 0; j < 3; ++j) {
        double x = pymod(xs[j] + G.half_width[j], G.width[j]) - G.half_step[j];
        double q = x / G.dxyz[j];
        if (!(q >= -1.0 && q < (double)nn[j])) q = 0.0;
        int64_t ind = (int64_t)floor(q);
        ix[j] = (ind == -1) ? nn[j] - 1 : ind;
        w[j] = 1 - (q - (double)ind);
    }
}

// What the gather reads of one bead, in one 32-byte record (one cache line per
// merge step instead of one per weight and per state row).
struct BeadRec {
    double w[3];
    int8_t st[CHROMO_MAX_BINDERS];
    int32_t pad;
};
static_assert(sizeof(BeadRec) == 32, "one bead record is 32 bytes");

__global__ void k_dens_cells(GridGeom G, int64_t N, PosView P, int B, const int8_t *states, int32_t *cell,
                             BeadRec *brec, int32_t *cnt)
{
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    int64_t ix[3];
    double w[3];
    grid_lower(G, P.p[b * P.sb], P.p[P.sc + b * P.sb], P.p[2 * P.sc + b * P.sb], ix, w);
    int32_t c = (int32_t)(ix[0] + ix[1] * G.nx + ix[2] * G.nx * G.ny);
    cell[b] = c;
    BeadRec q;
    q.w[0] = w[0];
    q.w[1] = w[1];
    q.w[2] = w[2];
#pragma unroll
    for (int m = 0; m < CHROMO_MAX_BINDERS; ++m) q.st[m] = m < B ? states[(int64_t)m * N + b] : (int8_t)0;
    q.pad = 0;
    brec[b] = q;
    atomicAdd(&cnt[c], 1);
}

// exclusive scan: one block of 1024 threads per 2048 elements, block totals in bsum
__global__ void k_scan_blocks(const int32_t *in, int32_t *out, int32_t *bsum, int64_t n)
{
    __shared__ int32_t s[2048];
    int64_t base = (int64_t)blockIdx.x * 2048;
    int t = threadIdx.x;
    for (int i = t; i < 2048; i += 1024) s[i] = (base + i < n) ? in[base + i] : 0;
    __syncthreads();
    for (int d = 1; d < 2048; d <<= 1) {
        int32_t v0 = 0, v1 = 0;
        int i0 = t, i1 = t + 1024;
        if (i0 >= d) v0 = s[i0 - d];
        if (i1 >= d) v1 = s[i1 - d];
        __syncthreads();
        s[i0] += v0;
        s[i1] += v1;
        __syncthreads();
    }
    for (int i = t; i < 2048; i += 1024) {
        if (base + i < n) out[base + i] = s[i] - in[base + i];
    }
    if (t == 0) bsum[blockIdx.x] = s[2047];
}

__global__ void k_scan_add(int32_t *out, const int32_t *boff, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += boff[i / 2048];
}

__global__ void k_dens_place(int64_t N, const int32_t *cell, const int32_t *off, int32_t *fill,
                             int32_t *list)
{
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    int32_t c = cell[b];
    int32_t pos = off[c] + atomicAdd(&fill[c], 1);
    list[pos] = (int32_t)b;
}

// Puts each bucket (filled in atomic order by k_dens_place) in bead order.  Small
// buckets, the usual case, by insertion sort; large ones (dense voxels: compact
// or collapsed chains, coarse grids) by heapsort, so the cost stays k log k.
__device__ __forceinline__ void sift_down(int32_t *a, int32_t i, int32_t n)
{
    const int32_t x = a[i];
    for (;;) {
        int32_t c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && a[c + 1] > a[c]) ++c;
        if (a[c] <= x) break;
        a[i] = a[c];
        i = c;
    }
    a[i] = x;
}

__global__ void k_dens_sort(int64_t nbins, const int32_t *off, const int32_t *cnt, int32_t *list)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nbins) return;
    const int32_t s0 = off[c], k = cnt[c];
    int32_t *a = list + s0;
    if (k <= 32) {
        for (int32_t p = 1; p < k; ++p) {
            const int32_t x = a[p];
            int32_t q = p - 1;
            while (q >= 0 && a[q] > x) {
                a[q + 1] = a[q];
                --q;
            }
            a[q + 1] = x;
        }
        return;
    }
    for (int32_t i = k / 2 - 1; i >= 0; --i) sift_down(a, i, k);
    for (int32_t e = k - 1; e > 0; --e) {
        const int32_t t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift_down(a, 0, e);
    }
}

// The buckets that reach voxel v: cell[l] holds the beads of which v is corner l;
// a bucket reached by several corners (grids one or two voxels wide) is kept once,
// in its first slot, whose mask collects every such corner.
__device__ __forceinline__ void voxel_cells(const GridGeom &G, int64_t v, int32_t (&cell)[8], int (&msk)[8])
{
    const int64_t nx = G.nx, ny = G.ny, nz = G.nz;
    const int64_t x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        int64_t cx = x - (l & 1), cy = y - ((l >> 1) & 1), cz = z - ((l >> 2) & 1);
        if (cx < 0) cx += nx;
        if (cy < 0) cy += ny;
        if (cz < 0) cz += nz;
        cell[l] = (int32_t)(cx + cy * nx + cz * nx * ny);
        msk[l] = 0;
    }
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        bool placed = false;
#pragma unroll
        for (int k = 0; k < l; ++k) {
            if (!placed && cell[k] == cell[l]) {
                msk[k] |= 1 << l;
                placed = true;
            }
        }
        if (!placed) msk[l] |= 1 << l;
    }
}

// On large grids (at least GATHER_BIG_BLOCKS blocks of 1024 voxels) one gather
// block takes 1024 consecutive voxels and hands them to its threads in order of
// merge length (the beads of their buckets, keys clamped to GATHER_KEYS - 1), so
// that the lanes of a wavefront have merges of similar length.  Smaller grids run
// in one round of 256-thread blocks, voxel t of the block on thread t: there the
// ranking's extra phase costs more than the balance saves.
constexpr int GATHER_KEYS = 64, GATHER_BIG_BLOCKS = 512;

// density[v * Bp1 + m]: m = 0 beads, m >= 1 binder m-1 (weight x state).  Voxel v
// is corner l of the beads whose lower-corner cell is v - (l & 1, l >> 1 & 1,
// l >> 2 & 1), wrapped.  Those <= 8 sorted buckets are merged in ascending bead
// index, and the corners of one bead that land on v (several only on grids one or
// two voxels wide, where buckets coincide) are folded in corner order: the
// reference's bead loop.  Every per-bucket value is indexed at compile time, so
// the merge state stays in registers.  Which thread merges which voxel of the
// block does not change any voxel's sum.
template <int T, bool RANK>
__global__ void __launch_bounds__(T)
k_dens_gather(GridGeom G, int Bp1, const int32_t *off, const int32_t *cnt, const int32_t *list,
              const BeadRec *brec, double *density, int cleanup)
{
    constexpr int32_t NONE = 0x7fffffff;
    __shared__ int32_t hist[GATHER_KEYS], hoff[GATHER_KEYS], perm[T];
    const int t = threadIdx.x;
    const int64_t nbv = G.nx * G.ny * G.nz, base = (int64_t)blockIdx.x * T;
    const int nreal = (int)(nbv - base < T ? nbv - base : T);
    int32_t cell[8], pos[8], end[8], head[8];
    int msk[8];
    int64_t v = base + t;
    if constexpr (RANK) {
        if (t < GATHER_KEYS) hist[t] = 0;
        __syncthreads();
        int key = 0;
        if (t < nreal) {
            voxel_cells(G, base + t, cell, msk);
            int64_t n = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) n += msk[k] ? cnt[cell[k]] : 0;
            key = (int)(n < GATHER_KEYS - 1 ? n : GATHER_KEYS - 1);
            atomicAdd(&hist[key], 1);
        }
        __syncthreads();
        if (t < GATHER_KEYS) {
            static_assert(GATHER_KEYS == 64, "one wavefront scans the key counts");
            const int c = hist[t];
            int incl = c;
            for (int o = 1; o < GATHER_KEYS; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (t >= o) incl += y;
            }
            hoff[t] = incl - c;
        }
        __syncthreads();
        if (t < nreal) perm[atomicAdd(&hoff[key], 1)] = t;
        __syncthreads();
        if (t < nreal) v = base + perm[t];
    }
    if (t >= nreal) return;
    voxel_cells(G, v, cell, msk);
    const double vvol = G.access_vol ? G.access_vol[v] : G.vol_bin;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        pos[k] = msk[k] ? off[cell[k]] : 0;
        end[k] = msk[k] ? pos[k] + cnt[cell[k]] : 0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) head[k] = pos[k] < end[k] ? list[pos[k]] : NONE;
    double acc[CHROMO_MAX_BINDERS + 1];
#pragma unroll
    for (int m = 0; m <= CHROMO_MAX_BINDERS; ++m) acc[m] = 0;
    for (;;) {
        // the smallest bead index at the buckets' heads; a bead lies in one bucket only
        int32_t bb = NONE;
        int bm = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (head[k] < bb) {
                bb = head[k];
                bm = msk[k];
            }
        }
        if (bb == NONE) break;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (head[k] == bb) {
                ++pos[k];
                head[k] = pos[k] < end[k] ? list[pos[k]] : NONE;
            }
        }
        const BeadRec q = brec[bb];
        const double wx = q.w[0], wy = q.w[1], wz = q.w[2];
        double st[CHROMO_MAX_BINDERS];
#pragma unroll
        for (int m = 1; m <= CHROMO_MAX_BINDERS; ++m) st[m - 1] = m < Bp1 ? (double)q.st[m - 1] : 0.0;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            if (!(bm >> l & 1)) continue;
            const double d = corner_weight(l, wx, wy, wz) / vvol;
            acc[0] += d;
#pragma unroll
            for (int m = 1; m <= CHROMO_MAX_BINDERS; ++m)
                if (m < Bp1) acc[m] += d * st[m - 1];
        }
    }
#pragma unroll
    for (int m = 0; m <= CHROMO_MAX_BINDERS; ++m) {
        if (m < Bp1) {
            double a = acc[m];
            if (cleanup && fabs(a) < 1E-18) a = 0;
            density[v * Bp1 + m] = a;
        }
    }
}

// The energy adjoint's placement: bead b's record goes straight into its cell's
// bucket, in the order the atomic counter hands out (the fixed-point gather below
// does not depend on it).
__global__ void k_dens_place_rec(int64_t N, const int32_t *cell, const int32_t *off, int32_t *fill,
                                 const BeadRec *brec, BeadRec *brec_bucketed)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    const int32_t c = cell[b];
    brec_bucketed[off[c] + atomicAdd(&fill[c], 1)] = brec[b];
}

// The density of the energy adjoint, which needs a deterministic sum but not the
// reference's bead-order fold.  Each corner weight w (in [0, 1]) is turned into
// the integer floor(w 2^60) (within 2^-60 of w) and the integers are summed in
// 128 bits, times the binder state (0..CHROMO_MAX_SITES) for the binder rows: the
// sums are exact, so they do not depend on the order of the beads in a bucket,
// and the atomic placement needs no sort.  The voxel's total is rounded to double
// once and divided by its volume.  Eight lanes per voxel, lane k summing bucket k
// (the beads whose lower-corner cell is v - (k & 1, k >> 1 & 1, k >> 2 & 1),
// wrapped; on grids one or two voxels wide, where such cells coincide, the first
// corner of a cell takes its bucket and adds every corner it stands for).  The
// bucket sums meet in LDS; lane m < B + 1 of the voxel adds species m's eight and
// writes it.  A block holds GF_VOX voxels.  The sums stay below 2^96 (2^31 beads,
// eight corners, states up to 3).
constexpr int GF_VOX = 32;

__device__ __forceinline__ uint64_t fixed60(double w) { return (uint64_t)(w * 0x1p60); }

__device__ __forceinline__ double from_fixed60(unsigned __int128 a)
{
    return ((double)(uint64_t)(a >> 64) * 0x1p64 + (double)(uint64_t)a) * 0x1p-60;
}

template <int NB>
__device__ __forceinline__ void add_fixed(unsigned __int128 (&acc)[NB + 1], uint64_t w, const int8_t *st)
{
    acc[0] += w;
#pragma unroll
    for (int m = 1; m <= NB; ++m) acc[m] += w * (uint64_t)(uint8_t)st[m - 1];
}

template <int NB>
__global__ void __launch_bounds__(8 * GF_VOX)
k_dens_gather_fixed(GridGeom G, const int32_t *off, const int32_t *cnt, const BeadRec *brec_bucketed,
                    double *density)
{
    constexpr int Bp1 = NB + 1;
    __shared__ unsigned __int128 part[8 * GF_VOX][Bp1];
    const int t = threadIdx.x, k = t & 7, j = t >> 3;
    // 32-bit index arithmetic: the grids hold fewer than 2^31 voxels
    const int32_t nx = (int32_t)G.nx, ny = (int32_t)G.ny, nz = (int32_t)G.nz, nbv = nx * ny * nz;
    const int32_t v = (int32_t)blockIdx.x * GF_VOX + j;
    unsigned __int128 acc[Bp1];
#pragma unroll
    for (int m = 0; m < Bp1; ++m) acc[m] = 0;
    if (v < nbv) {
        const uint32_t q = (uint32_t)v / (uint32_t)nx;
        const int32_t x = v - (int32_t)q * nx, y = (int32_t)(q % (uint32_t)ny), z = (int32_t)(q / (uint32_t)ny);
        auto cell_of = [&](int l) {
            int32_t cx = x - (l & 1), cy = y - ((l >> 1) & 1), cz = z - ((l >> 2) & 1);
            if (cx < 0) cx += nx;
            if (cy < 0) cy += ny;
            if (cz < 0) cz += nz;
            return cx + nx * (cy + ny * cz);
        };
        const int32_t ck = cell_of(k);
        int msk = 0;
        bool dup = false;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            if (cell_of(l) != ck) continue;
            if (l < k) dup = true;
            else msk |= 1 << l;
        }
        const int32_t pos = off[ck], n = dup ? 0 : cnt[ck];
        if (msk == 1 << k) {
            for (int32_t i = 0; i < n; ++i) {
                const BeadRec r = brec_bucketed[pos + i];
                add_fixed<NB>(acc, fixed60(corner_weight(k, r.w[0], r.w[1], r.w[2])), r.st);
            }
        } else {
            for (int32_t i = 0; i < n; ++i) {
                const BeadRec r = brec_bucketed[pos + i];
                for (int l = 0; l < 8; ++l)
                    if (msk >> l & 1) add_fixed<NB>(acc, fixed60(corner_weight(l, r.w[0], r.w[1], r.w[2])), r.st);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < Bp1; ++m) part[t][m] = acc[m];
    __syncthreads();
    if (v < nbv && k < Bp1) {
        unsigned __int128 a = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) a += part[8 * j + c][k];
        const double vvol = G.access_vol ? G.access_vol[v] : G.vol_bin;
        density[(int64_t)v * Bp1 + k] = from_fixed60(a) / vvol;
    }
}

// Scratch for one scatter: sized by the largest chain and grid.
struct DensityScratch {
    int32_t *cell = nullptr, *list = nullptr, *cnt = nullptr, *off = nullptr, *fill = nullptr;
    int32_t *bsum = nullptr, *boff = nullptr, *bsum2 = nullptr;
    BeadRec *brec = nullptr, *brec_bucketed = nullptr;
    int64_t maxN = 0, maxBins = 0;
};

constexpr size_t SCRATCH_ALIGN = 256;

inline size_t align_scratch(size_t o) { return (o + SCRATCH_ALIGN - 1) / SCRATCH_ALIGN * SCRATCH_ALIGN; }

// The scatter's scratch carved from one buffer at base (nullptr: sizes only);
// returns the bytes used, a multiple of SCRATCH_ALIGN.  bucketed_records: room for
// the bead records placed by bucket for the fixed-point gather.
inline size_t density_scratch_carve(DensityScratch &S, char *base, int64_t maxN, int64_t maxBins,
                                    bool bucketed_records)
{
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o = align_scratch(o + bytes);
        return p;
    };
    S.maxN = maxN;
    S.maxBins = maxBins;
    S.cell = (int32_t *)take(sizeof(int32_t) * maxN);
    S.list = (int32_t *)take(sizeof(int32_t) * maxN);
    S.brec = (BeadRec *)take(sizeof(BeadRec) * maxN);
    S.brec_bucketed = bucketed_records ? (BeadRec *)take(sizeof(BeadRec) * maxN) : nullptr;
    S.cnt = (int32_t *)take(sizeof(int32_t) * maxBins);
    S.off = (int32_t *)take(sizeof(int32_t) * maxBins);
    S.fill = (int32_t *)take(sizeof(int32_t) * maxBins);
    S.bsum = (int32_t *)take(sizeof(int32_t) * 2048);
    S.boff = (int32_t *)take(sizeof(int32_t) * 2048);
    S.bsum2 = (int32_t *)take(sizeof(int32_t) * 2);
    return o;
}

inline hipError_t density_scratch_alloc(DensityScratch &S, int64_t maxN, int64_t maxBins,
                                        hipStream_t st)
{
    hipError_t e;
    S.maxN = maxN;
    S.maxBins = maxBins;
#define CHROMO_ALLOC(ptr, bytes) \
    if ((e = hipMallocAsync((void **)&ptr, bytes, st)) != hipSuccess) return e
    CHROMO_ALLOC(S.cell, sizeof(int32_t) * maxN);
    CHROMO_ALLOC(S.list, sizeof(int32_t) * maxN);
    CHROMO_ALLOC(S.brec, sizeof(BeadRec) * maxN);
    CHROMO_ALLOC(S.cnt, sizeof(int32_t) * maxBins);
    CHROMO_ALLOC(S.off, sizeof(int32_t) * maxBins);
    CHROMO_ALLOC(S.fill, sizeof(int32_t) * maxBins);
    CHROMO_ALLOC(S.bsum, sizeof(int32_t) * 2048);
    CHROMO_ALLOC(S.boff, sizeof(int32_t) * 2048);
    CHROMO_ALLOC(S.bsum2, sizeof(int32_t) * 2);
#undef CHROMO_ALLOC
    return hipSuccess;
}

inline void density_scratch_free(DensityScratch &S, hipStream_t st)
{
    int32_t *ip[] = {S.cell, S.list, S.cnt, S.off, S.fill, S.bsum, S.boff, S.bsum2};
    for (auto p : ip)
        if (p) (void)hipFreeAsync(p, st);
    if (S.brec) (void)hipFreeAsync(S.brec, st);
    if (S.brec_bucketed) (void)hipFreeAsync(S.brec_bucketed, st);
    S = DensityScratch();
}

// Clears two int32 arrays of n elements.
__global__ void k_fill2_i32(int32_t *p, int32_t *q, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0;
        q[i] = 0;
    }
}

// Scatter N beads into density (n_bins x Bp1).  Requires N <= S.maxN,
// n_bins <= S.maxBins <= 2048 * 2048.  fixed_order: the adjoint's exact
// fixed-point sum over unsorted buckets (k_dens_place_rec, k_dens_gather_fixed; S
// carved with bucketed_records) instead of the reference's bead-order fold; cleanup
// (the 1e-18 rule of the initial densities) applies to the fold only.
inline hipError_t density_scatter(const GridGeom &G, int64_t N, int Bp1, PosView P,
                                  const int8_t *states, double *density, int cleanup,
                                  DensityScratch &S, hipStream_t st, bool fixed_order = false)
{
    const int64_t nb = G.nx * G.ny * G.nz;
    hipError_t e;
    unsigned gb = (unsigned)((N + 255) / 256);
    unsigned gv = (unsigned)((nb + 255) / 256);
    // cleared by a kernel rather than hipMemsetAsync: a caller may capture this
    // sequence in a HIP graph
    hipLaunchKernelGGL(k_fill2_i32, dim3(gv), dim3(256), 0, st, S.cnt, S.fill, nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_dens_cells, dim3(gb), dim3(256), 0, st, G, N, P, Bp1 - 1, states, S.cell, S.brec,
                       S.cnt);
    int64_t nbl = (nb + 2047) / 2048;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nbl), dim3(1024), 0, st, S.cnt, S.off, S.bsum, nb);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, S.bsum, S.boff, S.bsum2, nbl);
    hipLaunchKernelGGL(k_scan_add, dim3(gv), dim3(256), 0, st, S.off, S.boff, nb);
    if (fixed_order) {
        const unsigned gf = (unsigned)((nb + GF_VOX - 1) / GF_VOX);
        hipLaunchKernelGGL(k_dens_place_rec, dim3(gb), dim3(256), 0, st, N, S.cell, S.off, S.fill, S.brec,
                           S.brec_bucketed);
        switch (Bp1 - 1) {
        case 1: hipLaunchKernelGGL(k_dens_gather_fixed<1>, dim3(gf), dim3(8 * GF_VOX), 0, st, G, S.off, S.cnt, S.brec_bucketed, density); break;
        case 2: hipLaunchKernelGGL(k_dens_gather_fixed<2>, dim3(gf), dim3(8 * GF_VOX), 0, st, G, S.off, S.cnt, S.brec_bucketed, density); break;
        case 3: hipLaunchKernelGGL(k_dens_gather_fixed<3>, dim3(gf), dim3(8 * GF_VOX), 0, st, G, S.off, S.cnt, S.brec_bucketed, density); break;
        default: hipLaunchKernelGGL(k_dens_gather_fixed<4>, dim3(gf), dim3(8 * GF_VOX), 0, st, G, S.off, S.cnt, S.brec_bucketed, density); break;
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_dens_place, dim3(gb), dim3(256), 0, st, N, S.cell, S.off, S.fill, S.list);
    hipLaunchKernelGGL(k_dens_sort, dim3(gv), dim3(256), 0, st, nb, S.off, S.cnt, S.list);
    if (nb >= (int64_t)1024 * GATHER_BIG_BLOCKS)
        hipLaunchKernelGGL((k_dens_gather<1024, true>), dim3((unsigned)((nb + 1023) / 1024)), dim3(1024), 0, st, G, Bp1,
                           S.off, S.cnt, S.list, S.brec, density, cleanup);
    else
        hipLaunchKernelGGL((k_dens_gather<256, false>), dim3(gv), dim3(256), 0, st, G, Bp1, S.off, S.cnt, S.list, S.brec,
                           density, cleanup);
    return hipGetLastError();
}

}  // namespace
}  // namespace chromo
