// chromo_energy.hip -- absolute energies and their adjoint (gfx950).
//
// The Monte Carlo dE of chromo (chromo/polymers.pyx:1024-1538,
// chromo/fields.pyx:1149-1842) are differences of
//
//   E(r, theta) = E_elastic(r)                       sum over bonds of E_pair
//               + E_field(rho(r), theta)             voxel sums below
//               + E_bind(states, theta)              single-site free energies
//
//   E_field = sum_v [ sum_j pref_j rho_vj^2 + sum_jk X_jk rho_vj rho_vk + g(phi_v) ]
//             + sum_j intra_j D_j
//   pref_j = eps_jj v_j Vbin / 2,  X_jk = eps_jk v_j Vbin (cross-talk),
//   intra_j = eps_jj (1 - v_j / Vbin),  phi_v = rho_v0 V,
//   g(phi) = chi (Vbin / V) phi^2  (phi <= vf_limit)  else 1e99 phi,
//   E_bind = sum_b sum_j [ F_j(marks, s) - s mu_j f ]  (mu_j > 0: - s mu_j (2 - f)).
//
// (The reference's own compute_E uses chi phi (1 - phi) and omits cross-talk,
// so its differences are not the move dE; chromo_energy_reference reproduces
// that function separately.)  Gradients: with respect to bead positions
// through the trilinear weights (dw/dx = -+1/dx) and the SSWLC bond energy, and
// with respect to theta = (chi, eps_ij for i <= j, mu_j).  Densities are
// recomputed (deterministic scatter) rather than stored; all sums use fixed-
// order block reductions.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <string>

#include "chromo_density.hpp"
#include "chromo_device.hpp"

using namespace chromo;

namespace {

constexpr int EB = 256;   // block size of the energy kernels
constexpr int MAXT = 64;  // partial-sum slots per block

int eset(int code, const std::string &m) { return chromo::set_last_error(code, m); }

#define HIPE(x)                                                                            \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return eset(-2, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct EModel {
    int64_t N;
    int B;
    int twist;
    GridGeom G;
    const double *t3;
    const int8_t *states, *marks;
    const double *bonds;
    double bead_vol, vf_limit, mu_adjust;
    double v_int[CHROMO_MAX_BINDERS];
    int cross_mask[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS];
    double F[CHROMO_MAX_BINDERS][CHROMO_MAX_MARK][CHROMO_MAX_SITES + 1];
};

// theta layout: [chi, eps_00, eps_01, .., eps_0B-1, eps_11, .., eps_B-1B-1, mu_0 .. mu_B-1]
__host__ __device__ inline int eps_index(int B, int i, int j)
{
    if (i > j) {
        int t = i;
        i = j;
        j = t;
    }
    // rows 0..i-1 hold B, B-1, .. entries
    return 1 + i * B - (i * (i - 1)) / 2 + (j - i);
}

__host__ __device__ inline int n_theta(int B) { return 1 + B * (B + 1) / 2 + B; }

struct Coef {
    double chi;
    double pref[CHROMO_MAX_BINDERS], intra[CHROMO_MAX_BINDERS];
    double X[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS];
    double mu[CHROMO_MAX_BINDERS];
};

__device__ inline Coef coef_of(const EModel &M, const double *th)
{
    Coef c;
    const int B = M.B;
    const double vb = M.G.vol_bin;
    c.chi = th[0];
    for (int j = 0; j < B; ++j) {
        double e = th[eps_index(B, j, j)];
        c.pref[j] = 0.5 * e * M.v_int[j] * vb;
        c.intra[j] = e * (1 - M.v_int[j] / vb);
        c.mu[j] = th[1 + B * (B + 1) / 2 + j];
        for (int k = 0; k < B; ++k)
            c.X[j * B + k] = M.cross_mask[j * B + k] ? th[eps_index(B, j, k)] * M.v_int[j] * vb : 0.0;
    }
    return c;
}

// coef_of for a binder count known at compile time: the same arithmetic, with
// constant indices (registers instead of scratch in the per-voxel and per-bead loops)
template <int B>
__device__ inline Coef coef_of(const EModel &M, const double *th)
{
    Coef c;
    const double vb = M.G.vol_bin;
    c.chi = th[0];
#pragma unroll
    for (int j = 0; j < B; ++j) {
        double e = th[eps_index(B, j, j)];
        c.pref[j] = 0.5 * e * M.v_int[j] * vb;
        c.intra[j] = e * (1 - M.v_int[j] / vb);
        c.mu[j] = th[1 + B * (B + 1) / 2 + j];
#pragma unroll
        for (int k = 0; k < B; ++k)
            c.X[j * B + k] = M.cross_mask[j * B + k] ? th[eps_index(B, j, k)] * M.v_int[j] * vb : 0.0;
    }
    return c;
}

// Block partials are stored term-major: part[t * gridDim.x + block], so that the
// final sums read each term's partials contiguously.

// Fixed-order block sum of NT partial values per thread: a butterfly within each
// wavefront, then the EB / 64 wavefront sums in wavefront order.
template <int NT>
__device__ void block_reduce_store(const double (&vals)[NT], double *part)
{
    __shared__ double sh[EB / 64][NT];
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double s = vals[t];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) sh[wv][t] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < NT) {
        double s = sh[0][threadIdx.x];
        for (int w = 1; w < EB / 64; ++w) s += sh[w][threadIdx.x];
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// fixed-order block sum of nt partial values per thread
__device__ void block_reduce_store(double *vals, int nt, double *part)
{
    __shared__ double sh[EB];
    for (int t = 0; t < nt; ++t) {
        sh[threadIdx.x] = vals[t];
        __syncthreads();
        for (int s = EB / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)t * gridDim.x + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

// ---- voxel terms -----------------------------------------------------------
// slots: [0..B) S_jj, [B..B+B(B-1)/2) S_jk (j<k), then sum g, sum dg/dchi
template <int B>
__global__ void __launch_bounds__(EB) k_e_voxel(EModel M, const double *theta, const double *rho, double *dEdrho,
                                                double *part)
{
    constexpr int Bp1 = B + 1, NT = B + B * (B - 1) / 2 + 2;
    const Coef c = coef_of<B>(M, theta);
    const int64_t nb = M.G.nx * M.G.ny * M.G.nz;
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 0;
    const double V = M.bead_vol, vb = M.G.vol_bin;
    for (int64_t v = (int64_t)blockIdx.x * EB + threadIdx.x; v < nb; v += (int64_t)gridDim.x * EB) {
        double r[Bp1];
#pragma unroll
        for (int j = 0; j < Bp1; ++j) r[j] = rho[v * Bp1 + j];
        int s = B;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            acc[j] += r[j + 1] * r[j + 1];
#pragma unroll
            for (int k = j + 1; k < B; ++k) acc[s++] += r[j + 1] * r[k + 1];
        }
        double phi = r[0] * V;
        bool huge = phi > M.vf_limit;
        acc[s] += huge ? E_HUGE_FIELD * phi : c.chi * (vb / V) * (phi * phi);
        acc[s + 1] += huge ? 0.0 : (vb / V) * (phi * phi);
        if (dEdrho) {
            double *d = dEdrho + v * Bp1;
            d[0] = V * (huge ? E_HUGE_FIELD : 2 * c.chi * (vb / V) * phi);
#pragma unroll
            for (int j = 0; j < B; ++j) {
                double g = 2 * c.pref[j] * r[j + 1];
#pragma unroll
                for (int k = 0; k < B; ++k) g += (c.X[j * B + k] + c.X[k * B + j]) * r[k + 1];
                d[j + 1] = g;
            }
        }
    }
    block_reduce_store<NT>(acc, part);
}

// ---- bead / bond terms -------------------------------------------------------
// Elastic energy of bond (r0, t30) -> (r1, t31), and (grad) its derivative with
// respect to r1 - r0 with the tangents fixed, from the same intermediates.
__device__ inline double E_pair_and_grad(const Bond &bd, const double *r0, const double *r1, const double *t30,
                                         const double *t31, bool grad, double g[3])
{
    double dr[3], drp[3], be[3];
    for (int i = 0; i < 3; ++i) dr[i] = r1[i] - r0[i];
    double par = dot3(t30, dr);
    for (int i = 0; i < 3; ++i) {
        drp[i] = dr[i] - t30[i] * par;
        be[i] = t31[i] - t30[i] - drp[i] * bd.eta;
    }
    if (grad) {
        // perp = P dr with P = I - t t^T; bend = t31 - t30 - eta perp
        double a[3];
        for (int i = 0; i < 3; ++i) a[i] = bd.eperp * drp[i] - bd.eta * bd.eb * be[i];
        double ta = dot3(t30, a);
        for (int i = 0; i < 3; ++i) g[i] = (a[i] - t30[i] * ta) + bd.ep * (par - bd.gam) * t30[i];
    }
    return E_pair(bd, be, par, drp);
}

// slots: [0] E_elastic, [1] E_bind, [2..2+B) D_j, [2+B..2+2B) sum s_j
// cell and brec: each bead's lower-corner voxel and weights from the density
// scatter of the same positions (grid_lower's values, not recomputed).
template <int B>
__global__ void __launch_bounds__(EB) k_e_bead(EModel M, const double *theta, const double *r,
                                               const double *dEdrho, const int32_t *cell, const BeadRec *brec,
                                               double *grad, double *part)
{
    constexpr int Bp1 = B + 1, NT = 2 + 2 * B;
    const int64_t N = M.N;
    const Coef c = coef_of<B>(M, theta);
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 0;
    const double f = M.mu_adjust;
    // +-1 / dx is exactly +-(1 / dx)
    const double inv_d[3] = {1.0 / M.G.dxyz[0], 1.0 / M.G.dxyz[1], 1.0 / M.G.dxyz[2]};
    // bond i = (i, i + 1): each is evaluated once, by thread i - b0 of the block
    // that holds bead i (the block's first bond (b0 - 1, b0) by its last thread);
    // the derivatives meet in LDS, where beads b and b + 1 read them
    __shared__ double gbond[EB + 1][3];
    const int t = threadIdx.x;
    for (int64_t b0 = (int64_t)blockIdx.x * EB; b0 < N; b0 += (int64_t)gridDim.x * EB) {
        const int64_t b = b0 + t;
        const double *rb = r + 3 * b;
        const double *tb = M.t3 + 3 * b;
        if (b + 1 < N) {
            Bond bd = load_bond(M.bonds, N, b);
            double gd[3];
            acc[0] += E_pair_and_grad(bd, rb, r + 3 * (b + 1), tb, M.t3 + 3 * (b + 1), grad != nullptr, gd);
            if (grad)
                for (int j = 0; j < 3; ++j) gbond[t + 1][j] = gd[j];
        }
        if (grad && t == EB - 1 && b0 > 0) {
            Bond bd = load_bond(M.bonds, N, b0 - 1);
            double gd[3];
            E_pair_and_grad(bd, r + 3 * (b0 - 1), r + 3 * b0, M.t3 + 3 * (b0 - 1), M.t3 + 3 * b0, true, gd);
            for (int j = 0; j < 3; ++j) gbond[0][j] = gd[j];
        }
        double g[3] = {0, 0, 0};
        if (b < N) {
            int st[B];
#pragma unroll
            for (int j = 0; j < B; ++j) st[j] = M.states[(int64_t)j * N + b];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int s = st[j];
                int m = M.marks[(int64_t)j * N + b];
                double mu = c.mu[j];
                double e = M.F[j][m][s];
                e += (mu > 0) ? -(double)s * (mu * (2 - f)) : -(double)s * mu * f;
                acc[1] += e;
                acc[2 + j] += (s == 2);
                acc[2 + B + j] += (double)s * ((mu > 0) ? (2 - f) : f);
            }
            if (grad && dEdrho) {
                // 32-bit index arithmetic: the grids hold fewer than 2^31 voxels
                const int32_t nx = (int32_t)M.G.nx, ny = (int32_t)M.G.ny, nz = (int32_t)M.G.nz;
                const uint32_t c = (uint32_t)cell[b], cq = c / (uint32_t)nx;
                const int32_t ix[3] = {(int32_t)(c - cq * (uint32_t)nx), (int32_t)(cq % (uint32_t)ny),
                                       (int32_t)(cq / (uint32_t)ny)};
                const double w[3] = {brec[b].w[0], brec[b].w[1], brec[b].w[2]};
                for (int l = 0; l < 8; ++l) {
                    int32_t x = ix[0] + (l & 1), y = ix[1] + ((l >> 1) & 1), z = ix[2] + ((l >> 2) & 1);
                    if (x >= nx) x -= nx;
                    if (y >= ny) y -= ny;
                    if (z >= nz) z -= nz;
                    const int64_t v = x + nx * (y + ny * z);
                    const double *d = dEdrho + v * Bp1;
                    double h = d[0];
#pragma unroll
                    for (int j = 0; j < B; ++j) h += (double)st[j] * d[j + 1];
                    h /= M.G.vol_bin;
                    double a[3];
                    for (int j = 0; j < 3; ++j) a[j] = ((l >> j) & 1) ? (1 - w[j]) : w[j];
                    for (int j = 0; j < 3; ++j) {
                        double dw = ((l >> j) & 1) ? inv_d[j] : -inv_d[j];
                        double prod = dw;
                        for (int q = 0; q < 3; ++q)
                            if (q != j) prod *= a[q];
                        g[j] += prod * h;
                    }
                }
            }
        }
        if (grad) {
            // every thread of the block reaches both barriers
            __syncthreads();
            if (b < N) {
                if (b + 1 < N)   // bond (b, b+1): r0 = r_b
                    for (int j = 0; j < 3; ++j) g[j] -= gbond[t + 1][j];
                if (b > 0)       // bond (b-1, b): r1 = r_b
                    for (int j = 0; j < 3; ++j) g[j] += gbond[t][j];
                for (int j = 0; j < 3; ++j) grad[3 * b + j] = g[j];
            }
            __syncthreads();
        }
    }
    block_reduce_store<NT>(acc, part);
}

// The voxel and bead kernels of one energy evaluation, for B binders.
template <int B>
void launch_energy_terms(const EModel &M, const double *theta, const double *rho, double *dEdrho, double *vpart,
                         int gv, const double *r, const DensityScratch &S, double *grad, double *bpart, int gb,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_e_voxel<B>, dim3(gv), dim3(EB), 0, st, M, theta, rho, dEdrho, vpart);
    hipLaunchKernelGGL(k_e_bead<B>, dim3(gb), dim3(EB), 0, st, M, theta, r, dEdrho, S.cell, S.brec, grad, bpart);
}

// Sums of the block partials of ntv voxel terms and ntb bead terms, by one block of
// FINAL_THREADS: wavefront w takes terms w, w + 16, ...; lane i adds partials i,
// i + 64, ... in increasing order (loaded 16 at a time, so that one round trip
// covers 16 of them), then a butterfly over the 64 lanes.  The order depends on
// the sizes only, so the sums are the same on every run.
constexpr int FINAL_THREADS = 1024;

__device__ void final_sums(const double *vpart, int nbv, int ntv, const double *bpart, int nbb, int ntb,
                           double *sv, double *sb)
{
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int q = wv; q < ntv + ntb; q += FINAL_THREADS / 64) {
        const bool vox = q < ntv;
        const double *p = vox ? vpart : bpart;
        const int nbk = vox ? nbv : nbb, t = vox ? q : q - ntv;
        double s = 0;
        for (int i0 = lane; i0 < nbk; i0 += 64 * 16) {
            double x[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int i = i0 + 64 * j;
                x[j] = i < nbk ? p[(int64_t)t * nbk + i] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (i0 + 64 * j < nbk) s += x[j];
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) (vox ? sv : sb)[t] = s;
    }
    __syncthreads();
}

// Final reduction over block partials (fixed order) and assembly, for B binders
// (compile-time, so the coefficients and sums stay in registers).
template <int B>
__global__ void __launch_bounds__(FINAL_THREADS)
k_e_final(EModel M, const double *theta, const double *vpart, int nbv, int ntv, const double *bpart, int nbb,
          int ntb, double *E_out, double *gtheta)
{
    __shared__ double sv[MAXT], sb[MAXT];
    final_sums(vpart, nbv, ntv, bpart, nbb, ntb, sv, sb);
    if (threadIdx.x != 0) return;
    const Coef c = coef_of<B>(M, theta);
    const double vb = M.G.vol_bin;
    double S[CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS];
    int s = B;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        S[j * B + j] = sv[j];
#pragma unroll
        for (int k = j + 1; k < B; ++k) {
            S[j * B + k] = sv[s];
            S[k * B + j] = sv[s];
            ++s;
        }
    }
    const double Eg = sv[s], dchi = sv[s + 1];
    double Ef = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        Ef += c.pref[j] * S[j * B + j];
        Ef += c.intra[j] * sb[2 + j];
#pragma unroll
        for (int k = 0; k < B; ++k) Ef += c.X[j * B + k] * S[j * B + k];
    }
    Ef += Eg;
    const double Eel = sb[0], Ebd = sb[1];
    E_out[0] = Eel + Ef + Ebd;
    E_out[1] = Eel;
    E_out[2] = Ef;
    E_out[3] = Ebd;
    if (gtheta) {
        gtheta[0] = dchi;
#pragma unroll
        for (int j = 0; j < B; ++j)
#pragma unroll
            for (int k = j; k < B; ++k) {
                double g;
                if (j == k) {
                    g = 0.5 * M.v_int[j] * vb * S[j * B + j] + (1 - M.v_int[j] / vb) * sb[2 + j];
                    if (M.cross_mask[j * B + j]) g += M.v_int[j] * vb * S[j * B + j];
                } else {
                    g = ((M.cross_mask[j * B + k] ? M.v_int[j] : 0.0) +
                         (M.cross_mask[k * B + j] ? M.v_int[k] : 0.0)) * vb * S[j * B + k];
                }
                gtheta[eps_index(B, j, k)] = g;
            }
#pragma unroll
        for (int j = 0; j < B; ++j) gtheta[1 + B * (B + 1) / 2 + j] = -sb[2 + B + j];
    }
}

// The final reduction and assembly of one energy evaluation, for B binders.
template <int B>
void launch_energy_final(const EModel &M, const double *theta, const double *vpart, int gv, int ntv,
                         const double *bpart, int gb, int ntb, double *E, double *gtheta, hipStream_t st)
{
    hipLaunchKernelGGL(k_e_final<B>, dim3(1), dim3(FINAL_THREADS), 0, st, M, theta, vpart, gv, ntv, bpart, gb, ntb,
                       E, gtheta);
}

// ---- the reference's compute_E (fields.pyx:1939-2286, polymers.pyx:1348-1381) ----

// Python round(x, 2) > lim, evaluated on the exact binary value of x.
__device__ inline double py_round2(double x)
{
    double p = x * 100.0;
    double e = fma(x, 100.0, -p);   // exact residual: x * 100 = p + e
    double k = rint(p);
    double d = p - k;
    if (d == 0.5 && e > 0) k += 1.0;
    else if (d == -0.5 && e < 0) k -= 1.0;
    return k / 100.0;
}

// slots: [0..B) sum rho_j^2, [B] nonspecific E
__global__ void k_ref_voxel(EModel M, const double *pref, double chi, const double *rho, double *part,
                            int nt)
{
    const int B = M.B, Bp1 = B + 1;
    const int64_t nb = M.G.nx * M.G.ny * M.G.nz;
    double acc[MAXT];
    for (int t = 0; t < nt; ++t) acc[t] = 0;
    const double V = M.bead_vol, vb = M.G.vol_bin;
    for (int64_t v = (int64_t)blockIdx.x * EB + threadIdx.x; v < nb; v += (int64_t)gridDim.x * EB) {
        const double *r = rho + v * Bp1;
        for (int j = 0; j < B; ++j) acc[j] += r[j + 1] * r[j + 1];
        double vf = r[0] * V;
        if (py_round2(vf) > M.vf_limit) acc[B] += E_HUGE_FIELD * vf;
        else acc[B] += chi * (vb / V) * vf * (1 - vf);
    }
    block_reduce_store(acc, nt, part);
}

// slots: [0] SSWLC.compute_E bonds (SSTWLC.compute_E with the twist term), [1..1+B)
// doubly bound counts
__global__ void k_ref_bead(EModel M, const double *pos, int64_t sc, int64_t sb, double *part, int nt)
{
    const int B = M.B;
    const int64_t N = M.N;
    double acc[MAXT];
    for (int t = 0; t < nt; ++t) acc[t] = 0;
    for (int64_t b = (int64_t)blockIdx.x * EB + threadIdx.x; b < N; b += (int64_t)gridDim.x * EB) {
        if (b + 1 < N) {
            double r0[3], r1[3], t30[3], t31[3];
            for (int j = 0; j < 3; ++j) {
                r0[j] = pos[j * sc + b * sb];
                r1[j] = pos[j * sc + (b + 1) * sb];
                t30[j] = pos[(3 + j) * sc + b * sb];
                t31[j] = pos[(3 + j) * sc + (b + 1) * sb];
            }
            Bond bd = load_bond(M.bonds, N, b);
            double dr[3], drp[3], be[3];
            for (int j = 0; j < 3; ++j) dr[j] = r1[j] - r0[j];
            double par = dot3(t30, dr);
            for (int j = 0; j < 3; ++j) drp[j] = dr[j] - t30[j] * par;
            for (int j = 0; j < 3; ++j) {   // bend = t3_1; bend += -t3_0 - eta dr_perp
                be[j] = t31[j];
                be[j] += -t30[j] - bd.eta * drp[j];
            }
            double tw = 0;
            if (M.twist) {
                double t20[3], t21[3], et, tw0;
                for (int j = 0; j < 3; ++j) {
                    t20[j] = pos[(6 + j) * sc + b * sb];
                    t21[j] = pos[(6 + j) * sc + (b + 1) * sb];
                }
                load_twist(M.bonds, N, b, et, tw0);
                tw = twist_energy(et, tw0, twist_omega(t20, t30, t21, t31));
            }
            acc[0] += E_pair(bd, be, par, drp) + tw;
        }
        for (int j = 0; j < B; ++j) acc[1 + j] += (M.states[(int64_t)j * N + b] == 2);
    }
    block_reduce_store(acc, nt, part);
}

__global__ void __launch_bounds__(FINAL_THREADS)
k_ref_final(EModel M, const double *pref, const double *intra, const double *vpart, int nbv,
            const double *bpart, int nbb, double *out)
{
    __shared__ double sv[MAXT], sb[MAXT];
    const int B = M.B;
    final_sums(vpart, nbv, B + 1, bpart, nbb, B + 1, sv, sb);
    if (threadIdx.x != 0) return;
    double E = 0;
    for (int i = 0; i < B; ++i) {
        E += pref[i] * sv[i];
        E += intra[i] * sb[1 + i];
    }
    E += sv[B];
    out[0] = sb[0];
    out[1] = E;
}

// Blocks of a voxel or bead kernel: at most 2048 (eight waves per SIMD of the
// device), so that the final sums read few partials per term.
int grid_blocks(int64_t n)
{
    int64_t g = (n + EB - 1) / EB;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return (int)g;
}

int make_model(const chromo_energy_model *m, EModel &M)
{
    if (m == nullptr) return eset(-1, "model is NULL");
    if (m->B < 1 || m->B > CHROMO_MAX_BINDERS) return eset(-1, "B outside 1..4");
    if (m->N < 2) return eset(-1, "N < 2");
    // the scatter stores bead indices as int32 and uses 0x7fffffff as its empty marker
    if (m->N >= 0x7fffffffLL) return eset(-1, "N must be below 2^31 - 1");
    if (n_theta(m->B) > MAXT || 2 + 2 * m->B > MAXT) return eset(-1, "too many terms");
    M.N = m->N;
    M.B = m->B;
    M.twist = m->twist != 0;
    M.G.nx = m->nx;
    M.G.ny = m->ny;
    M.G.nz = m->nz;
    const double W[3] = {m->x_width, m->y_width, m->z_width};
    const int64_t n[3] = {m->nx, m->ny, m->nz};
    for (int j = 0; j < 3; ++j) {
        M.G.width[j] = W[j];
        M.G.dxyz[j] = W[j] / (double)n[j];
        M.G.half_width[j] = 0.5 * W[j];
        M.G.half_step[j] = 0.5 * M.G.dxyz[j];
    }
    M.G.vol_bin = m->x_width * m->y_width * m->z_width / (double)(m->nx * m->ny * m->nz);
    M.t3 = m->t3;
    M.states = m->states;
    M.marks = m->marks;
    M.bonds = m->bonds;
    M.bead_vol = m->bead_vol;
    M.vf_limit = m->vf_limit;
    M.mu_adjust = m->mu_adjust;
    for (int j = 0; j < CHROMO_MAX_BINDERS; ++j) M.v_int[j] = m->v_int[j];
    for (int j = 0; j < CHROMO_MAX_BINDERS * CHROMO_MAX_BINDERS; ++j) M.cross_mask[j] = 0;
    for (int j = 0; j < m->B; ++j)
        for (int k = 0; k < m->B; ++k) M.cross_mask[j * m->B + k] = m->cross_mask[j * CHROMO_MAX_BINDERS + k];
    memcpy(M.F, m->F, sizeof(M.F));
    int64_t nb = m->nx * m->ny * m->nz;
    if (nb < 1 || nb > 2048LL * 2048LL) return eset(-1, "grid size outside 1..2^22 voxels");
    return 0;
}

// Device buffers of one energy evaluation, carved from one allocation (or from
// the file-scope scratch array of chromo_energy_adjoint_noalloc).
struct EnergyBufs {
    DensityScratch S;
    double *rho = nullptr, *dEdrho = nullptr, *vpart = nullptr, *bpart = nullptr;
    int gv = 0, gb = 0;
};

// Carves the buffers of an evaluation of M from base (nullptr: sizes only) and
// returns the bytes used.  rho_given: the density goes to the caller's array.
// fixed_order: the adjoint's density (k_dens_gather_fixed); otherwise the
// reference's bead-order fold.
size_t energy_carve(const EModel &M, bool grad, bool rho_given, bool fixed_order, char *base, EnergyBufs &Bf)
{
    const int64_t nb = M.G.nx * M.G.ny * M.G.nz;
    const int Bp1 = M.B + 1;
    size_t o = density_scratch_carve(Bf.S, base, M.N, nb, fixed_order);
    auto take = [&](size_t bytes) {
        double *p = base ? (double *)(base + o) : nullptr;
        o = align_scratch(o + bytes);
        return p;
    };
    Bf.gv = grid_blocks(nb);
    Bf.gb = grid_blocks(M.N);
    Bf.rho = rho_given ? nullptr : take(sizeof(double) * nb * Bp1);
    Bf.dEdrho = grad ? take(sizeof(double) * nb * Bp1) : nullptr;
    Bf.vpart = take(sizeof(double) * MAXT * Bf.gv);
    Bf.bpart = take(sizeof(double) * MAXT * Bf.gb);
    return o;
}

// One allocation (stream-ordered) freed on every return path.
struct StreamBuffer {
    char *p = nullptr;
    hipStream_t st = nullptr;
    ~StreamBuffer()
    {
        if (p) (void)hipFreeAsync(p, st);
    }
};

// E, dE/dr and dE/dtheta of M at (r, theta) with the buffers Bf.
int energy_adjoint_run(const EModel &M, const double *d_r, const double *d_theta, double *d_E, double *d_grad_r,
                       double *d_grad_theta, EnergyBufs &Bf, hipStream_t st)
{
    const int B = M.B, Bp1 = B + 1;
    PosView P{d_r, 1, 3};
    // a deterministic density, not the reference's bead-order fold (nothing compares
    // the adjoint's densities with the reference's bit for bit)
    hipError_t e = density_scatter(M.G, M.N, Bp1, P, M.states, Bf.rho, 0, Bf.S, st, true);
    if (e != hipSuccess) return eset(-2, std::string("density scatter: ") + hipGetErrorString(e));
    const int ntv = B + B * (B - 1) / 2 + 2, ntb = 2 + 2 * B;
    const int gv = Bf.gv, gb = Bf.gb;
    switch (B) {
    case 1: launch_energy_terms<1>(M, d_theta, Bf.rho, Bf.dEdrho, Bf.vpart, gv, d_r, Bf.S, d_grad_r, Bf.bpart, gb, st); break;
    case 2: launch_energy_terms<2>(M, d_theta, Bf.rho, Bf.dEdrho, Bf.vpart, gv, d_r, Bf.S, d_grad_r, Bf.bpart, gb, st); break;
    case 3: launch_energy_terms<3>(M, d_theta, Bf.rho, Bf.dEdrho, Bf.vpart, gv, d_r, Bf.S, d_grad_r, Bf.bpart, gb, st); break;
    default: launch_energy_terms<4>(M, d_theta, Bf.rho, Bf.dEdrho, Bf.vpart, gv, d_r, Bf.S, d_grad_r, Bf.bpart, gb, st); break;
    }
    switch (B) {
    case 1: launch_energy_final<1>(M, d_theta, Bf.vpart, gv, ntv, Bf.bpart, gb, ntb, d_E, d_grad_theta, st); break;
    case 2: launch_energy_final<2>(M, d_theta, Bf.vpart, gv, ntv, Bf.bpart, gb, ntb, d_E, d_grad_theta, st); break;
    case 3: launch_energy_final<3>(M, d_theta, Bf.vpart, gv, ntv, Bf.bpart, gb, ntb, d_E, d_grad_theta, st); break;
    default: launch_energy_final<4>(M, d_theta, Bf.vpart, gv, ntv, Bf.bpart, gb, ntb, d_E, d_grad_theta, st); break;
    }
    HIPE(hipGetLastError());
    return 0;
}

// ---- the allocation-free evaluation -----------------------------------------

// Device memory of chromo_energy_adjoint_noalloc: one array per GPU, held until
// the process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
constexpr size_t NOALLOC_BYTES = size_t(32) << 20;
__device__ __align__(256) char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The scratch array of the current device, looked up once per device.
int noalloc_scratch(char **out)
{
    int dev = 0;
    HIPE(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return eset(-1, "device index outside 0..63");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HIPE(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return eset(-2, "hipGetSymbolAddress gave no address for the scratch array");
        g_noalloc_ptr[dev] = (char *)p;
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int chromo_energy_adjoint(const chromo_energy_model *model, const double *d_r, const double *d_theta,
                          double *d_E, double *d_grad_r, double *d_grad_theta, void *stream)
{
    EModel M;
    int rc = make_model(model, M);
    if (rc) return rc;
    if (M.twist) return eset(-1, "the energy adjoint has no twist term (SSTWLC)");
    if (!d_r || !d_theta || !d_E) return eset(-1, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    EnergyBufs Bf;
    const size_t bytes = energy_carve(M, d_grad_r != nullptr, false, true, nullptr, Bf);
    StreamBuffer buf;
    buf.st = st;
    HIPE(hipMallocAsync((void **)&buf.p, bytes, st));
    energy_carve(M, d_grad_r != nullptr, false, true, buf.p, Bf);
    return energy_adjoint_run(M, d_r, d_theta, d_E, d_grad_r, d_grad_theta, Bf, st);
}

int chromo_energy_adjoint_noalloc(const chromo_energy_model *model, const double *d_r, const double *d_theta,
                                  double *d_E, double *d_grad_r, double *d_grad_theta, void *stream)
{
    EModel M;
    int rc = make_model(model, M);
    if (rc) return rc;
    if (M.twist) return eset(-1, "the energy adjoint has no twist term (SSTWLC)");
    if (!d_r || !d_theta || !d_E) return eset(-1, "NULL argument");
    EnergyBufs Bf;
    const size_t bytes = energy_carve(M, d_grad_r != nullptr, false, true, nullptr, Bf);
    if (bytes > NOALLOC_BYTES)
        return eset(-1, "the evaluation needs " + std::to_string(bytes) + " bytes of scratch; the allocation-free "
                        "entry point holds " + std::to_string(NOALLOC_BYTES) + " (use chromo_energy_adjoint)");
    char *base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    energy_carve(M, d_grad_r != nullptr, false, true, base, Bf);
    return energy_adjoint_run(M, d_r, d_theta, d_E, d_grad_r, d_grad_theta, Bf, (hipStream_t)stream);
}

int chromo_energy_reference(const chromo_energy_model *model, const double *d_pos, int32_t soa,
                            const double *d_pref, const double *d_intra, double chi,
                            double *d_density_out, double *d_out, void *stream)
{
    EModel M;
    int rc = make_model(model, M);
    if (rc) return rc;
    if (!d_pos || !d_pref || !d_intra || !d_out) return eset(-1, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const int B = M.B, Bp1 = B + 1;
    EnergyBufs Bf;
    const size_t bytes = energy_carve(M, false, d_density_out != nullptr, false, nullptr, Bf);
    StreamBuffer buf;
    buf.st = st;
    HIPE(hipMallocAsync((void **)&buf.p, bytes, st));
    energy_carve(M, false, d_density_out != nullptr, false, buf.p, Bf);
    double *rho = d_density_out ? d_density_out : Bf.rho;
    const int gv = Bf.gv, gb = Bf.gb;
    // UniformDensityField.update_all_densities: bead-order scatter without the 1e-18 clean-up
    const int64_t sc = soa ? M.N : 1, sbd = soa ? 1 : 9;
    PosView P{d_pos, sc, sbd};
    hipError_t e = density_scatter(M.G, M.N, Bp1, P, M.states, rho, 0, Bf.S, st);
    if (e != hipSuccess) return eset(-2, std::string("density scatter: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(k_ref_voxel, dim3(gv), dim3(EB), 0, st, M, d_pref, chi, rho, Bf.vpart, B + 1);
    hipLaunchKernelGGL(k_ref_bead, dim3(gb), dim3(EB), 0, st, M, d_pos, sc, sbd, Bf.bpart, B + 1);
    hipLaunchKernelGGL(k_ref_final, dim3(1), dim3(FINAL_THREADS), 0, st, M, d_pref, d_intra, Bf.vpart, gv, Bf.bpart,
                       gb, d_out);
    HIPE(hipGetLastError());
    return 0;
}

}  // extern "C"
