// chromo_spec.hip -- the speculative persistent kernel k_mc_spec<W, NB> for one
// W = CHROMO_SPEC_W (wavefronts per replica); built once per W so that the
// instantiations compile in parallel.
#include <hip/hip_runtime.h>

#include "chromo_spec.hpp"

#ifndef CHROMO_SPEC_W
#error "compile with -DCHROMO_SPEC_W=2, 4 or 8"
#endif

#define CHROMO_CAT2(a, b) a##b
#define CHROMO_CAT(a, b) CHROMO_CAT2(a, b)

namespace {

// Persistent mc_sim with W wavefronts per replica (one workgroup per replica),
// for the replicas with NB binder types; the others are left to the launch of
// their own NB.  Two workgroups share a compute unit (LDS), so a SIMD runs W / 2
// wavefronts: the register budget is sized for that.
template <int W, int NB>
__global__ void __launch_bounds__(WAVE * W) __attribute__((amdgpu_waves_per_eu(W / 2 > 1 ? W / 2 : 1)))
k_mc_spec(const chromo_replica *__restrict__ reps, int n_reps, int64_t steps)
{
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    __shared__ chromo_replica Rs;
    load_record(&Rs, &reps[r]);
    __syncthreads();
    const chromo_replica *R = &Rs;
    if (R->B != NB) return;
    if (R->counters[CHROMO_CNT_ERROR] != 0) return;
    if (R->twist) {
        // the speculative evaluator has no twist terms: a twisting chain runs
        // with one wavefront (k_mc_run)
        counters_error(R, ERR_TWIST_WAVES);
        return;
    }
    spec_run<NB, W>(R, steps);
}

template <int NB>
hipError_t launch_nb(const chromo_replica *d_reps, int n_reps, int64_t steps, hipStream_t st, int lim)
{
    constexpr int W = CHROMO_SPEC_W;
    const size_t bytes = spec_layout(NB + 1, W).total;
    if (bytes > (size_t)lim) return hipErrorInvalidValue;
    // the attribute is set once per process: setting it costs far more than a launch
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_mc_spec<W, NB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_mc_spec<W, NB>), dim3(n_reps), dim3(WAVE * W), bytes, st, d_reps, n_reps, steps);
    return hipGetLastError();
}

}  // namespace

namespace chromo {

// Launches k_mc_spec<CHROMO_SPEC_W, B> for every binder count B whose bit is set
// in binder_mask (bit B for B = 1..4).  Returns hipErrorInvalidValue when the
// LDS of a launch exceeds the device limit.
hipError_t CHROMO_CAT(spec_launch_w, CHROMO_SPEC_W)(const chromo_replica *d_reps, int n_reps, int64_t steps,
                                                    int binder_mask, hipStream_t st)
{
    static int lim = 0;
    hipError_t e = hipSuccess;
    if (lim == 0) {
        int dev = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
            return e;
    }
    if ((binder_mask & 2) && (e = launch_nb<1>(d_reps, n_reps, steps, st, lim)) != hipSuccess) return e;
    if ((binder_mask & 4) && (e = launch_nb<2>(d_reps, n_reps, steps, st, lim)) != hipSuccess) return e;
    if ((binder_mask & 8) && (e = launch_nb<3>(d_reps, n_reps, steps, st, lim)) != hipSuccess) return e;
    if ((binder_mask & 16) && (e = launch_nb<4>(d_reps, n_reps, steps, st, lim)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace chromo
