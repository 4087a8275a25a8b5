// tm_kernels_f64.hip -- FP64 instantiations of the step kernels, plus the
// precision-independent rho(t) combine.  Compiled with -ffp-contract=off so
// every FP64 operation rounds exactly as the CPU restatement's does.
#define TM_DEFINE_COMBINE
#include "tm_launch.h"

namespace tmk {

template <typename T, int TX, int TY, int NT, int MODE = 0, int MINW = 1, int Q = 0>
static void launch_dp(const StepArgs& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((dp_step_kernel<T, TX, TY, NT, MODE, MINW, Q>), dim3(grid), dim3(NT), 0, s, a);
}

template <typename T, int TX, int TY, int NT, int K, int MINW = 1>
static void launch_dpk(const StepArgs& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((dp_stepk_kernel<T, TX, TY, NT, K, MINW>), dim3(grid), dim3(NT), 0, s, a);
}

template <typename T, int TX, int TY, int NT, bool SPLIT = false>
static void launch_rk(const StepArgs& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((rk_step_kernel<T, TX, TY, NT, SPLIT>), dim3(grid), dim3(NT), 0, s, a);
}

static const Variant kDP[] = {
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256", &launch_dp<double, 64, 32, 256>},
    {128, 32, 256, 0, 4, "dp_f64_128x32_t256", &launch_dp<double, 128, 32, 256>},
    {64, 64, 256, 0, 4, "dp_f64_64x64_t256", &launch_dp<double, 64, 64, 256>},
    {256, 32, 256, 0, 4, "dp_f64_256x32_t256", &launch_dp<double, 256, 32, 256>},
    // _xK: K steps per launch (temporal blocking); the single-step kernel of the same
    // tile covers a block's last steps when fewer than K remain
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x2", &launch_dpk<double, 64, 32, 256, 2>, 0, 0, nullptr, 0, 2,
     &launch_dp<double, 64, 32, 256>},
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x4", &launch_dpk<double, 64, 32, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<double, 64, 32, 256>},
    {64, 16, 256, 0, 4, "dp_f64_64x16_t256_x4", &launch_dpk<double, 64, 16, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<double, 64, 16, 256>},
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x8", &launch_dpk<double, 64, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<double, 64, 32, 256>},
    {64, 8, 128, 0, 2, "dp_f64_64x8_t128_x16", &launch_dpk<double, 64, 8, 128, 16>, 0, 0, nullptr, 0, 16,
     &launch_dp<double, 64, 8, 128>},
    // _xKa: adaptive -- single steps while more than the given per-mille of the cells are
    // occupied, K-step launches below
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x4a", &launch_dpk<double, 64, 32, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<double, 64, 32, 256>, 20},
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x8a", &launch_dpk<double, 64, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<double, 64, 32, 256>, 20},
    // _q: the single steps defer cells whose Gamma trial is rejected to a second
    // pass over a queue in LDS (dp_tile_deferred)
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x8aq", &launch_dpk<double, 64, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<double, 64, 32, 256, 0, 1, 1>, 20},
    // _w6: its single step bounded to 80 VGPRs (six waves per SIMD)
    {64, 32, 256, 0, 4, "dp_f64_64x32_t256_x8aq_w6", &launch_dpk<double, 64, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<double, 64, 32, 256, 0, 6, 1>, 20},
};

static const Variant kRK[] = {
    {16, 16, 256, 0, 1, "rk_f64_16x16_t256", &launch_rk<double, 16, 16, 256>},
    {32, 16, 256, 0, 1, "rk_f64_32x16_t256", &launch_rk<double, 32, 16, 256>},
    {32, 32, 256, 0, 1, "rk_f64_32x32_t256", &launch_rk<double, 32, 32, 256>},
    // split step: Dornic kernel (dp_step_kernel MODE 1) on 64 x 32 tiles, then the RK4 kernel
    {16, 16, 256, 0, 1, "rk_f64_split_16x16_t256", &launch_rk<double, 16, 16, 256, true>, 64, 32,
     &launch_dp<double, 64, 32, 256, 1>},
    {32, 16, 256, 0, 1, "rk_f64_split_32x16_t256", &launch_rk<double, 32, 16, 256, true>, 64, 32,
     &launch_dp<double, 64, 32, 256, 1>},
    {32, 32, 256, 0, 1, "rk_f64_split_32x32_t256", &launch_rk<double, 32, 32, 256, true>, 64, 32,
     &launch_dp<double, 64, 32, 256, 1>},
};

const Variant* dp_variants_f64(int* n) { *n = (int)(sizeof(kDP) / sizeof(kDP[0])); return kDP; }
const Variant* rk_variants_f64(int* n) { *n = (int)(sizeof(kRK) / sizeof(kRK[0])); return kRK; }

void launch_measure_f64(int nf, const DevInst* inst, const DevTile* tiles, int ntiles, int tx, int ty,
                        const void* fields, double2* partials, hipStream_t s)
{
    if (nf == 1)
        hipLaunchKernelGGL((measure_kernel<double, 1>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const double*)fields, partials);
    else if (nf == 3)
        hipLaunchKernelGGL((measure_kernel<double, 3>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const double*)fields, partials);
    else if (nf == 5)
        hipLaunchKernelGGL((measure_kernel<double, 5>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const double*)fields, partials);
    else
        hipLaunchKernelGGL((measure_kernel<double, 7>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const double*)fields, partials);
}

void launch_combine(const CombineArgs& a, int n_inst, int nsteps, hipStream_t s)
{
    hipLaunchKernelGGL(combine_rows_kernel, dim3(n_inst, nsteps), dim3(256), 0, s, a);
}

void launch_prep(int model, const DevInst* inst, const void* params, const DevState* state, int n_inst,
                 int nsteps, void* table, hipStream_t s)
{
    const int64_t n = (int64_t)n_inst * nsteps;
    const int grid = (int)((n + 255) / 256);
    if (model == TM_MODEL_DP)
        hipLaunchKernelGGL(prep_consts_kernel<TM_MODEL_DP>, dim3(grid), dim3(256), 0, s, inst, params, state,
                           n_inst, nsteps, table);
    else if (model == TM_MODEL_DP2 || model == TM_MODEL_DP3)
        hipLaunchKernelGGL(prep_consts_kernel<TM_MODEL_DP2>, dim3(grid), dim3(256), 0, s, inst, params, state,
                           n_inst, nsteps, table);
    else if (model == TM_MODEL_FKE2 || model == TM_MODEL_FKE3)
        hipLaunchKernelGGL(prep_consts_kernel<TM_MODEL_FKE2>, dim3(grid), dim3(256), 0, s, inst, params, state,
                           n_inst, nsteps, table);
    else
        hipLaunchKernelGGL(prep_consts_kernel<TM_MODEL_RIETKERK>, dim3(grid), dim3(256), 0, s, inst, params,
                           state, n_inst, nsteps, table);
}

void launch_bump(DevState* state, int n, hipStream_t s)
{
    hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(64), 0, s, state, n);
}

size_t const_size(int model)
{
    if (model == TM_MODEL_DP) return sizeof(DPConst);
    if (model == TM_MODEL_DP2 || model == TM_MODEL_DP3) return sizeof(DPCConst);
    return sizeof(RKConstD);
}

}  // namespace tmk
