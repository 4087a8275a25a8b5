// tm_kernels_f32.hip -- FP32-storage instantiations of the step kernels.
// Fields, stencil and RK4 stages are FP32 (contraction allowed, hardware
// reciprocal); the Poisson mean and draw, the ramp and the rho(t) sums stay FP64
// (their helpers pin contraction off locally).
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

// _w6: at most 80 VGPRs (six waves per SIMD); _rm: tiles ordered along rows.
static const Variant kDP[] = {
    {64, 32, 256, 0, 4, "dp_f32_64x32_t256", &launch_dp<float, 64, 32, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256", &launch_dp<float, 128, 32, 256>},
    {64, 64, 256, 0, 4, "dp_f32_64x64_t256", &launch_dp<float, 64, 64, 256>},
    {256, 32, 256, 0, 4, "dp_f32_256x32_t256", &launch_dp<float, 256, 32, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_w6", &launch_dp<float, 128, 32, 256, 0, 6>},
    {64, 64, 256, 0, 4, "dp_f32_64x64_t256_w6", &launch_dp<float, 64, 64, 256, 0, 6>},
    {128, 16, 256, 0, 4, "dp_f32_128x16_t256_rm", &launch_dp<float, 128, 16, 256>, 0, 0, nullptr, 1},
    {256, 32, 256, 0, 4, "dp_f32_256x32_t256_rm", &launch_dp<float, 256, 32, 256>, 0, 0, nullptr, 1},
    // _xK: K steps per launch (temporal blocking); the single-step kernel of the same
    // tile covers a block's last steps when fewer than K remain
    {128, 16, 256, 0, 4, "dp_f32_128x16_t256_x2", &launch_dpk<float, 128, 16, 256, 2>, 0, 0, nullptr, 0, 2,
     &launch_dp<float, 128, 16, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x2", &launch_dpk<float, 128, 32, 256, 2>, 0, 0, nullptr, 0, 2,
     &launch_dp<float, 128, 32, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x3", &launch_dpk<float, 128, 32, 256, 3>, 0, 0, nullptr, 0, 3,
     &launch_dp<float, 128, 32, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x4", &launch_dpk<float, 128, 32, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<float, 128, 32, 256>},
    {128, 16, 256, 0, 4, "dp_f32_128x16_t256_x4", &launch_dpk<float, 128, 16, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<float, 128, 16, 256>},
    {128, 24, 256, 0, 4, "dp_f32_128x24_t256_x4", &launch_dpk<float, 128, 24, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<float, 128, 24, 256>},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x8", &launch_dpk<float, 128, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<float, 128, 32, 256>},
    // _xKa: adaptive -- single steps (<= 80 VGPRs) while more than the given per-mille of
    // the cells are occupied, K-step launches below
    {128, 24, 256, 0, 4, "dp_f32_128x24_t256_x4a", &launch_dpk<float, 128, 24, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<float, 128, 24, 256, 0, 6>, 20},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x4a", &launch_dpk<float, 128, 32, 256, 4>, 0, 0, nullptr, 0, 4,
     &launch_dp<float, 128, 32, 256, 0, 6>, 20},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x8a", &launch_dpk<float, 128, 32, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<float, 128, 32, 256, 0, 6>, 20},
    {128, 24, 256, 0, 4, "dp_f32_128x24_t256_x8a", &launch_dpk<float, 128, 24, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<float, 128, 24, 256, 0, 6>, 20},
    {128, 16, 256, 0, 4, "dp_f32_128x16_t256_x8a", &launch_dpk<float, 128, 16, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<float, 128, 16, 256, 0, 6>, 20},
    {128, 32, 256, 0, 4, "dp_f32_128x32_t256_x6a", &launch_dpk<float, 128, 32, 256, 6>, 0, 0, nullptr, 0, 6,
     &launch_dp<float, 128, 32, 256, 0, 6>, 20},
    {128, 16, 256, 0, 4, "dp_f32_128x16_t256_x16a", &launch_dpk<float, 128, 16, 256, 16>, 0, 0, nullptr, 0, 16,
     &launch_dp<float, 128, 16, 256, 0, 6>, 20},
    // _q: the single steps defer cells whose Gamma trial is rejected to a second
    // pass over a queue in LDS (dp_tile_deferred)
    {128, 24, 256, 0, 4, "dp_f32_128x24_t256_w6q", &launch_dp<float, 128, 24, 256, 0, 6, 1>},
    {128, 24, 256, 0, 4, "dp_f32_128x24_t256_x8aq", &launch_dpk<float, 128, 24, 256, 8>, 0, 0, nullptr, 0, 8,
     &launch_dp<float, 128, 24, 256, 0, 6, 1>, 20},
};

static const Variant kRK[] = {
    {16, 16, 256, 0, 1, "rk_f32_16x16_t256", &launch_rk<float, 16, 16, 256>},
    {32, 16, 256, 0, 1, "rk_f32_32x16_t256", &launch_rk<float, 32, 16, 256>},
    {32, 32, 256, 0, 1, "rk_f32_32x32_t256", &launch_rk<float, 32, 32, 256>},
    // split step: Dornic kernel (dp_step_kernel MODE 1) on 128 x 32 tiles, then the RK4 kernel
    {32, 16, 256, 0, 1, "rk_f32_split_32x16_t256", &launch_rk<float, 32, 16, 256, true>, 128, 32,
     &launch_dp<float, 128, 32, 256, 1>},
    {32, 32, 256, 0, 1, "rk_f32_split_32x32_t256", &launch_rk<float, 32, 32, 256, true>, 128, 32,
     &launch_dp<float, 128, 32, 256, 1>},
    {32, 16, 256, 0, 1, "rk_f32_split_32x16_t256_w6", &launch_rk<float, 32, 16, 256, true>, 128, 32,
     &launch_dp<float, 128, 32, 256, 1, 6>},
    // _q: the Dornic kernel in two passes (one PTRS trial settled by its box and
    // one Gamma trial by its squeeze first; cells needing more go to a queue in
    // LDS, dp_tile_deferred)
    {32, 16, 256, 0, 1, "rk_f32_split_32x16_t256_q", &launch_rk<float, 32, 16, 256, true>, 128, 32,
     &launch_dp<float, 128, 32, 256, 1, 1, 1>},
    {32, 16, 256, 0, 1, "rk_f32_split_32x16_t256_q24", &launch_rk<float, 32, 16, 256, true>, 128, 24,
     &launch_dp<float, 128, 24, 256, 1, 1, 1>},
};

const Variant* dp_variants_f32(int* n) { *n = (int)(sizeof(kDP) / sizeof(kDP[0])); return kDP; }
const Variant* rk_variants_f32(int* n) { *n = (int)(sizeof(kRK) / sizeof(kRK[0])); return kRK; }

void launch_measure_f32(int nf, const DevInst* inst, const DevTile* tiles, int ntiles, int tx, int ty,
                        const void* fields, double2* partials, hipStream_t s)
{
    if (nf == 1)
        hipLaunchKernelGGL((measure_kernel<float, 1>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const float*)fields, partials);
    else
        hipLaunchKernelGGL((measure_kernel<float, 3>), dim3(ntiles), dim3(256), 0, s, inst, tiles, ntiles, tx,
                           ty, (const float*)fields, partials);
}

}  // namespace tmk
