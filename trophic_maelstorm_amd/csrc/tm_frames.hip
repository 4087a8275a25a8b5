// tm_frames.hip -- initial frames generated on the device, one thread per cell.
//
// * Speckle frames: each cell draws one uniform u from Philox(seed, instance_id;
//   cell, 0xFFFFFFFE, 0xFFFFFFFF, 0) (word 0, (w + 0.5) 2^-32) and every field
//   takes its "high" value where u <= perc, its "low" value elsewhere -- the
//   rule of init_randconstframe (rietkerk_bjork_FKE.cpp:409-431), which the
//   reference's bistable and mean-field speckle frames reuse with other values.
// * Gaussian frames: fields f < n_gauss hold amp_f exp(-((p - cx_f)^2 + (q -
//   cy_f)^2) / (2 sd_f sd_f)) at cell i = p L + q, the others amp_f
//   (init_gaussiantear, rietkerk_bjork_FKE.cpp:1185-1199).
// Values are formed in FP64, in the reference's operation order with this
// library's own exp, and stored in the field precision.
#include <hip/hip_runtime.h>
#include "tm_launch.h"

namespace tmk {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void init_speckle_kernel(int L, int nf, uint32_t seed, uint32_t inst_id,
                                                           FrameVals v, T* __restrict__ fields)
{
#pragma clang fp contract(off)
    const int64_t N = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const Philox4 w = philox4x32_10((uint32_t)i, 0xFFFFFFFEu, 0xFFFFFFFFu, 0u, seed, inst_id);
    const double u = ((double)w.v0 + 0.5) * 0x1p-32;
    const bool high = u <= v.perc;
    for (int f = 0; f < nf; ++f) fields[f * N + i] = (T)(high ? v.a[f] : v.b[f]);
}

template <typename T>
__global__ __launch_bounds__(256) void init_gauss_kernel(int L, int nf, int n_gauss, FrameVals v,
                                                         T* __restrict__ fields)
{
#pragma clang fp contract(off)
    const int64_t N = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int p = (int)(i / L), q = (int)(i % L);
    for (int f = 0; f < nf; ++f) {
        double x = v.a[f];
        if (f < n_gauss) {
            const double dp = p - v.cx[f], dq = q - v.cy[f];
            x = v.a[f] * dexp(-(dp * dp + dq * dq) / (2 * v.sd[f] * v.sd[f]));
        }
        fields[f * N + i] = (T)x;
    }
}

inline int cells_grid(int L) { return (int)(((int64_t)L * L + 255) / 256); }

}  // namespace

void launch_init_speckle(int prec, int nf, int L, uint32_t seed, uint32_t inst_id, const FrameVals& v,
                         void* fields, hipStream_t s)
{
    if (prec == TM_F64)
        hipLaunchKernelGGL(init_speckle_kernel<double>, dim3(cells_grid(L)), dim3(256), 0, s, L, nf, seed, inst_id,
                           v, (double*)fields);
    else
        hipLaunchKernelGGL(init_speckle_kernel<float>, dim3(cells_grid(L)), dim3(256), 0, s, L, nf, seed, inst_id,
                           v, (float*)fields);
}

void launch_init_gauss(int prec, int nf, int L, int n_gauss, const FrameVals& v, void* fields, hipStream_t s)
{
    if (prec == TM_F64)
        hipLaunchKernelGGL(init_gauss_kernel<double>, dim3(cells_grid(L)), dim3(256), 0, s, L, nf, n_gauss, v,
                           (double*)fields);
    else
        hipLaunchKernelGGL(init_gauss_kernel<float>, dim3(cells_grid(L)), dim3(256), 0, s, L, nf, n_gauss, v,
                           (float*)fields);
}

}  // namespace tmk
