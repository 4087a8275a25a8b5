// tm_frames.hip -- initial frames generated on the device, every instance of a
// batch in one launch, one thread per cell.
//
// * Speckle frames: each cell draws one uniform u from Philox(seed, instance_id;
//   cell, 0xFFFFFFFE, 0xFFFFFFFF, 0) (word 0, (w + 0.5) 2^-32) and every field
//   takes its "high" value where u <= perc, its "low" value elsewhere -- the
//   rule of init_randconstframe (rietkerk_bjork_FKE.cpp:409-431), which the
//   reference's bistable and mean-field speckle frames reuse with other values.
// * Gaussian frames: fields f < n_gauss hold amp_f exp(-((p - cx_f)^2 + (q -
//   cy_f)^2) / (2 sd_f sd_f)) at cell i = p L + q, the others amp_f
//   (init_gaussiantear, rietkerk_bjork_FKE.cpp:1185-1199).  Centres go round the
//   corners from the top left, cx = [L/8, 7L/8, 7L/8, L/8] and cy = [L/8, L/8,
//   7L/8, 7L/8] in integer division (:1110-1115); sd_f = L / sd_div_f.
// Values are formed in FP64, in the reference's operation order with this
// library's own exp, and stored in the field precision.
// * Uniform frames (tm_init_uniform): word f % 4 of Philox(seed, instance_id;
//   cell, 0xFFFFFFFF, 0xFFFFFFFF, f / 4) gives field f = lo_f + (hi_f - lo_f) u.
// Workgroup b covers kFrameCellsPerBlock consecutive cells of the instance k
// with blk0[k] <= b < blk0[k+1] (its 256 threads stride through them, so every
// store is a contiguous wave); the values of the instance's parameter point
// come from vals[params_index].
#include <hip/hip_runtime.h>
#include "tm_launch.h"

namespace tmk {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void init_frames_kernel(int mode, const DevInst* __restrict__ inst,
                                                          const int64_t* __restrict__ blk0, int n_inst, int nf,
                                                          int n_gauss, FrameSd sd_div,
                                                          const FrameVals* __restrict__ vals,
                                                          T* __restrict__ fields)
{
#pragma clang fp contract(off)
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n_inst;            // blk0[lo] <= b < blk0[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid;
    }
    const DevInst d = inst[lo];
    const int L = d.L;
    const int64_t N = (int64_t)L * L;
    T* f0 = fields + (int64_t)nf * d.cell_off;
    const FrameVals v = vals[d.pidx];
    const int64_t base = (b - blk0[lo]) * kFrameCellsPerBlock;
    for (int r = 0; r < kFrameCellsPerBlock / 256; ++r) {
        const int64_t i = base + r * 256 + threadIdx.x;
        if (i >= N) return;
        if (mode != 1) {
            const uint32_t c1 = mode == 0 ? 0xFFFFFFFEu : 0xFFFFFFFFu;
            const Philox4 w = philox4x32_10((uint32_t)i, c1, 0xFFFFFFFFu, 0u, d.seed, d.inst_id);
            if (mode == 0) {            // speckle: word 0 decides high or low for every field
                const double u = ((double)w.v0 + 0.5) * 0x1p-32;
                const bool high = u <= v.perc;
                for (int f = 0; f < nf; ++f) f0[f * N + i] = (T)(high ? v.a[f] : v.b[f]);
            } else {                    // uniform: word f % 4 of block f / 4 gives field f in [a_f, b_f)
                Philox4 wb = w;
                for (int f = 0; f < nf; ++f) {
                    if (f > 0 && f % 4 == 0)
                        wb = philox4x32_10((uint32_t)i, c1, 0xFFFFFFFFu, (uint32_t)(f / 4), d.seed, d.inst_id);
                    const uint32_t wf = f % 4 == 0 ? wb.v0 : (f % 4 == 1 ? wb.v1 : (f % 4 == 2 ? wb.v2 : wb.v3));
                    const double u = ((double)wf + 0.5) * 0x1p-32;
                    f0[f * N + i] = (T)(v.a[f] + (v.b[f] - v.a[f]) * u);
                }
            }
            continue;
        }
        // cell indices fit 32 bits (the Philox counter word is the 32-bit cell index)
        const uint32_t iu = (uint32_t)i;
        const int p = (int)(iu / (uint32_t)L), q = (int)(iu - (uint32_t)p * (uint32_t)L);
        for (int f = 0; f < nf; ++f) {
            double x = v.a[f];
            if (f < n_gauss) {
                const double cx = (f % 4) % 3 == 0 ? (double)(L / 8) : (double)(7 * L / 8);
                const double cy = (f % 4) < 2 ? (double)(L / 8) : (double)(7 * L / 8);
                const double sd = L / sd_div.v[f];
                const double dp = p - cx, dq = q - cy;
                x = v.a[f] * dexp(-(dp * dp + dq * dq) / (2 * sd * sd));
            }
            f0[f * N + i] = (T)x;
        }
    }
}

}  // namespace

void launch_init_frames(int prec, int mode, const DevInst* inst, const int64_t* blk0, int n_inst, int64_t nblocks,
                        int nf, int n_gauss, const double* sd_div, const FrameVals* vals, void* fields,
                        hipStream_t s)
{
    if (nblocks <= 0) return;
    FrameSd sd;
    for (int f = 0; f < kMaxFields; ++f) sd.v[f] = sd_div[f];
    if (prec == TM_F64)
        hipLaunchKernelGGL(init_frames_kernel<double>, dim3((unsigned)nblocks), dim3(256), 0, s, mode, inst, blk0,
                           n_inst, nf, n_gauss, sd, vals, (double*)fields);
    else
        hipLaunchKernelGGL(init_frames_kernel<float>, dim3((unsigned)nblocks), dim3(256), 0, s, mode, inst, blk0,
                           n_inst, nf, n_gauss, sd, vals, (float*)fields);
}

}  // namespace tmk
