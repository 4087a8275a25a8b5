// tm_ensemble.hip -- measurement series and across-replicate statistics of a
// parameter sweep, on the device.
//
// A sweep batch holds every (parameter point, replicate) as one instance.  The
// step kernels already fold each instance's rho(t) row every step; the series
// kernel keeps the rows at the measurement steps, and the statistics kernel
// folds them over the replicates of each parameter point in replicate order,
// with the recurrences of the reference's running statistics.  All of it is
// FP64 with contraction off, in the reference's operation order, so it agrees
// bit for bit with a CPU evaluation of the same recurrences.
//
// Data are small (instances x measurements x 2 nfields doubles); these kernels
// are latency-bound and exist so the series never leaves the device.
#include <hip/hip_runtime.h>
#include "tm_launch.h"

namespace tmk {
namespace {

// series[i][m][:] = rows[i][steps[m] - s0 - 1][:] for m in [m0, m1): the rows of
// one advance() chunk that began s0 steps into the series.
__global__ __launch_bounds__(256) void gather_series_kernel(const double* __restrict__ rows, int64_t rows_cap,
                                                            int w, const int64_t* __restrict__ steps, int m0,
                                                            int m1, int64_t s0, int n_meas, int n_inst,
                                                            double* __restrict__ series)
{
    const int64_t nm = m1 - m0;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (int64_t)n_inst * nm * w) return;
    const int c = (int)(k % w);
    const int64_t q = k / w;
    const int m = m0 + (int)(q % nm);
    const int i = (int)(q / nm);
    const int64_t row = steps[m] - s0 - 1;
    series[((int64_t)i * n_meas + m) * w + c] = rows[((int64_t)i * rows_cap + row) * w + c];
}

// First measurement at which field f (f < nsp) has no occupied cell (n_meas if
// none): first[i * nsp + f].
__global__ __launch_bounds__(256) void first_extinct_kernel(const double* __restrict__ series, int n_inst,
                                                            int n_meas, int w, int nsp, int32_t* __restrict__ first)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n_inst * nsp) return;
    const int i = k / nsp, f = k % nsp;
    const double* s = series + (int64_t)i * n_meas * w + 2 * f;
    int m = 0;
    while (m < n_meas && s[(int64_t)m * w] != 0.0) ++m;
    first[k] = m;
}

// One thread per (parameter point, measurement).  Replicates j = 0, 1, ... are
// the point's instances in instance order (grp_inst[grp_off[g] .. grp_off[g+1])).
//
//   j == 0 (rietkerk_bjork_FKE.cpp:4578-4597): mean <- x, var <- 0,
//          active <- n, surviving <- (n > 0), other-field means <- their values.
//   j >= 1 (rietkerk_bjork_FKE.cpp:1319-1369): other-field means
//          <- (x_f + j mean_f) / (j + 1) always; the field-0 update only while the
//          replicate has been occupied at every measurement so far:
//          r = surviving + 1; first survivor: mean <- x, var <- 0, active <- n;
//          else mean <- (x + (r-1) mean)/r, active <- (n + (r-1) active)/r,
//          var <- (mean - mean_prev)^2 + k ((r-2) var + (x - mean)^2) with
//          k = 1/(r-1) in integer arithmetic (reference) or real division;
//          surviving <- r.
// out row: [means of fields 1..nf-1, mean, var, surviving, active].
__global__ __launch_bounds__(256) void replicate_stats_kernel(const double* __restrict__ series, int n_meas,
                                                              int nf, const int32_t* __restrict__ grp_off,
                                                              const int32_t* __restrict__ grp_inst,
                                                              const int32_t* __restrict__ first_extinct,
                                                              int n_groups, int flags, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (int64_t)n_groups * n_meas) return;
    const int g = (int)(k / n_meas), m = (int)(k % n_meas);
    const int w = 2 * nf;
    const bool intdiv = (flags & TM_STATS_INTDIV_VARIANCE) != 0;
    const bool stop = (flags & TM_STATS_STOP_AT_EXTINCTION) != 0;
    double other[2] = {0.0, 0.0};
    double mean = 0.0, var = 0.0, surv = 0.0, act = 0.0;
    const int b = grp_off[g], e = grp_off[g + 1];
    for (int j = 0; j < e - b; ++j) {
        const int i = grp_inst[b + j];
        const int fe = first_extinct[i];
        const bool zero = stop && m > fe;
        const double* x = series + ((int64_t)i * n_meas + m) * w;
        const double xn = zero ? 0.0 : x[0], xm = zero ? 0.0 : x[1];
        if (j == 0) {
            for (int f = 1; f < nf; ++f) other[f - 1] = zero ? 0.0 : x[2 * f + 1];
            mean = xm;
            var = 0.0;
            act = xn;
            surv = xn > 0.0 ? 1.0 : 0.0;
            continue;
        }
        for (int f = 1; f < nf; ++f)
            other[f - 1] = ((zero ? 0.0 : x[2 * f + 1]) + (double)j * other[f - 1]) / (double)(j + 1);
        if (m >= fe) continue;   // not occupied at m or earlier: no longer a survivor
        const int r = (int)(surv + 1.0);
        const double mp = mean;
        if (surv == 0.0) {
            mean = xm;
            var = 0.0;
            act = xn;
        } else {
            mean = (xm + (double)(r - 1) * mean) / (double)r;
            act = (xn + (double)(r - 1) * act) / (double)r;
            const double d = xm - mean;
            const double inner = (double)(r - 2) * var + d * d;
            var = intdiv ? (mean - mp) * (mean - mp) + (double)(1 / (r - 1)) * inner
                         : (mean - mp) * (mean - mp) + inner / (double)(r - 1);
        }
        surv = (double)r;
    }
    double* o = out + k * (nf + 3);
    for (int f = 1; f < nf; ++f) o[f - 1] = other[f - 1];
    o[nf - 1] = mean;
    o[nf] = var;
    o[nf + 1] = surv;
    o[nf + 2] = act;
}

// The multi-species form (multiSPDP.cpp:3376-3391 first replicate,
// :1080-1123 later ones): every species f < nsp gets its own (mean, variance,
// surviving count, mean occupied count) over the replicates still occupied by f,
// each stopping at its own first extinct measurement.  With
// TM_STATS_STOP_AT_EXTINCTION a replicate's rows after the vegetation's (field
// 0's) first extinct measurement read as zero for every species.
// One thread per (parameter point, measurement, species); out row [4 nsp].
__global__ __launch_bounds__(256) void species_stats_kernel(const double* __restrict__ series, int n_meas, int nf,
                                                            int nsp, const int32_t* __restrict__ grp_off,
                                                            const int32_t* __restrict__ grp_inst,
                                                            const int32_t* __restrict__ first_extinct,
                                                            int n_groups, int flags, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (int64_t)n_groups * n_meas * nsp) return;
    const int f = (int)(k % nsp);
    const int m = (int)((k / nsp) % n_meas);
    const int g = (int)(k / ((int64_t)nsp * n_meas));
    const int w = 2 * nf;
    const bool intdiv = (flags & TM_STATS_INTDIV_VARIANCE) != 0;
    const bool stop = (flags & TM_STATS_STOP_AT_EXTINCTION) != 0;
    double mean = 0.0, var = 0.0, surv = 0.0, act = 0.0;
    const int b = grp_off[g], e = grp_off[g + 1];
    for (int j = 0; j < e - b; ++j) {
        const int i = grp_inst[b + j];
        const int fe0 = first_extinct[i * nsp];
        int fe = first_extinct[i * nsp + f];
        if (stop && fe0 + 1 < fe) fe = fe0 + 1;     // zero rows after the vegetation's end
        const bool zero = stop && m > fe0;
        const double* x = series + ((int64_t)i * n_meas + m) * w + 2 * f;
        const double xn = zero ? 0.0 : x[0], xm = zero ? 0.0 : x[1];
        if (j == 0) {
            mean = xm;
            var = 0.0;
            act = xn;
            surv = xn > 0.0 ? 1.0 : 0.0;
            continue;
        }
        if (m >= fe) continue;
        const int r = (int)(surv + 1.0);
        const double mp = mean;
        if (surv == 0.0) {
            mean = xm;
            var = 0.0;
            act = xn;
        } else {
            mean = (xm + (double)(r - 1) * mean) / (double)r;
            act = (xn + (double)(r - 1) * act) / (double)r;
            const double d = xm - mean;
            const double inner = (double)(r - 2) * var + d * d;
            var = intdiv ? (mean - mp) * (mean - mp) + (double)(1 / (r - 1)) * inner
                         : (mean - mp) * (mean - mp) + inner / (double)(r - 1);
        }
        surv = (double)r;
    }
    double* o = out + ((int64_t)g * n_meas + m) * 4 * nsp + 4 * f;
    o[0] = mean;
    o[1] = var;
    o[2] = surv;
    o[3] = act;
}

// Sum over the instances of field 0's occupied count in row `row` of the rows
// buffer, in a fixed order (one workgroup): the live measure of the adaptive
// multi-step launches.
__global__ __launch_bounds__(256) void live_count_kernel(const double* __restrict__ rows, int64_t rows_cap, int w,
                                                         int64_t row, int n_inst, double* __restrict__ out)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_inst; i += 256) acc += rows[((int64_t)i * rows_cap + row) * w];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

inline int grid_of(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

void launch_live_count(const double* rows, int64_t rows_cap, int w, int64_t row, int n_inst, double* out,
                       hipStream_t s)
{
    hipLaunchKernelGGL(live_count_kernel, dim3(1), dim3(256), 0, s, rows, rows_cap, w, row, n_inst, out);
}

void launch_gather_series(const double* rows, int64_t rows_cap, int w, const int64_t* steps, int m0, int m1,
                          int64_t s0, int n_meas, int n_inst, double* series, hipStream_t s)
{
    const int64_t n = (int64_t)n_inst * (m1 - m0) * w;
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_series_kernel, dim3(grid_of(n)), dim3(256), 0, s, rows, rows_cap, w, steps, m0, m1,
                       s0, n_meas, n_inst, series);
}

void launch_replicate_stats(const double* series, int n_inst, int n_meas, int nf, int nsp, const int32_t* grp_off,
                            const int32_t* grp_inst, int n_groups, int flags, int32_t* first_extinct,
                            double* out, hipStream_t s)
{
    const int nfe = nsp > 0 ? nsp : 1;
    if (n_inst > 0)
        hipLaunchKernelGGL(first_extinct_kernel, dim3(grid_of((int64_t)n_inst * nfe)), dim3(256), 0, s, series,
                           n_inst, n_meas, 2 * nf, nfe, first_extinct);
    if (nsp > 0) {
        const int64_t n = (int64_t)n_groups * n_meas * nsp;
        if (n > 0)
            hipLaunchKernelGGL(species_stats_kernel, dim3(grid_of(n)), dim3(256), 0, s, series, n_meas, nf, nsp,
                               grp_off, grp_inst, first_extinct, n_groups, flags, out);
        return;
    }
    const int64_t n = (int64_t)n_groups * n_meas;
    if (n > 0)
        hipLaunchKernelGGL(replicate_stats_kernel, dim3(grid_of(n)), dim3(256), 0, s, series, n_meas, nf, grp_off,
                           grp_inst, first_extinct, n_groups, flags, out);
}

}  // namespace tmk
