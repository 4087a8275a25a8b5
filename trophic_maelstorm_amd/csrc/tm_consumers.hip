// tm_consumers.hip -- the vegetation + consumer step (multiSPDP.cpp built with
// SPB = 2: vegetation and grazer; SPB = 3: and predator), FP64, performing the
// CPU restatement's operations in its order with contraction off.
//
// Fields per instance: rho_0 .. rho_{NS-1}, then the perception fields
// gamma_1 .. gamma_{NS-1} (NF = 2 NS - 1), species-major [NF][L*L].
//
// A step is three launches:
//   dpc_chunk_kernel  one workgroup per 4096-cell chunk: the chunk's sums of
//                     every density and the occupied count of the vegetation,
//                     in a fixed order (lane strides, then a pairwise fold);
//   dpc_state_kernel  one thread per instance: the chunk sums in chunk order ->
//                     spatial means, the cover factor and the number of
//                     perception neighbours each consumer uses this step
//                     (multiSPDP.cpp:2962-2975, :1705-1730);
//   dpc_step_kernel   one workgroup per tile: the tile and a halo of the
//                     perception radius of every density in LDS, then per cell
//                     the vegetation's Dornic step (:2901-2945), each consumer's
//                     perception gamma (:1699-1786, :2166-2284) and its Dornic
//                     step with upwind advection in a random direction
//                     (:3041-3190), the RK4 of the trophic terms (:1821-1898),
//                     checks and flush (:3228-3262), and rho(t) partials.
// The means steer the step (they scale gamma and switch consumers off), so they
// come from the fixed-order reduction rather than from the compensated rho(t)
// rows; the rows stay the reporting path.
#include <hip/hip_runtime.h>
#include "tm_launch.h"

namespace tmk {

namespace {

constexpr int kChunk = kDpcChunkCells;
constexpr double kGammaEps = 1.0e-12;   // calc_gamma's eps (multiSPDP.cpp:1715, :2182)

template <int NS>
__global__ __launch_bounds__(256) void dpc_chunk_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    constexpr int NF = 2 * NS - 1;
    __shared__ double red[NS + 1][256];
    const DPCWork* w = (const DPCWork*)a.work;
    const int2 ch = w->chunks[blockIdx.x];
    const DevInst di = a.inst[ch.x];
    const int64_t N = (int64_t)di.L * di.L;
    const double* f0 = (const double*)a.in + (int64_t)NF * di.cell_off;
    const int64_t base = (int64_t)ch.y * kChunk;
    const int t = threadIdx.x;
    double s[NS + 1];
#pragma unroll
    for (int f = 0; f <= NS; ++f) s[f] = 0.0;
    for (int64_t i = base + t; i < base + kChunk && i < N; i += 256) {
#pragma unroll
        for (int f = 0; f < NS; ++f) s[f] += f0[f * N + i];
        s[NS] += f0[i] > 0.0 ? 1.0 : 0.0;
    }
#pragma unroll
    for (int f = 0; f <= NS; ++f) red[f][t] = s[f];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
#pragma unroll
            for (int f = 0; f <= NS; ++f) red[f][t] += red[f][t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int f = 0; f <= NS; ++f) w->chunk[(int64_t)blockIdx.x * 4 + f] = red[f][0];
    }
}

template <int NS>
__global__ __launch_bounds__(64) void dpc_state_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.n_inst) return;
    const DPCWork* w = (const DPCWork*)a.work;
    const DPCInstInfo in = w->info[i];
    const DevInst di = a.inst[i];
    const double N = (double)((int64_t)di.L * di.L);
    double tot[NS], occ = 0.0;
#pragma unroll
    for (int f = 0; f < NS; ++f) tot[f] = 0.0;
    for (int c = 0; c < in.nchunks; ++c) {
        const double* p = w->chunk + (int64_t)(in.chunk0 + c) * 4;
#pragma unroll
        for (int f = 0; f < NS; ++f) tot[f] += p[f];
        occ += p[NS];
    }
    DPCStepState st{};
    for (int f = 0; f < NS; ++f) st.mean[f] = tot[f] / N;
    double nfac = 1 - occ / N;                       // multiSPDP.cpp:2972-2974
    if (nfac < 0.35) nfac = 0.35;
    const int new_len = (int)(nfac * nfac * (double)in.disk_len);
    for (int f = 0; f < NS; ++f) st.inv[f] = st.mean[f] >= kGammaEps ? 1 / st.mean[f] : 1 / kGammaEps;
    for (int f = 0; f < NS; ++f) {
        int e = (int)(in.fr[f] * in.fr[f] * (double)new_len);
        if (NS == 3 && e < 1 && in.fr[f] > 0) e = 1;     // multiSPDP.cpp:2215-2216
        st.eff[f] = e;
    }
    w->st[i] = st;
}

// f_DP_Dor_2Sp / f_DP_Dor_3Sp (multiSPDP.cpp:1670-1697, :1789-1819) at one cell.
template <int NS>
__device__ __forceinline__ void dpc_f(const tm_dpc_params& p, const double* y, double* f)
{
#pragma clang fp contract(off)
    const double A01H01 = p.A[0][1] * p.H[0][1];
    const double EA1 = p.E[1] * p.A[0][1];
    f[0] = -p.b * y[0] * y[0] - (p.A[0][1] * y[0] / (1 + A01H01 * y[0])) * y[1];
    if constexpr (NS == 2) {
        f[1] = (EA1 * y[0] / (1 + A01H01 * y[0])) * y[1];
    } else {
        const double A12H12 = p.A[1][2] * p.H[1][2];
        const double EA2 = p.E[2] * p.A[1][2];
        f[1] = (EA1 * y[0] / (1 + A01H01 * y[0])) * y[1] - p.A[1][2] * y[1] / (1 + A12H12 * y[1]) * y[2];
        f[2] = (EA2 * y[1] / (1 + A12H12 * y[1])) * y[2];
    }
}

// RK4_Integrate_Stochastic_MultiSp (multiSPDP.cpp:1821-1898): pointwise.
template <int NS>
__device__ __forceinline__ void dpc_rk4(const tm_dpc_params& p, double* y)
{
#pragma clang fp contract(off)
    const double dt = p.dt, dt2 = dt / 2.0, dt6 = dt / 6.0;
    double k1[NS], k2[NS], k3[NS], k4[NS], ys[NS];
    dpc_f<NS>(p, y, k1);
#pragma unroll
    for (int s = 0; s < NS; ++s) ys[s] = y[s] + (dt2) * k1[s];
    dpc_f<NS>(p, ys, k2);
#pragma unroll
    for (int s = 0; s < NS; ++s) ys[s] = y[s] + (dt2) * k2[s];
    dpc_f<NS>(p, ys, k3);
#pragma unroll
    for (int s = 0; s < NS; ++s) ys[s] = y[s] + (dt) * k3[s];
    dpc_f<NS>(p, ys, k4);
#pragma unroll
    for (int s = 0; s < NS; ++s) y[s] += (dt6) * (k1[s] + 2.0 * k2[s] + 2.0 * k3[s] + k4[s]);
}

template <int NS, int TX, int TY, int NT, int MINW>
__global__ __launch_bounds__(NT, MINW) void dpc_step_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    constexpr int NF = 2 * NS - 1;
    extern __shared__ __align__(16) double s_lds[];
    __shared__ double s_red[2 * NF * (NT / 64)];

    const int64_t counter = a.state->counter + a.k;
    const int64_t slot = counter & (int64_t)(a.nslots - 1);
    const int tile = blockIdx.x;
    if (tile >= a.ntiles) return;
    const DevTile tl = a.tiles[tile];
    const DevInst di = a.inst[tl.inst];
    const DPCConst c = ((const DPCConst*)a.consts)[(int64_t)a.k * a.n_inst + tl.inst];
    const DPCWork* w = (const DPCWork*)a.work;
    const tm_dpc_params p = w->params[di.pidx];
    const DPCInstInfo info = w->info[tl.inst];
    const DPCStepState st = w->st[tl.inst];
    const int H = a.halo;
    const int RW = TX + 2 * H, RH = TY + 2 * H, IMG = RW * RH;
    double* img = s_lds;                        // [NS][RH][RW]
    int* offs = (int*)(s_lds + NS * IMG);       // perception offsets as LDS strides
    const int L = tl.L;
    const int64_t N = (int64_t)L * L;
    const double* fin = (const double*)a.in + (int64_t)NF * tl.cell_off;
    double* fout = (double*)a.out + (int64_t)NF * tl.cell_off;

    for (int e = threadIdx.x; e < IMG; e += NT) {
        const int ry = e / RW, rx = e - (e / RW) * RW;
        const int gy = wrap(tl.y0 + ry - H, L), gx = wrap(tl.x0 + rx - H, L);
        const int64_t g = (int64_t)gy * L + gx;
#pragma unroll
        for (int s = 0; s < NS; ++s) img[s * IMG + e] = fin[s * N + g];
    }
    for (int k = threadIdx.x; k < info.disk_len; k += NT) {
        const int2 o = w->disk[info.disk_off + k];
        offs[k] = o.x * RW + o.y;
    }
    __syncthreads();

    const double dx = p.dx, dt = p.dt;
    const double dx2 = dx * dx, dx1 = 1 / dx, dx1_2 = 1 / (dx2), dxb2 = dx / 2.0;
    const double Dc0 = p.D[0] / (dx2);
    double sum[NF], cnt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) { sum[f] = 0.0; cnt[f] = 0.0; }

    for (int e = threadIdx.x; e < TX * TY; e += NT) {
        const int ly = e / TX, lx = e % TX;
        const int gy = tl.y0 + ly, gx = tl.x0 + lx;
        if (gy >= L || gx >= L) continue;
        const uint32_t cell = (uint32_t)gy * (uint32_t)L + (uint32_t)gx;
        const int r = (ly + H) * RW + (lx + H);
        double y[NS];
        // vegetation: Dornic (multiSPDP.cpp:2901-2945)
        {
            const double cur = img[r];
            const double alpha = Dc0 * (img[r - RW] + img[r + RW] + img[r - 1] + img[r + 1]);
            y[0] = cur;
            if (!(alpha == 0.0 && cur == 0.0)) {
                Stream s0(di.seed, di.inst_id, cell, c.step);
                int64_t n;
                y[0] = dornic_cell(s0, alpha, cur, c.k, c.compat != 0, n);
            }
        }
        // perception gamma of each consumer, from the old fields
        double gam[NS];
        gam[0] = 0.0;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            double g = fin[(NS + s - 1) * N + cell];
            if (info.fr[s] == 0.0) { gam[s] = g; continue; }
            if (st.mean[s] < kGammaEps || st.mean[s - 1] < kGammaEps) { gam[s] = 0.0; continue; }
            const double* src = img + (s - 1) * IMG + r;
            const int eff = st.eff[s];
            if constexpr (NS == 2) {            // accumulates onto the previous value (:1756-1760)
                for (int k = 0; k < eff; ++k) g += src[offs[k]];
                g = (g * st.inv[s - 1]) / (double)eff;
            } else if (s == 1) {                // grazers, weighed down by predators (:2228-2249)
                double acc = 0.0;
                if (st.mean[2] < kGammaEps) {
                    for (int k = 0; k < eff; ++k) acc += src[offs[k]];
                } else {
                    const double* pr = img + 2 * IMG + r;
                    for (int k = 0; k < eff; ++k) acc += src[offs[k]] * (1 - pr[offs[k]] * st.inv[2]);
                }
                g = (acc * st.inv[s - 1]) / (double)eff;
            } else {                            // predators (:2251-2258)
                double acc = 0.0;
                for (int k = 0; k < eff; ++k) acc += src[offs[k]];
                g = (acc / (double)eff) * st.inv[s - 1];
            }
            if (g > 1) g = 1.0;
            else if (g < 0) g = 0.0;
            gam[s] = g;
        }
        // consumers: Dornic with upwind advection in a random direction (:3041-3190)
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            const double* is = img + s * IMG;
            y[s] = is[r];
            if (st.mean[s] == 0.0) continue;
            if (gam[s] > 1) gam[s] = 1;
            Stream sc(di.seed, di.inst_id, cell, c.step);
            sc.blk = (uint32_t)s << 24;
            double sn, cs;
            dsincos2pi(sc.u64(), sn, cs);
            const double vx = p.v[s] * cs, vy = p.v[s] * sn;
            const double vxa = vx * (double)((0 < vx) - (vx < 0));
            const double vya = vy * (double)((0 < vy) - (vy < 0));
            const double gp = (1 - gam[s]);
            double Dx = (p.D[s] - (dxb2) * (gp * vxa * (1 - dt * dx1 * gp * vxa))) * dx1_2;
            double Dy = (p.D[s] - (dxb2) * (gp * vya * (1 - dt * dx1 * gp * vya))) * dx1_2;
            if (Dx < Dc0) Dx = Dc0;
            if (Dy < Dc0) Dy = Dc0;
            const double ux = vx > 0 ? is[r - 1] : is[r + 1];
            const double uy = vy > 0 ? is[r - RW] : is[r + RW];
            const double alpha = gam[s] * (Dy * (is[r - RW] + is[r + RW]) + Dx * (is[r - 1] + is[r + 1]))
                                 + dx1 * (gp) * (vxa * ux + vya * uy);
            if (alpha == 0 && is[r] == 0) continue;
            const double beta = -p.M[s] - 2 * gam[s] * (Dx + Dy) - (1 - gam[s]) * (vxa + vya) * dx1;
            const double lexp = dexp((beta) * dt);
            const double sig2_1 = 1 / (p.sigma[s] * p.sigma[s]);
            const double lam = 2 * (beta) * (sig2_1) / (lexp - 1.0);
            DornicConst k{};
            k.lambda = lam;
            k.ll = lam * lexp;
            k.sigma2_1 = sig2_1;
            int64_t n;
            y[s] = dornic_cell(sc, alpha, is[r], k, c.compat != 0, n);
        }
        dpc_rk4<NS>(p, y);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double v = y[s];
            if (v < 0 || !__builtin_isfinite(v)) record_fault(a.state, tl.inst, s * N + cell, counter);
            if (v < c.eps && v >= c.neg) v = 0.0;
            fout[s * N + cell] = v;
            sum[s] += v;
            cnt[s] += v > 0.0 ? 1.0 : 0.0;
        }
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            fout[(NS + s - 1) * N + cell] = gam[s];
            sum[NS + s - 1] += gam[s];
            cnt[NS + s - 1] += gam[s] > 0.0 ? 1.0 : 0.0;
        }
    }
    block_partials<NT, NF>(sum, cnt, a.partials + (slot * a.ntiles + tile) * NF, s_red);
}

template <int NS, int TX, int TY, int NT, int MINW = 1>
void launch_dpc(const StepArgs& a, int grid, hipStream_t s)
{
    const size_t lds = dpc_lds_bytes(NS, TX, TY, a.halo);
    if (lds > 65536) {   // above the default dynamic-LDS limit (CDNA4 has 160 KiB per CU)
        static size_t granted = 0;
        if (lds > granted &&
            hipFuncSetAttribute((const void*)dpc_step_kernel<NS, TX, TY, NT, MINW>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess)
            granted = lds;
    }
    hipLaunchKernelGGL(dpc_chunk_kernel<NS>, dim3(a.nchunks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(dpc_state_kernel<NS>, dim3((a.n_inst + 63) / 64), dim3(64), 0, s, a);
    hipLaunchKernelGGL((dpc_step_kernel<NS, TX, TY, NT, MINW>), dim3(grid), dim3(NT), lds, s, a);
}

// _wN: at most 512/N VGPRs, N waves per SIMD
const Variant kDPC2[] = {
    {32, 16, 256, 0, 1, "dpc2_f64_32x16_t256", &launch_dpc<2, 32, 16, 256>, 0, 0, nullptr, 1},
    {64, 8, 256, 0, 1, "dpc2_f64_64x8_t256", &launch_dpc<2, 64, 8, 256>, 0, 0, nullptr, 1},
    {32, 16, 256, 0, 1, "dpc2_f64_32x16_t256_w4", &launch_dpc<2, 32, 16, 256, 4>, 0, 0, nullptr, 1},
    {32, 16, 256, 0, 1, "dpc2_f64_32x16_t256_w5", &launch_dpc<2, 32, 16, 256, 5>, 0, 0, nullptr, 1},
    {32, 16, 256, 0, 1, "dpc2_f64_32x16_t256_w6", &launch_dpc<2, 32, 16, 256, 6>, 0, 0, nullptr, 1},
};
const Variant kDPC3[] = {
    {32, 16, 256, 0, 1, "dpc3_f64_32x16_t256", &launch_dpc<3, 32, 16, 256>, 0, 0, nullptr, 1},
    {64, 8, 256, 0, 1, "dpc3_f64_64x8_t256", &launch_dpc<3, 64, 8, 256>, 0, 0, nullptr, 1},
    {32, 16, 256, 0, 1, "dpc3_f64_32x16_t256_w4", &launch_dpc<3, 32, 16, 256, 4>, 0, 0, nullptr, 1},
    {32, 16, 256, 0, 1, "dpc3_f64_32x16_t256_w5", &launch_dpc<3, 32, 16, 256, 5>, 0, 0, nullptr, 1},
};

}  // namespace

const Variant* dpc2_variants_f64(int* n) { *n = (int)(sizeof(kDPC2) / sizeof(kDPC2[0])); return kDPC2; }
const Variant* dpc3_variants_f64(int* n) { *n = (int)(sizeof(kDPC3) / sizeof(kDPC3[0])); return kDPC3; }

}  // namespace tmk
