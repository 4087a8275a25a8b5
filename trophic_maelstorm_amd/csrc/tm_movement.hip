// tm_movement.hip -- movement of the Rietkerk model's consumer species: the
// analytic Fokker-Planck step of advdiff_FKE_MultiSp
// (Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3614-3695), whose CUDA twin
// updateMovement (kernel_advdiff_FKE2D.cu:92-180) is the reference's only GPU
// code.
//
// Each source cell i with rho_i >= 1e-10 spreads rho_i over the lattice with
// weight A_i(x) B_i(y), where the row kernel is
//   A_i(x) = gamma_i SD(x - c_i) + (1 - gamma_i) SV(x - mu_x(i))
// (SD, SV: the diffusion and advection half-stencils, mu_x(i) the rounded
// advected mean) and B_i the same in columns.  Expanded, that is the
// reference's four scatter terms (gamma^2, twice gamma - gamma^2, (1 - gamma)^2).
// The reference scatters with atomicAdd; this kernel gathers instead: a first
// pass forms each source's term coefficients and advected mean once, then every
// output cell loops over the sources within reach -- those with |x - c_i| <=
// W and |y - c_j| <= W, W = max(rD, rV + largest shift) -- and adds their
// terms in FP64, in a fixed order.  No atomics, and the result is
// deterministic.  The sums differ from the reference's sequential scatter only
// in their order.
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <vector>

#include "tm/tm.h"
#include "tm_device.h"
#include "tm_kernels.h"

namespace {

constexpr double kMoveEps = 1.0e-10;   // epsilon, rietkerk_bjork_basic.h:119

struct MoveSpecies {
    int rD, rV, W, offD, offV;
    double pref;
};
struct MoveArgs {
    MoveSpecies sp[3];
    int L;
};

// x - m (both in [0, L)) as its representative in (-L/2, L/2]: the one offset that
// matters while every reach stays below L/2.
__device__ __forceinline__ int signed_gap(int x, int m, int L)
{
    int e = x - m;
    if (2 * e > L) e -= L;
    if (2 * e <= -L) e += L;
    return e;
}

__device__ __forceinline__ int wrapi(int v, int L)
{
    v %= L;
    return v < 0 ? v + L : v;
}

// Per source cell: its three term coefficients (gamma^2 rho, (gamma - gamma^2) rho,
// (1 - gamma)^2 rho, formed as the reference forms them) and its advected mean
// cell, once per call; mu.x = -1 marks a source below the 1e-10 floor.
__global__ __launch_bounds__(256) void fke_move_prep_kernel(MoveArgs a, const double* __restrict__ rho,
                                                            const double* __restrict__ gamma,
                                                            const double2* __restrict__ veff,
                                                            double4* __restrict__ coef, int2* __restrict__ mu)
{
#pragma clang fp contract(off)
    const int s = blockIdx.y;
    const int L = a.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= L2) return;
    const double r = rho[s * L2 + i];
    if (r < kMoveEps) {
        mu[s * L2 + i] = make_int2(-1, -1);
        return;
    }
    const double pref = a.sp[s].pref;
    const int ci = (int)(i / L), cj = (int)(i % L);
    const double2 v = veff[s * L2 + i];
    const int mux = (int)(__builtin_round(ci + v.x * (pref)) + L) % L;
    const int muy = (int)(__builtin_round(cj + v.y * (pref)) + L) % L;
    const double g = gamma[s * L2 + i], g2 = g * g;
    coef[s * L2 + i] = make_double4((g2) * r, (g - g2) * r, (1 - g) * (1 - g) * r, 0.0);
    mu[s * L2 + i] = make_int2(mux, muy);
}

// v in [-L, 2L) -> v mod L
__device__ __forceinline__ int wrap1(int v, int L)
{
    v += v < 0 ? L : 0;
    return v >= L ? v - L : v;
}

__global__ __launch_bounds__(256) void fke_move_kernel(MoveArgs a, const double4* __restrict__ coef,
                                                       const int2* __restrict__ mu,
                                                       const double* __restrict__ SDall,
                                                       const double* __restrict__ SVall, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int s = blockIdx.y;
    const int L = a.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= L2) return;
    const MoveSpecies sp = a.sp[s];
    const double* SD = SDall + sp.offD;
    const double* SV = SVall + sp.offV;
    const double4* cs = coef + s * L2;
    const int2* ms = mu + s * L2;
    const int x = (int)(o / L), y = (int)(o % L);
    double acc = 0.0;
    for (int dr = -sp.W; dr <= sp.W; ++dr) {
        const int ci = wrap1(x - dr, L);
        const int adr = dr < 0 ? -dr : dr;
        const bool rowD = adr <= sp.rD;
        for (int dc = -sp.W; dc <= sp.W; ++dc) {
            const int cj = wrap1(y - dc, L);
            const int64_t i = (int64_t)ci * L + cj;
            const int2 m = ms[i];
            if (m.x < 0) continue;
            const int ex = signed_gap(x, m.x, L), ey = signed_gap(y, m.y, L);
            const int adc = dc < 0 ? -dc : dc;
            const int aex = ex < 0 ? -ex : ex, aey = ey < 0 ? -ey : ey;
            const bool rowV = aex <= sp.rV, colD = adc <= sp.rD, colV = aey <= sp.rV;
            if (!((rowD || rowV) && (colD || colV))) continue;
            const double4 c = cs[i];
            double t = 0.0;
            if (rowD && colD) t += c.x * (SD[adr] * SD[adc]);
            if (rowD && colV) t += c.y * (SD[adr] * SV[aey]);
            if (rowV && colD) t += c.y * (SV[aex] * SD[adc]);
            if (rowV && colV) t += c.z * (SV[aex] * SV[aey]);
            acc += t;
        }
    }
    out[s * L2 + o] = acc;
}

// The largest |vx|, |vy| of each species (one workgroup per species); a NaN or
// infinite velocity gives +inf, which the host refuses.
__global__ __launch_bounds__(256) void vmax_kernel(const double2* __restrict__ veff, int64_t L2, double* out)
{
    __shared__ double red[256];
    const int s = blockIdx.x;
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < L2; i += 256) {
        const double2 v = veff[s * L2 + i];
        const double a = __builtin_fabs(v.x), b = __builtin_fabs(v.y);
        double c = a > b ? a : b;
        if (!(a <= 1e300) || !(b <= 1e300)) c = __builtin_inf();   // NaN or infinite
        m = c > m ? c : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + o] ? red[threadIdx.x]
                                                                                         : red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

// Kernel arguments of ns species from their constants and largest speeds; the
// stencils' offsets in the concatenated arrays.  TM_ERR_ARG when a reach does
// not fit the lattice.
int move_args(int ns, int L, const tm_fke_move_species* sp, const double* vmax, MoveArgs* a, int* nD, int* nV)
{
    *a = MoveArgs{};
    a->L = L;
    int offD = 0, offV = 0;
    for (int s = 0; s < ns; ++s) {
        if (sp[s].rD < 0 || sp[s].rV < 0) return TM_ERR_ARG;
        const double shift = vmax[s] * std::fabs(sp[s].mu_prefactor);   // the largest shift of any source
        if (!(shift < (double)L)) return TM_ERR_ARG;                    // NaN and inf included
        const int S = (int)std::floor(shift + 0.5) + 1;
        MoveSpecies& m = a->sp[s];
        m.rD = sp[s].rD;
        m.rV = sp[s].rV;
        m.pref = sp[s].mu_prefactor;
        m.W = std::max(m.rD, m.rV + S);
        m.offD = offD;
        m.offV = offV;
        if (2 * m.W + 1 > L) return TM_ERR_ARG;   // reaches must not wrap onto themselves
        offD += m.rD + 1;
        offV += m.rV + 1;
    }
    *nD = offD;
    *nV = offV;
    return TM_OK;
}

// Perception gamma and velocities (calc_gamma_vel_2Sp/3Sp_NonRefugia,
// FKE:2534-2606, :3371-3482): one thread per cell, the perception offsets
// gathered from the old densities.  Per-species numbers of the step (means,
// inverses, neighbour counts) come from the host.
struct PerceptArgs {
    int ns, L;
    double v[3], mean[3], inv[3], fr[3];
    int eff[3];
    uint32_t seed, inst;
    int64_t step;
};

__global__ __launch_bounds__(256) void fke_perception_kernel(PerceptArgs a, const int2* __restrict__ disk,
                                                             const double* __restrict__ rho,
                                                             double* __restrict__ gamma, double2* __restrict__ veff)
{
#pragma clang fp contract(off)
    const int L = a.L;
    const int64_t N = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int p = (int)(i / L), q = (int)(i % L);
    if (a.ns == 3)
        for (int s = 0; s < 3; ++s) {
            tmk::Stream st(a.seed, a.inst, (uint32_t)i, a.step);
            st.blk = ((uint32_t)s << 24) | 0x800000u;
            double sn, cs;
            tmk::dsincos2pi(st.u64(), sn, cs);
            veff[s * N + i] = make_double2(a.v[s] * cs, a.v[s] * sn);
        }
    for (int s = 1; s < a.ns; ++s) {
        double* g = gamma + (s - 1) * N + i;
        if (a.fr[s] == 0.0) continue;
        if (a.mean[s] < 1.0e-12) { *g = 1.0; continue; }
        if (a.mean[s - 1] < 1.0e-12) { *g = 0.0; continue; }
        const double* src = rho + (s - 1) * N;
        const bool weigh = a.ns == 3 && s == 1 && !(a.mean[2] < 1.0e-12);
        double acc = *g;
        for (int k = 0; k < a.eff[s]; ++k) {
            const int2 o = disk[k];
            const int64_t j = (int64_t)wrapi(p + o.x, L) * L + wrapi(q + o.y, L);
            if (weigh) acc += src[j] * (1 - rho[2 * N + j] * a.inv[2]);
            else acc += src[j];
        }
        if (a.ns == 2 || s == 1) acc = (acc * a.inv[s - 1]) / (double)a.eff[s];
        else acc = (acc / (double)a.eff[s]) * a.inv[s - 1];
        if (acc > 1) acc = 1.0;
        else if (acc < 0) acc = 0.0;
        *g = acc;
    }
}

// The fixed-order sum the consumer models use for their spatial means (chunks
// of 4096 cells, 256 lane sums folded pairwise, chunk sums in order), so that
// the same means come out on the host and in the device reduction.
double fixed_sum(const double* x, int64_t n)
{
    double total = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += 4096) {
        double s[256];
        for (int t = 0; t < 256; ++t) {
            double a = 0.0;
            for (int64_t i = c0 + t; i < c0 + 4096 && i < n; i += 256) a += x[i];
            s[t] = a;
        }
        for (int o = 128; o > 0; o >>= 1)
            for (int t = 0; t < o; ++t) s[t] += s[t + o];
        total += s[0];
    }
    return total;
}

// precompute_gaussian (FKE:1735-1749): R + 1 normalised half-stencil values.
void gauss_stencil(double sigma, int R, double* st)
{
    const double pi = 3.141592653589793;
    double sum = 0.0;
    const double factor = 1.0 / (std::sqrt(2 * pi) * sigma);
    for (int x = 0; x <= R; ++x) {
        st[x] = factor * std::exp(-0.5 * (x * x) / (sigma * sigma));
        sum += (x > 0) ? 2 * st[x] : st[0];
    }
    for (int x = 0; x <= R; ++x) st[x] /= sum;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// ---------------------------------------------------------------------------
// Trophic RK4 of the Rietkerk consumer model: RK4_Integrate_Stochastic_2Sp /
// _MultiSp with f_DorFKE_2Sp / f_DorFKE_3Sp (rietkerk_bjork_FKE.cpp:2449-2533,
// :2611-2720).  Fields in the reference's order: P, G, (Pr,) W, O.  W and O
// diffuse, so every stage needs its neighbours' stage values: one grid-wide
// pass per stage.  The K's are folded into one accumulator in the reference's
// order ((K1 + 2 K2) + 2 K3) + K4, so no K array is kept.
// ---------------------------------------------------------------------------
struct FkeRk {
    int ns, L;                       // consumer species (1 or 2), lattice side
    double a, cgmax, gmax, alpha, rW, K1, K2, K2W0;
    double A01, A01H01, EA1, M1, A12, A12H12, EA2, M2;
    double dW, dO;                   // D / dx^2 of W and O
    double dt, dt2, dt6;             // the RK4's dt; dt2 and dt6 from Dt_Analytic (FKE:206-210)
};

// f at cell i of the NF fields y (stride L^2), in the reference's operation order.
template <int NS>
__device__ __forceinline__ void fke_rhs(const FkeRk& c, const double* y, int64_t L2, int i, int up, int dn, int lf,
                                        int rt, double (&f)[NS + 3])
{
#pragma clang fp contract(off)
    constexpr int W = NS + 1, O = NS + 2;
    const double P = y[i], G = y[L2 + i], Wi = y[W * L2 + i], Oi = y[O * L2 + i];
    const double* yW = y + W * L2;
    const double* yO = y + O * L2;
    const double A = c.alpha * (P + c.K2W0) / (P + c.K2) * Oi;
    const double X = c.EA1 * P / (1 + c.A01H01 * P) * G;
    f[0] = c.cgmax * Wi * P / (Wi + c.K1) - (c.A01 * P / (1 + c.A01H01 * P)) * G;
    if constexpr (NS == 1) {
        f[1] = X - c.M1 * G;
    } else {
        const double Pr = y[2 * L2 + i];
        f[1] = X - c.M1 * G - c.A12 * G / (1 + c.A12H12 * G) * Pr;
        f[2] = (c.EA2 * G / (1 + c.A12H12 * G)) * Pr - c.M2 * Pr;
    }
    f[W] = A - c.gmax * Wi * P / (Wi + c.K1) - c.rW * Wi + c.dW * (yW[up] + yW[dn] + yW[lf] + yW[rt] - 4 * Wi);
    f[O] = c.a - A + c.dO * (yO[up] + yO[dn] + yO[lf] + yO[rt] - 4 * Oi);
}

// Stage s (1..4) from the stage values src: K_s = f(src).  Stages 1-3 fold K_s
// into acc and write the next stage's values rho + h K_s to nxt; stage 4 writes
// rho + dt6 (acc + K4) over rho and counts the cells that turned negative or
// non-finite (the reference logs them, FKE:2511-2529, and carries on).
template <int NS>
__global__ __launch_bounds__(256) void fke_rk_stage_kernel(FkeRk c, int stage, double* rho, const double* src,
                                                           double* nxt, double* acc, unsigned long long* n_bad)
{
#pragma clang fp contract(off)
    constexpr int NF = NS + 3;
    const int L = c.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L2) return;
    const int y = (int)(i / L), x = (int)(i % L);
    const int up = (y == 0 ? L - 1 : y - 1) * L + x, dn = (y == L - 1 ? 0 : y + 1) * L + x;
    const int lf = y * L + (x == 0 ? L - 1 : x - 1), rt = y * L + (x == L - 1 ? 0 : x + 1);
    double f[NF];
    fke_rhs<NS>(c, src, L2, (int)i, up, dn, lf, rt, f);
    uint32_t bad = 0;
#pragma unroll
    for (int s = 0; s < NF; ++s) {
        const int64_t j = s * L2 + i;
        if (stage == 1) {
            acc[j] = f[s];
            nxt[j] = rho[j] + c.dt2 * f[s];
        } else if (stage < 4) {
            acc[j] = acc[j] + 2.0 * f[s];
            nxt[j] = rho[j] + (stage == 2 ? c.dt2 : c.dt) * f[s];
        } else {
            const double r = rho[j] + c.dt6 * (acc[j] + f[s]);
            rho[j] = r;
            bad += (r < 0 || !__builtin_isfinite(r)) ? 1u : 0u;
        }
    }
    if (stage == 4 && bad) atomicAdd(n_bad, (unsigned long long)bad);
}

bool fke_rk_args(int ns, int L, const tm_fke_rk_params* p, FkeRk* c)
{
#pragma clang fp contract(off)
    if ((ns != 1 && ns != 2) || L < 1 || !p || (int64_t)L * L > 0x7FFFFFFFll) return false;
    if (!(p->dx > 0.0) || !(p->dt_rk > 0.0) || !std::isfinite(p->dt)) return false;
    const double dx2 = p->dx * p->dx;
    c->ns = ns;
    c->L = L;
    c->a = p->a;
    c->cgmax = p->c * p->gmax;
    c->gmax = p->gmax;
    c->alpha = p->alpha;
    c->rW = p->rW;
    c->K1 = p->K[1];
    c->K2 = p->K[2];
    c->K2W0 = p->K[2] * p->W0;
    c->A01 = p->A[0][1];
    c->A01H01 = p->A[0][1] * p->H[0][1];
    c->EA1 = p->E[1] * p->A[0][1];
    c->M1 = p->M[1];
    c->A12 = p->A[1][2];
    c->A12H12 = p->A[1][2] * p->H[1][2];
    c->EA2 = p->E[2] * p->A[1][2];
    c->M2 = p->M[2];
    c->dW = p->D_W / dx2;
    c->dO = p->D_O / dx2;
    c->dt = p->dt;
    c->dt2 = p->dt_rk / 2.0;
    c->dt6 = p->dt_rk / 6.0;
    return true;
}

// The four stage launches on stream st; scratch holds 3 NF L^2 doubles.
void fke_rk_launch(const FkeRk& c, double* fields, double* scratch, unsigned long long* n_bad, hipStream_t st)
{
    const int64_t L2 = (int64_t)c.L * c.L;
    const int64_t nf = (int64_t)(c.ns + 3) * L2;
    double *sa = scratch, *sb = scratch + nf, *acc = scratch + 2 * nf;
    const dim3 grid((unsigned)((L2 + 255) / 256));
    const double* src[4] = {fields, sa, sb, sa};
    double* nxt[4] = {sa, sb, sa, nullptr};
    for (int s = 1; s <= 4; ++s) {
        if (c.ns == 1)
            hipLaunchKernelGGL(fke_rk_stage_kernel<1>, grid, dim3(256), 0, st, c, s, fields, src[s - 1], nxt[s - 1], acc,
                               n_bad);
        else
            hipLaunchKernelGGL(fke_rk_stage_kernel<2>, grid, dim3(256), 0, st, c, s, fields, src[s - 1], nxt[s - 1], acc,
                               n_bad);
    }
}

// ---------------------------------------------------------------------------
// Vegetation Dornic sub-step of the Rietkerk consumer model (FKE:4150-4211):
// the same transition, constants and Philox stream as the single-species
// step's P Dornic (cell_update MODE 1), with the vegetation's mortality M[0] as
// d (beta_0 = -M[0] - 4 D_0 / dx^2, FKE:3847-3849).  No flush: the reference
// flushes after the movement step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fke_veg_dornic_kernel(tm_rietkerk_params p, int L, uint32_t seed,
                                                             uint32_t inst, int64_t step, const double* P,
                                                             double* Pstar)
{
    const int64_t L2 = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L2) return;
    const tmk::RKConstD c = tmk::rk_constants(p, step);
    tmk::DevInst di = {};
    di.seed = seed;
    di.inst_id = inst;
    const int y = (int)(i / L), x = (int)(i % L);
    const int up = (y == 0 ? L - 1 : y - 1) * L + x, dn = (y == L - 1 ? 0 : y + 1) * L + x;
    const int lf = y * L + (x == 0 ? L - 1 : x - 1), rt = y * L + (x == L - 1 ? 0 : x + 1);
    int64_t n;
    bool bad;
    Pstar[i] = tmk::cell_update<double, 1>(c, di, (uint32_t)i, P[i], P[up], P[dn], P[lf], P[rt], n, bad);
}

bool veg_dornic_ok(int L, const tm_rietkerk_params* p)
{
    return L >= 1 && (int64_t)L * L <= 0x7FFFFFFFll && p && p->dx > 0.0 && p->sigma > 0.0 && p->dt > 0.0;
}

}  // namespace

extern "C" {

int tm_fke_move_constants(double D, double v, double sigma, double dt_analytic, double dt_adv, int mv_invariance,
                          double dx, int CI, tm_fke_move_species* out, double* stencil_D, int cap_D,
                          double* stencil_V, int cap_V)
{
    if (!out || !stencil_D || !stencil_V || !(dx > 0.0) || !(dt_analytic > 0.0) || !(sigma > 0.0)) return -1;
    if (mv_invariance == 1 && !(dt_adv > 0.0)) return -1;
    // set_global_FKE_params, FKE:222-250
    const double Dt_ay = dt_analytic;
    const double beta_vel = mv_invariance == 1 ? 3.0 / dt_adv : 3.0 / dt_analytic;
    const double exp_beta_Dt = std::exp(-beta_vel * dt_analytic);
    const double exp_2beta_Dt = std::exp(-2 * beta_vel * dt_analytic);
    const double sigma_v = std::sqrt(2 * beta_vel * 1 * std::fabs(v));
    const double sigma_D = std::sqrt(2 * D * dt_analytic);
    const double sigma_vx = (sigma_v / beta_vel) * std::sqrt((Dt_ay + (1 / (2 * beta_vel)) * (1 - exp_2beta_Dt)
                                                               - (2 / beta_vel) * (1 - exp_beta_Dt)));
    out->rD = (int)(CI * sigma_D / dx);
    out->rV = (int)(CI * sigma_vx / dx);
    out->mu_prefactor = (1 / dx) * (Dt_ay - ((1 - exp_beta_Dt) / beta_vel));
    if (out->rD < 0 || out->rV < 0 || out->rD + 1 > cap_D || out->rV + 1 > cap_V) return -1;
    gauss_stencil(sigma, out->rD, stencil_D);
    gauss_stencil(sigma, out->rV, stencil_V);
    return 0;
}

int tm_fke_movement(int ns, int L, const tm_fke_move_species* sp, const double* stencil_D, const double* stencil_V,
                    const double* rho, const double* gamma, const double* v_eff, double* out, void* stream)
{
    if (ns < 1 || ns > 3 || L < 1 || !sp || !stencil_D || !stencil_V || !rho || !gamma || !v_eff || !out)
        return TM_ERR_ARG;
    const int64_t L2 = (int64_t)L * L;
    if (L2 > 0x7FFFFFFFll) return TM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    DevBuf d_vmax;
    if (hipMalloc(&d_vmax.p, sizeof(double) * ns) != hipSuccess) return TM_ERR_NOMEM;
    hipLaunchKernelGGL(vmax_kernel, dim3(ns), dim3(256), 0, st, (const double2*)v_eff, L2, (double*)d_vmax.p);
    double vmax[3] = {0.0, 0.0, 0.0};
    if (hipMemcpyAsync(vmax, d_vmax.p, sizeof(double) * ns, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return TM_ERR_HIP;
    MoveArgs a;
    int nD, nV;
    const int rc = move_args(ns, L, sp, vmax, &a, &nD, &nV);
    if (rc != TM_OK) return rc;
    DevBuf d_coef, d_mu;
    if (hipMalloc(&d_coef.p, sizeof(double4) * (size_t)ns * L2) != hipSuccess ||
        hipMalloc(&d_mu.p, sizeof(int2) * (size_t)ns * L2) != hipSuccess)
        return TM_ERR_NOMEM;
    const dim3 grid((unsigned)((L2 + 255) / 256), (unsigned)ns);
    hipLaunchKernelGGL(fke_move_prep_kernel, grid, dim3(256), 0, st, a, rho, gamma, (const double2*)v_eff,
                       (double4*)d_coef.p, (int2*)d_mu.p);
    hipLaunchKernelGGL(fke_move_kernel, grid, dim3(256), 0, st, a, (const double4*)d_coef.p, (const int2*)d_mu.p,
                       stencil_D, stencil_V, out);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    // the scratch is released when this returns: wait for the kernels that use it
    return hipStreamSynchronize(st) == hipSuccess ? TM_OK : TM_ERR_HIP;
}

int tm_fke_movement_host(int device, int ns, int L, const tm_fke_move_species* sp, const double* stencil_D,
                         const double* stencil_V, const double* rho, const double* gamma, const double* v_eff,
                         double* out)
{
    if (ns < 1 || ns > 3 || L < 1 || !sp || !stencil_D || !stencil_V || !rho || !gamma || !v_eff || !out)
        return TM_ERR_ARG;
    const int64_t L2 = (int64_t)L * L;
    if (L2 > 0x7FFFFFFFll) return TM_ERR_ARG;
    double vmax[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < ns; ++s)
        for (int64_t i = 0; i < L2; ++i) {
            const double vx = std::fabs(v_eff[2 * (s * L2 + i)]), vy = std::fabs(v_eff[2 * (s * L2 + i) + 1]);
            if (!(vx <= 1e300) || !(vy <= 1e300)) return TM_ERR_ARG;   // NaN or inf velocity
            vmax[s] = std::max(vmax[s], std::max(vx, vy));
        }
    MoveArgs a;
    int nD, nV;
    int rc = move_args(ns, L, sp, vmax, &a, &nD, &nV);
    if (rc != TM_OK) return rc;
    if (hipSetDevice(device) != hipSuccess) return TM_ERR_NODEV;
    DevBuf d_rho, d_gam, d_v, d_out, d_sd, d_sv, d_coef, d_mu;
    const size_t fb = sizeof(double) * (size_t)ns * L2;
    if (hipMalloc(&d_rho.p, fb) != hipSuccess || hipMalloc(&d_gam.p, fb) != hipSuccess ||
        hipMalloc(&d_v.p, 2 * fb) != hipSuccess || hipMalloc(&d_out.p, fb) != hipSuccess ||
        hipMalloc(&d_sd.p, sizeof(double) * nD) != hipSuccess || hipMalloc(&d_sv.p, sizeof(double) * nV) != hipSuccess ||
        hipMalloc(&d_coef.p, 4 * fb) != hipSuccess || hipMalloc(&d_mu.p, sizeof(int2) * (size_t)ns * L2) != hipSuccess)
        return TM_ERR_NOMEM;
    if (hipMemcpy(d_rho.p, rho, fb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_gam.p, gamma, fb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_v.p, v_eff, 2 * fb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_sd.p, stencil_D, sizeof(double) * nD, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_sv.p, stencil_V, sizeof(double) * nV, hipMemcpyHostToDevice) != hipSuccess)
        return TM_ERR_HIP;
    const dim3 grid((unsigned)((L2 + 255) / 256), (unsigned)ns);
    hipLaunchKernelGGL(fke_move_prep_kernel, grid, dim3(256), 0, nullptr, a, (const double*)d_rho.p,
                       (const double*)d_gam.p, (const double2*)d_v.p, (double4*)d_coef.p, (int2*)d_mu.p);
    hipLaunchKernelGGL(fke_move_kernel, grid, dim3(256), 0, nullptr, a, (const double4*)d_coef.p,
                       (const int2*)d_mu.p, (const double*)d_sd.p, (const double*)d_sv.p, (double*)d_out.p);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    if (hipMemcpy(out, d_out.p, fb, hipMemcpyDeviceToHost) != hipSuccess) return TM_ERR_HIP;
    return TM_OK;
}

int tm_fke_perception_host(int device, int ns, int L, const double* pR, double dx, const double* v, uint32_t seed,
                           uint32_t inst, int64_t step, const double* rho, double* gamma, double* v_eff)
{
    if ((ns != 2 && ns != 3) || L < 1 || !pR || !v || !rho || !gamma || !v_eff || !(dx > 0.0)) return TM_ERR_ARG;
    const int64_t N = (int64_t)L * L;
    if (N > 0x7FFFFFFFll) return TM_ERR_ARG;
    double rmax = pR[0];
    for (int s = 1; s < ns; ++s) rmax = pR[s] > rmax ? pR[s] : rmax;
    rmax = rmax / dx;
    if (!(rmax > 0.0) || !(rmax < TM_DPC_MAX_RADIUS + 1.0) || (int)rmax >= L) return TM_ERR_ARG;
    const int R = (int)rmax;
    std::vector<int32_t> off(2 * (size_t)(2 * R + 1) * (2 * R + 1));
    const int disk_len = tm_perception_offsets(R, off.data(), (int)off.size() / 2);
    if (disk_len < 1) return TM_ERR_ARG;
    PerceptArgs a{};
    a.ns = ns;
    a.L = L;
    a.seed = seed;
    a.inst = inst;
    a.step = step;
    double occ = 0.0;
    for (int64_t i = 0; i < N; ++i) occ += rho[i] > 0 ? 1.0 : 0.0;
    double nfac = 1 - occ / (double)N;                 // FKE:4109-4112
    if (nfac < 0.35) nfac = 0.35;
    int new_len = (int)(nfac * nfac * (double)disk_len);
    if (ns == 3 && new_len < 1) new_len = 1;           // FKE:3386
    for (int s = 0; s < ns; ++s) {
        a.v[s] = v[s];
        a.mean[s] = fixed_sum(rho + s * N, N) / (double)N;
        a.inv[s] = a.mean[s] >= 1.0e-12 ? 1 / a.mean[s] : 1 / 1.0e-12;
        a.fr[s] = pR[s] / (dx * rmax);
        int e = (int)(a.fr[s] * a.fr[s] * (double)new_len);
        if (ns == 3 && e < 1 && a.fr[s] > 0) e = 1;     // FKE:3415-3416
        a.eff[s] = e;
    }
    if (hipSetDevice(device) != hipSuccess) return TM_ERR_NODEV;
    DevBuf d_rho, d_gam, d_v, d_disk;
    const size_t fb = sizeof(double) * (size_t)N;
    if (hipMalloc(&d_rho.p, fb * ns) != hipSuccess || hipMalloc(&d_gam.p, fb * (ns - 1)) != hipSuccess ||
        hipMalloc(&d_v.p, 2 * fb * ns) != hipSuccess || hipMalloc(&d_disk.p, sizeof(int2) * disk_len) != hipSuccess)
        return TM_ERR_NOMEM;
    if (hipMemcpy(d_rho.p, rho, fb * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_gam.p, gamma, fb * (ns - 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_v.p, v_eff, 2 * fb * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_disk.p, off.data(), sizeof(int2) * disk_len, hipMemcpyHostToDevice) != hipSuccess)
        return TM_ERR_HIP;
    hipLaunchKernelGGL(fke_perception_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, nullptr, a,
                       (const int2*)d_disk.p, (const double*)d_rho.p, (double*)d_gam.p, (double2*)d_v.p);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    if (hipMemcpy(gamma, d_gam.p, fb * (ns - 1), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(v_eff, d_v.p, 2 * fb * ns, hipMemcpyDeviceToHost) != hipSuccess)
        return TM_ERR_HIP;
    return TM_OK;
}

int tm_fke_rk4(int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad, void* stream)
{
    FkeRk c;
    if (!fields || !fke_rk_args(ns, L, p, &c)) return TM_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    const size_t nf = (size_t)(ns + 3) * L * L;
    DevBuf d_s, d_bad;
    if (hipMalloc(&d_s.p, sizeof(double) * 3 * nf) != hipSuccess ||
        hipMalloc(&d_bad.p, sizeof(unsigned long long)) != hipSuccess)
        return TM_ERR_NOMEM;
    unsigned long long bad = 0;
    if (hipMemcpyAsync(d_bad.p, &bad, sizeof(bad), hipMemcpyHostToDevice, st) != hipSuccess) return TM_ERR_HIP;
    fke_rk_launch(c, fields, (double*)d_s.p, (unsigned long long*)d_bad.p, st);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    // the scratch is freed on return: wait for the stages that use it
    if (hipMemcpyAsync(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return TM_ERR_HIP;
    if (n_bad) *n_bad = (int64_t)bad;
    return TM_OK;
}

// ---------------------------------------------------------------------------
// tm_fke_rk4_noalloc: tm_fke_rk4 with its stage buffers and its negative-value
// counter in `scratch`, a file-scope device array, instead of per-call device
// allocations.  Each GPU the library runs on holds its own copy until the
// process ends.  In host code the name is only a handle: the array's device
// address on a device comes from hipGetSymbolAddress with that device current,
// taken on the first call on the device and kept per device.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU race; the
// counter is cleared by a kernel at the start of each call.
// ---------------------------------------------------------------------------
constexpr int kRk4NoAllocL = 256;            // largest lattice side (ns <= 2 consumers)
constexpr size_t kRk4ScratchDoubles = 3 * 5 * (size_t)kRk4NoAllocL * kRk4NoAllocL;   // 3 (ns + 3) L^2
static __device__ __align__(256) double scratch[kRk4ScratchDoubles + 1];             // + the counter

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (nullptr on failure).
static double* rk4_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return g_scratch_ptr[dev];
}

__global__ void clear_counter_kernel(unsigned long long* n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *n = 0ull;
}

int tm_fke_rk4_noalloc(int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad, void* stream)
{
    FkeRk c;
    if (!fields || !fke_rk_args(ns, L, p, &c)) return TM_ERR_ARG;
    if (L > kRk4NoAllocL) return TM_ERR_TOO_LARGE;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return TM_ERR_NODEV;
    double* s = rk4_scratch(dev);
    if (!s) return TM_ERR_HIP;
    unsigned long long* d_bad = (unsigned long long*)(s + kRk4ScratchDoubles);
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(64), 0, st, d_bad);
    fke_rk_launch(c, fields, s, d_bad, st);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    unsigned long long bad = 0;
    if (hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return TM_ERR_HIP;
    if (n_bad) *n_bad = (int64_t)bad;
    return TM_OK;
}

int tm_fke_rk4_host(int device, int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad)
{
    FkeRk c;
    if (!fields || !fke_rk_args(ns, L, p, &c)) return TM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return TM_ERR_NODEV;
    const size_t nf = (size_t)(ns + 3) * L * L;
    DevBuf d_f, d_s, d_bad;
    if (hipMalloc(&d_f.p, sizeof(double) * nf) != hipSuccess ||
        hipMalloc(&d_s.p, sizeof(double) * 3 * nf) != hipSuccess ||
        hipMalloc(&d_bad.p, sizeof(unsigned long long)) != hipSuccess)
        return TM_ERR_NOMEM;
    unsigned long long bad = 0;
    if (hipMemcpy(d_f.p, fields, sizeof(double) * nf, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_bad.p, &bad, sizeof(bad), hipMemcpyHostToDevice) != hipSuccess)
        return TM_ERR_HIP;
    fke_rk_launch(c, (double*)d_f.p, (double*)d_s.p, (unsigned long long*)d_bad.p, nullptr);
    if (hipGetLastError() != hipSuccess) return TM_ERR_HIP;
    if (hipMemcpy(fields, d_f.p, sizeof(double) * nf, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost) != hipSuccess)
        return TM_ERR_HIP;
    if (n_bad) *n_bad = (int64_t)bad;
    return TM_OK;
}

int tm_fke_veg_dornic(int L, const tm_rietkerk_params* p, uint32_t seed, uint32_t inst, int64_t step,
                      const double* P, double* Pstar, void* stream)
{
    if (!P || !Pstar || P == Pstar || !veg_dornic_ok(L, p)) return TM_ERR_ARG;
    const int64_t L2 = (int64_t)L * L;
    hipLaunchKernelGGL(fke_veg_dornic_kernel, dim3((unsigned)((L2 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *p, L, seed, inst, step, P, Pstar);
    return hipGetLastError() == hipSuccess ? TM_OK : TM_ERR_HIP;
}

int tm_fke_veg_dornic_host(int device, int L, const tm_rietkerk_params* p, uint32_t seed, uint32_t inst,
                           int64_t step, const double* P, double* Pstar)
{
    if (!P || !Pstar || !veg_dornic_ok(L, p)) return TM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return TM_ERR_NODEV;
    const size_t fb = sizeof(double) * (size_t)L * L;
    DevBuf d_p, d_s;
    if (hipMalloc(&d_p.p, fb) != hipSuccess || hipMalloc(&d_s.p, fb) != hipSuccess) return TM_ERR_NOMEM;
    if (hipMemcpy(d_p.p, P, fb, hipMemcpyHostToDevice) != hipSuccess) return TM_ERR_HIP;
    int rc = tm_fke_veg_dornic(L, p, seed, inst, step, (const double*)d_p.p, (double*)d_s.p, nullptr);
    if (rc != TM_OK) return rc;
    return hipMemcpy(Pstar, d_s.p, fb, hipMemcpyDeviceToHost) == hipSuccess ? TM_OK : TM_ERR_HIP;
}

}  // extern "C"
