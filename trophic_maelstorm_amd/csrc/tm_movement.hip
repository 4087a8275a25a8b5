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
// The reference scatters with atomicAdd.  Here the four terms are factored by
// rows and columns instead: a first pass forms each source's term coefficients
// and advected mean, a second sums, per cell, the coefficients of the sources
// whose advected mean lands on it (within the largest shift), and two 1-D
// passes apply the row and column stencils (move_disp_at, move_row_at,
// move_col_at).  No atomics, every sum in FP64 in a fixed order, and the
// result is deterministic; the sums differ from the reference's sequential
// scatter only in their order.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "tm/tm.h"
#include "tm_device.h"
#include "tm_kernels.h"
#include "tm_launch.h"

namespace {

constexpr double kMoveEps = 1.0e-10;   // epsilon, rietkerk_bjork_basic.h:119

struct MoveSpecies {
    int rD, rV, W, S, offD, offV;   // S: the largest shift of an advected mean, plus one
    double pref;
};
struct MoveArgs {
    MoveSpecies sp[3];
    int L;
};

__device__ __forceinline__ int wrapi(int v, int L)
{
    v %= L;
    return v < 0 ? v + L : v;
}

// Per source cell: its three term coefficients (gamma^2 rho, (gamma - gamma^2) rho,
// (1 - gamma)^2 rho, formed as the reference forms them) and its advected mean
// cell, once per call; mu.x = -1 marks a source below the 1e-10 floor.
__device__ __forceinline__ void move_prep_at(double pref, int L, int64_t i, double r, double g, double2 v,
                                             double4& coef, int2& mu)
{
#pragma clang fp contract(off)
    if (r < kMoveEps) {
        mu = make_int2(-1, -1);
        return;
    }
    const int ci = (int)(i / L), cj = (int)(i % L);
    const int mux = (int)(__builtin_round(ci + v.x * (pref)) + L) % L;
    const int muy = (int)(__builtin_round(cj + v.y * (pref)) + L) % L;
    const double g2 = g * g;
    coef = make_double4((g2) * r, (g - g2) * r, (1 - g) * (1 - g) * r, 0.0);
    mu = make_int2(mux, muy);
}

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
    move_prep_at(a.sp[s].pref, L, i, rho[s * L2 + i], gamma[s * L2 + i], veff[s * L2 + i], coef[s * L2 + i],
                 mu[s * L2 + i]);
}

// v in [-L, 2L) -> v mod L
__device__ __forceinline__ int wrap1(int v, int L)
{
    v += v < 0 ? L : 0;
    return v >= L ? v - L : v;
}

// The separable form, in three passes over one species' lattice.  With each
// source's terms (c.x, c.y, c.z) and advected mean (mx, my) from move_prep_at:
//   1. move_disp_at: cx(p, q) = c.x of source (p, q) (0 if empty), and over the
//      (2S + 1)^2 sources around (p, q), rows then columns ascending,
//      Y1 = sum of c.y of the sources of row p with my = q,
//      Y2 = sum of c.y of the sources of column q with mx = p,
//      Z  = sum of c.z of the sources with (mx, my) = (p, q);
//   2. move_row_at, along the row: P = sum_b SD|b| cx(p, q - b) + sum_b SV|b| Y1(p, q - b),
//      Q = sum_b SD|b| Y2(p, q - b) + sum_b SV|b| Z(p, q - b);
//   3. move_col_at, along the column: out(x, y) = sum_a SD|a| P(x - a, y) + sum_a SV|a| Q(x - a, y);
// taps from -r to r, indices mod L.  Summed in FP64 in exactly this order, which
// is the oracle's tmo_fke_movement_sep.  The four terms of the scatter (D x D,
// D x V, V x D, V x V) factor by rows and columns: per output cell 2(2 rD + 1) +
// 2(2 rV + 1) taps and (2S + 1)^2 shift tests instead of a (2W + 1)^2 window.
__device__ __forceinline__ double4 move_disp_at(const double4* __restrict__ cs, const int2* __restrict__ ms, int L,
                                                int S, int p, int q)
{
#pragma clang fp contract(off)
    const int64_t o = (int64_t)p * L + q;
    double c0 = 0.0, a1 = 0.0, a2 = 0.0, z = 0.0;
    if (ms[o].x >= 0) c0 = cs[o].x;
    for (int di = -S; di <= S; ++di) {
        const int ci = wrap1(p + di, L);
        for (int dj = -S; dj <= S; ++dj) {
            const int64_t i = (int64_t)ci * L + wrap1(q + dj, L);
            const int2 m = ms[i];
            const bool hx = m.x == p, hy = m.y == q;   // an empty source has m.x = -1
            if (!(hx || hy)) continue;
            if (di == 0 && hy) a1 += cs[i].y;
            if (dj == 0 && hx) a2 += cs[i].y;
            if (hx && hy) z += cs[i].z;
        }
    }
    return make_double4(c0, a1, a2, z);
}

// mv: the four planes cx, Y1, Y2, Z of one lattice, N = L^2 apart, so that
// neighbouring threads read neighbouring words.
__device__ __forceinline__ double2 move_row_at(const double* __restrict__ mv, int64_t N,
                                               const double* __restrict__ SD, const double* __restrict__ SV,
                                               int rD, int rV, int L, int p, int q)
{
#pragma clang fp contract(off)
    const double* row = mv + (int64_t)p * L;
    double P = 0.0, Q = 0.0;
    for (int b = -rD; b <= rD; ++b) {
        const int j = wrap1(q - b, L);
        const double w = SD[b < 0 ? -b : b];
        P += w * row[j];
        Q += w * row[2 * N + j];
    }
    for (int b = -rV; b <= rV; ++b) {
        const int j = wrap1(q - b, L);
        const double w = SV[b < 0 ? -b : b];
        P += w * row[N + j];
        Q += w * row[3 * N + j];
    }
    return make_double2(P, Q);
}

// pq: the planes P and Q of one lattice, N apart.
__device__ __forceinline__ double move_col_at(const double* __restrict__ pq, int64_t N,
                                              const double* __restrict__ SD, const double* __restrict__ SV,
                                              int rD, int rV, int L, int x, int y)
{
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int a = -rD; a <= rD; ++a) acc += SD[a < 0 ? -a : a] * pq[(int64_t)wrap1(x - a, L) * L + y];
    for (int a = -rV; a <= rV; ++a) acc += SV[a < 0 ? -a : a] * pq[N + (int64_t)wrap1(x - a, L) * L + y];
    return acc;
}

__device__ __forceinline__ void put4(double* __restrict__ planes, int64_t N, int64_t o, double4 v)
{
    planes[o] = v.x;
    planes[N + o] = v.y;
    planes[2 * N + o] = v.z;
    planes[3 * N + o] = v.w;
}

// The three passes of the host-array operator: one thread per (cell, species).
// mv: 4 planes per species, pq: 2 planes per species.
__global__ __launch_bounds__(256) void fke_move_disp_kernel(MoveArgs a, const double4* __restrict__ coef,
                                                            const int2* __restrict__ mu, double* __restrict__ mv)
{
    const int s = blockIdx.y;
    const int L = a.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= L2) return;
    put4(mv + 4 * s * L2, L2, o, move_disp_at(coef + s * L2, mu + s * L2, L, a.sp[s].S, (int)(o / L), (int)(o % L)));
}

__global__ __launch_bounds__(256) void fke_move_row_kernel(MoveArgs a, const double* __restrict__ mv,
                                                           const double* __restrict__ SDall,
                                                           const double* __restrict__ SVall, double* __restrict__ pq)
{
    const int s = blockIdx.y;
    const int L = a.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= L2) return;
    const MoveSpecies sp = a.sp[s];
    const double2 r = move_row_at(mv + 4 * s * L2, L2, SDall + sp.offD, SVall + sp.offV, sp.rD, sp.rV, L,
                                  (int)(o / L), (int)(o % L));
    pq[2 * s * L2 + o] = r.x;
    pq[2 * s * L2 + L2 + o] = r.y;
}

__global__ __launch_bounds__(256) void fke_move_col_kernel(MoveArgs a, const double* __restrict__ pq,
                                                           const double* __restrict__ SDall,
                                                           const double* __restrict__ SVall, double* __restrict__ out)
{
    const int s = blockIdx.y;
    const int L = a.L;
    const int64_t L2 = (int64_t)L * L;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= L2) return;
    const MoveSpecies sp = a.sp[s];
    out[s * L2 + o] = move_col_at(pq + 2 * s * L2, L2, SDall + sp.offD, SVall + sp.offV, sp.rD, sp.rV, L,
                                  (int)(o / L), (int)(o % L));
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
        m.S = S;
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

// Velocity of species s at a cell in the three-species perception step: v (cos,
// sin)(2 pi u), u = word 0 of Philox(seed, inst; cell, step, block (s << 24) | 0x800000).
__device__ __forceinline__ double2 fke_velocity(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step, int s,
                                                double v)
{
#pragma clang fp contract(off)
    tmk::Stream st(seed, inst, cell, step);
    st.blk = ((uint32_t)s << 24) | 0x800000u;
    double sn, cs;
    tmk::dsincos2pi(st.u64(), sn, cs);
    return make_double2(v * cs, v * sn);
}

// Perception gamma of consumer s (1 .. ns-1) at cell (p, q) from its previous
// value g: the old densities rho [ns][N] summed over the first eff[s] perception
// offsets (weighed down by the predators for the grazer of the three-species
// form), scaled by the prey's inverse mean and clamped to [0, 1]; 1 when the
// consumer is extinct, 0 when its prey is; unchanged when it perceives nothing.
__device__ __forceinline__ double fke_gamma_at(int ns, int s, const double* mean, const double* inv, const double* fr,
                                               const int* eff, const int2* __restrict__ disk,
                                               const double* __restrict__ rho, int64_t N, int L, int p, int q,
                                               double g)
{
#pragma clang fp contract(off)
    if (fr[s] == 0.0) return g;
    if (mean[s] < 1.0e-12) return 1.0;
    if (mean[s - 1] < 1.0e-12) return 0.0;
    const double* src = rho + (s - 1) * N;
    const bool weigh = ns == 3 && s == 1 && !(mean[2] < 1.0e-12);
    double acc = g;
    for (int k = 0; k < eff[s]; ++k) {
        const int2 o = disk[k];
        const int64_t j = (int64_t)wrapi(p + o.x, L) * L + wrapi(q + o.y, L);
        if (weigh) acc += src[j] * (1 - rho[2 * N + j] * inv[2]);
        else acc += src[j];
    }
    if (ns == 2 || s == 1) acc = (acc * inv[s - 1]) / (double)eff[s];
    else acc = (acc / (double)eff[s]) * inv[s - 1];
    if (acc > 1) acc = 1.0;
    else if (acc < 0) acc = 0.0;
    return acc;
}

__global__ __launch_bounds__(256) void fke_perception_kernel(PerceptArgs a, const int2* __restrict__ disk,
                                                             const double* __restrict__ rho,
                                                             double* __restrict__ gamma, double2* __restrict__ veff)
{
    const int L = a.L;
    const int64_t N = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int p = (int)(i / L), q = (int)(i % L);
    if (a.ns == 3)
        for (int s = 0; s < 3; ++s) veff[s * N + i] = fke_velocity(a.seed, a.inst, (uint32_t)i, a.step, s, a.v[s]);
    for (int s = 1; s < a.ns; ++s) {
        double* g = gamma + (s - 1) * N + i;
        *g = fke_gamma_at(a.ns, s, a.mean, a.inv, a.fr, a.eff, disk, rho, N, L, p, q, *g);
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
__device__ __forceinline__ uint32_t fke_stage_at(const FkeRk& c, int stage, double* rho, const double* src,
                                                 double* nxt, double* acc, int64_t L2, int i)
{
#pragma clang fp contract(off)
    constexpr int NF = NS + 3;
    const int L = c.L;
    const int y = i / L, x = i % L;
    const int up = (y == 0 ? L - 1 : y - 1) * L + x, dn = (y == L - 1 ? 0 : y + 1) * L + x;
    const int lf = y * L + (x == 0 ? L - 1 : x - 1), rt = y * L + (x == L - 1 ? 0 : x + 1);
    double f[NF];
    fke_rhs<NS>(c, src, L2, i, up, dn, lf, rt, f);
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
    return bad;
}

template <int NS>
__global__ __launch_bounds__(256) void fke_rk_stage_kernel(FkeRk c, int stage, double* rho, const double* src,
                                                           double* nxt, double* acc, unsigned long long* n_bad)
{
    const int64_t L2 = (int64_t)c.L * c.L;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L2) return;
    const uint32_t bad = fke_stage_at<NS>(c, stage, rho, src, nxt, acc, L2, (int)i);
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

// ---------------------------------------------------------------------------
// The Rietkerk consumer model as one step of a batch (TM_MODEL_FKE2: vegetation
// P and grazer G; TM_MODEL_FKE3: and predator Pr), FP64, the loop body of
// rietkerk_bjork_FKE.cpp:4100-4462.  Fields per instance, species-major
// [NF][L*L] at NF cell_off: P, G, (Pr,) W, O, then gamma_1 (, gamma_2);
// NF = 2 SP + 1 for SP = 2 or 3 biota species.  One step, every launch over the
// whole batch:
//   fke_chunk / fke_state  spatial means of the old densities and the cover
//                          factor (fixed-order sums, FKE:4100-4112), per instance
//   fke_percept            gamma of each consumer from the old densities
//                          (FKE:4114-4118); velocities are drawn where used
//   fke_veg                the vegetation's Dornic step P* (FKE:4150-4211), the
//                          other densities carried into the output buffer
//   fke_stage x 4          RK4 of the trophic and water terms (FKE:4233-4238),
//                          one grid-wide pass per stage (W and O diffuse)
//   fke_mprep, fke_move_pass x 3  the consumers' Fokker-Planck movement
//                          (FKE:4276-4309), in separable form
//   fke_final              checks and flush of the densities, rho(t) partials of
//                          every field (FKE:4352-4462)
// Every operation is the one the per-operator entry points (tm_fke_perception,
// tm_fke_veg_dornic, tm_fke_rk4, tm_fke_movement) perform, in their order.
// ---------------------------------------------------------------------------
namespace tmk {
namespace {

struct FkeModelConst {     // per parameter point
    FkeRk rk;              // L is set per instance
    MoveSpecies ms[2];     // consumers 1 and 2
    double v[3];
    double eps;
};

struct FkeWork {
    const FkeModelConst* mc;   // [n_params]
    const DPCInstInfo* info;   // [n_inst]: chunks, perception table, fr
    DPCStepState* st;          // [n_inst]: means of the step
    double* chunk;             // [nchunks][8]
    const int2* chunks;        // [nchunks]: (instance, chunk within it)
    const int2* disk;          // perception offsets of every table
    const int2* blocks;        // [nblocks]: (instance, first cell) of 256 cells
    const double* SD;          // movement stencils, offsets in MoveSpecies
    const double* SV;
    double* sa;                // RK4 stage values, (SP + 2) fields at (SP + 2) cell_off
    double* sb;
    double* acc;
    double4* coef;             // movement terms, (SP - 1) fields at (SP - 1) cell_off
    int2* mu;
    double* mv;                // separable movement: planes cx, Y1, Y2, Z of each lattice (4 per
    double* pq;                // coef entry, at 4 x its offset), then planes P, Q (2, at 2 x)
    double* mstat;             // [n_inst][8] means of the movement statistics
};

constexpr int kFkeChunkCells = kDpcChunkCells;

// Sums over the chunks of NV per-cell values in the consumer models' fixed order
// (lanes of 256 strided over the chunk, then a pairwise fold).
template <int NV>
__device__ __forceinline__ void chunk_fold(double (&v)[NV], double (&red)[NV][256], double* dst)
{
#pragma clang fp contract(off)
    const int t = threadIdx.x;
#pragma unroll
    for (int f = 0; f < NV; ++f) red[f][t] = v[f];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
#pragma unroll
            for (int f = 0; f < NV; ++f) red[f][t] += red[f][t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int f = 0; f < NV; ++f) dst[f] = red[f][0];
    }
}

template <int SP>
__global__ __launch_bounds__(256) void fke_chunk_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    constexpr int NF = 2 * SP + 1;
    __shared__ double red[SP + 1][256];
    const FkeWork* w = (const FkeWork*)a.work;
    const int2 ch = w->chunks[blockIdx.x];
    const DevInst di = a.inst[ch.x];
    const int64_t N = (int64_t)di.L * di.L;
    const double* f0 = (const double*)a.in + (int64_t)NF * di.cell_off;
    const int64_t base = (int64_t)ch.y * kFkeChunkCells;
    double s[SP + 1];
#pragma unroll
    for (int f = 0; f <= SP; ++f) s[f] = 0.0;
    for (int64_t i = base + threadIdx.x; i < base + kFkeChunkCells && i < N; i += 256) {
#pragma unroll
        for (int f = 0; f < SP; ++f) s[f] += f0[f * N + i];
        s[SP] += f0[i] > 0 ? 1.0 : 0.0;
    }
    chunk_fold<SP + 1>(s, red, w->chunk + (int64_t)blockIdx.x * 8);
}

// The per-instance numbers of the perception step, as tm_fke_perception_host
// forms them (FKE:4100-4112, :2534-2606, :3371-3482).
template <int SP>
__global__ __launch_bounds__(64) void fke_state_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.n_inst) return;
    const FkeWork* w = (const FkeWork*)a.work;
    const DPCInstInfo in = w->info[i];
    const DevInst di = a.inst[i];
    const double N = (double)((int64_t)di.L * di.L);
    double tot[SP], occ = 0.0;
#pragma unroll
    for (int f = 0; f < SP; ++f) tot[f] = 0.0;
    for (int c = 0; c < in.nchunks; ++c) {
        const double* p = w->chunk + (int64_t)(in.chunk0 + c) * 8;
#pragma unroll
        for (int f = 0; f < SP; ++f) tot[f] += p[f];
        occ += p[SP];
    }
    DPCStepState st{};
    double nfac = 1 - occ / N;                       // FKE:4109-4112
    if (nfac < 0.35) nfac = 0.35;
    int new_len = (int)(nfac * nfac * (double)in.disk_len);
    if (SP == 3 && new_len < 1) new_len = 1;         // FKE:3386
    for (int f = 0; f < SP; ++f) {
        st.mean[f] = tot[f] / N;
        st.inv[f] = st.mean[f] >= 1.0e-12 ? 1 / st.mean[f] : 1 / 1.0e-12;
        int e = (int)(in.fr[f] * in.fr[f] * (double)new_len);
        if (SP == 3 && e < 1 && in.fr[f] > 0) e = 1;     // FKE:3415-3416
        st.eff[f] = e;
    }
    w->st[i] = st;
}

// The cell of the thread in the block table; false past its instance's end.
__device__ __forceinline__ bool fke_cell(const StepArgs& a, const FkeWork* w, int& inst, int& cell, DevInst& di)
{
    const int2 b = w->blocks[blockIdx.x];
    inst = b.x;
    cell = b.y + (int)threadIdx.x;
    di = a.inst[inst];
    return (int64_t)cell < (int64_t)di.L * di.L;
}

template <int SP>
__global__ __launch_bounds__(256) void fke_percept_kernel(StepArgs a)
{
    constexpr int NF = 2 * SP + 1;
    const FkeWork* w = (const FkeWork*)a.work;
    int inst, cell;
    DevInst di;
    if (!fke_cell(a, w, inst, cell, di)) return;
    const int L = di.L;
    const int64_t N = (int64_t)L * L;
    const double* fin = (const double*)a.in + (int64_t)NF * di.cell_off;
    double* fout = (double*)a.out + (int64_t)NF * di.cell_off;
    const DPCStepState st = w->st[inst];
    const DPCInstInfo info = w->info[inst];
    const int p = cell / L, q = cell % L;
#pragma unroll
    for (int s = 1; s < SP; ++s) {
        const int64_t j = (int64_t)(SP + 1 + s) * N + cell;
        fout[j] = fke_gamma_at(SP, s, st.mean, st.inv, info.fr, st.eff, w->disk + info.disk_off, fin, N, L, p, q,
                               fin[j]);
    }
}

template <int SP>
__global__ __launch_bounds__(256) void fke_veg_kernel(StepArgs a)
{
    constexpr int NF = 2 * SP + 1;
    const FkeWork* w = (const FkeWork*)a.work;
    int inst, cell;
    DevInst di;
    if (!fke_cell(a, w, inst, cell, di)) return;
    const int L = di.L;
    const int64_t N = (int64_t)L * L;
    const double* fin = (const double*)a.in + (int64_t)NF * di.cell_off;
    double* fout = (double*)a.out + (int64_t)NF * di.cell_off;
    const RKConstD c = ((const RKConstD*)a.consts)[(int64_t)a.k * a.n_inst + inst];
    const int y = cell / L, x = cell % L;
    const int up = (y == 0 ? L - 1 : y - 1) * L + x, dn = (y == L - 1 ? 0 : y + 1) * L + x;
    const int lf = y * L + (x == 0 ? L - 1 : x - 1), rt = y * L + (x == L - 1 ? 0 : x + 1);
    int64_t n;
    bool bad;
    fout[cell] = cell_update<double, 1>(c, di, (uint32_t)cell, fin[cell], fin[up], fin[dn], fin[lf], fin[rt], n, bad);
#pragma unroll
    for (int f = 1; f < SP + 2; ++f) fout[f * N + cell] = fin[f * N + cell];
}

template <int SP>
__global__ __launch_bounds__(256) void fke_stage_kernel(StepArgs a, int stage)
{
    constexpr int NF = 2 * SP + 1, ND = SP + 2;
    const FkeWork* w = (const FkeWork*)a.work;
    int inst, cell;
    DevInst di;
    if (!fke_cell(a, w, inst, cell, di)) return;
    FkeRk c = w->mc[di.pidx].rk;
    c.L = di.L;
    const int64_t N = (int64_t)di.L * di.L;
    double* rho = (double*)a.out + (int64_t)NF * di.cell_off;
    double* sa = w->sa + (int64_t)ND * di.cell_off;
    double* sb = w->sb + (int64_t)ND * di.cell_off;
    double* acc = w->acc + (int64_t)ND * di.cell_off;
    const double* src = stage == 1 ? rho : (stage == 3 ? sb : sa);
    double* nxt = stage == 2 ? sb : sa;   // unused by stage 4
    (void)fke_stage_at<SP - 1>(c, stage, rho, src, nxt, acc, N, cell);
}

template <int SP>
__global__ __launch_bounds__(256) void fke_mprep_kernel(StepArgs a)
{
    constexpr int NF = 2 * SP + 1;
    const FkeWork* w = (const FkeWork*)a.work;
    int inst, cell;
    DevInst di;
    if (!fke_cell(a, w, inst, cell, di)) return;
    const int s = blockIdx.y + 1;
    const int L = di.L;
    const int64_t N = (int64_t)L * L;
    const double* f = (const double*)a.out + (int64_t)NF * di.cell_off;
    const FkeModelConst& mc = w->mc[di.pidx];
    double2 v = make_double2(0.0, 0.0);   // the two-species step draws no velocity
    if (SP == 3) {
        const int64_t step = ((const RKConstD*)a.consts)[(int64_t)a.k * a.n_inst + inst].step;
        v = fke_velocity(di.seed, di.inst_id, (uint32_t)cell, step, s, mc.v[s]);
    }
    const int64_t j = (int64_t)(SP - 1) * di.cell_off + (s - 1) * N + cell;
    move_prep_at(mc.ms[s - 1].pref, L, cell, f[s * N + cell], f[(SP + 1 + s) * N + cell], v, w->coef[j], w->mu[j]);
}

// The separable movement's passes (move_disp_at, move_row_at, move_col_at) on
// each consumer's lattice; the last writes the moved density.
template <int SP, int PASS>
__global__ __launch_bounds__(256) void fke_move_pass_kernel(StepArgs a)
{
    constexpr int NF = 2 * SP + 1;
    const FkeWork* w = (const FkeWork*)a.work;
    int inst, cell;
    DevInst di;
    if (!fke_cell(a, w, inst, cell, di)) return;
    const int s = blockIdx.y + 1;
    const int L = di.L;
    const int64_t N = (int64_t)L * L;
    const MoveSpecies sp = w->mc[di.pidx].ms[s - 1];
    const int64_t base = (int64_t)(SP - 1) * di.cell_off + (s - 1) * N;
    const int p = cell / L, q = cell % L;
    if (PASS == 0) {
        put4(w->mv + 4 * base, N, cell, move_disp_at(w->coef + base, w->mu + base, L, sp.S, p, q));
    } else if (PASS == 1) {
        const double2 r = move_row_at(w->mv + 4 * base, N, w->SD + sp.offD, w->SV + sp.offV, sp.rD, sp.rV, L, p, q);
        w->pq[2 * base + cell] = r.x;
        w->pq[2 * base + N + cell] = r.y;
    } else {
        double* f = (double*)a.out + (int64_t)NF * di.cell_off;
        f[s * N + cell] = move_col_at(w->pq + 2 * base, N, w->SD + sp.offD, w->SV + sp.offV, sp.rD, sp.rV, L, p, q);
    }
}

// Checks and flush of the densities (the perception fields are left as they
// are), then the (sum, occupied count) partials of every field of the tile.
template <int SP, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void fke_final_kernel(StepArgs a)
{
#pragma clang fp contract(off)
    constexpr int NF = 2 * SP + 1, ND = SP + 2;
    __shared__ double s_red[2 * NF * (NT / 64)];
    const FkeWork* w = (const FkeWork*)a.work;
    const int tile = blockIdx.x;
    if (tile >= a.ntiles) return;
    const int64_t counter = a.state->counter + a.k;
    const int64_t slot = counter & (int64_t)(a.nslots - 1);
    const DevTile tl = a.tiles[tile];
    const DevInst di = a.inst[tl.inst];
    const double eps = w->mc[di.pidx].eps;
    const int L = tl.L;
    const int64_t N = (int64_t)L * L;
    double* f = (double*)a.out + (int64_t)NF * tl.cell_off;
    double sum[NF], cnt[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) { sum[k] = 0.0; cnt[k] = 0.0; }
    for (int e = threadIdx.x; e < TX * TY; e += NT) {
        const int gy = tl.y0 + e / TX, gx = tl.x0 + e % TX;
        if (gy >= L || gx >= L) continue;
        const int64_t cell = (int64_t)gy * L + gx;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            double v = f[k * N + cell];
            if (k < ND) {
                if (v < 0 || !__builtin_isfinite(v)) record_fault(a.state, tl.inst, k * N + cell, counter);
                if (v < eps && v >= 0.0) {   // FKE:4459-4461
                    v = 0.0;
                    f[k * N + cell] = v;
                }
            }
            sum[k] += v;
            cnt[k] += v > 0.0 ? 1.0 : 0.0;
        }
    }
    block_partials<NT, NF>(sum, cnt, a.partials + (slot * a.ntiles + tile) * NF, s_red);
}

template <int SP, int TX, int TY, int NT>
void launch_fke(const StepArgs& a, int grid, hipStream_t s)
{
    const dim3 blocks((unsigned)a.nblocks), moves((unsigned)a.nblocks, SP - 1);
    hipLaunchKernelGGL(fke_chunk_kernel<SP>, dim3(a.nchunks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(fke_state_kernel<SP>, dim3((a.n_inst + 63) / 64), dim3(64), 0, s, a);
    hipLaunchKernelGGL(fke_percept_kernel<SP>, blocks, dim3(256), 0, s, a);
    hipLaunchKernelGGL(fke_veg_kernel<SP>, blocks, dim3(256), 0, s, a);
    for (int st = 1; st <= 4; ++st) hipLaunchKernelGGL(fke_stage_kernel<SP>, blocks, dim3(256), 0, s, a, st);
    hipLaunchKernelGGL(fke_mprep_kernel<SP>, moves, dim3(256), 0, s, a);
    hipLaunchKernelGGL((fke_move_pass_kernel<SP, 0>), moves, dim3(256), 0, s, a);
    hipLaunchKernelGGL((fke_move_pass_kernel<SP, 1>), moves, dim3(256), 0, s, a);
    hipLaunchKernelGGL((fke_move_pass_kernel<SP, 2>), moves, dim3(256), 0, s, a);
    hipLaunchKernelGGL((fke_final_kernel<SP, TX, TY, NT>), dim3(grid), dim3(NT), 0, s, a);
}

// ---- movement statistics (Rho_Mov, FKE:4034-4060) ---------------------------
// pass 0: chunk sums of gamma, vx, vy of each consumer; pass 1: of (gamma - mean)^2.
template <int SP>
__global__ __launch_bounds__(256) void fke_mstat_chunk_kernel(StepArgs a, int pass)
{
#pragma clang fp contract(off)
    constexpr int NF = 2 * SP + 1, NV = 3 * (SP - 1);
    __shared__ double red[NV][256];
    const FkeWork* w = (const FkeWork*)a.work;
    const int2 ch = w->chunks[blockIdx.x];
    const DevInst di = a.inst[ch.x];
    const FkeModelConst& mc = w->mc[di.pidx];
    const int64_t N = (int64_t)di.L * di.L;
    const double* f0 = (const double*)a.in + (int64_t)NF * di.cell_off;
    const int64_t base = (int64_t)ch.y * kFkeChunkCells;
    // the velocities of the last step taken (none before the first)
    const int64_t last = di.step0 + a.state->counter - 1;
    const bool drawn = SP == 3 && last >= 0;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int64_t i = base + threadIdx.x; i < base + kFkeChunkCells && i < N; i += 256) {
#pragma unroll
        for (int s = 1; s < SP; ++s) {
            const double g = f0[(SP + 1 + s) * N + i];
            if (pass == 0) {
                v[3 * (s - 1)] += g;
                if (drawn) {
                    const double2 u = fke_velocity(di.seed, di.inst_id, (uint32_t)i, last, s, mc.v[s]);
                    v[3 * (s - 1) + 1] += u.x;
                    v[3 * (s - 1) + 2] += u.y;
                }
            } else {
                const double m = w->mstat[(int64_t)ch.x * 8 + 3 * (s - 1)];
                v[3 * (s - 1)] += (g - m) * (g - m);
            }
        }
    }
    chunk_fold<NV>(v, red, w->chunk + (int64_t)blockIdx.x * 8);
}

template <int SP>
__global__ __launch_bounds__(64) void fke_mstat_fold_kernel(StepArgs a, int pass, double* out)
{
#pragma clang fp contract(off)
    constexpr int NV = 3 * (SP - 1);
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.n_inst) return;
    const FkeWork* w = (const FkeWork*)a.work;
    const DPCInstInfo in = w->info[i];
    const DevInst di = a.inst[i];
    const double N = (double)((int64_t)di.L * di.L);
    double tot[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = 0.0;
    for (int c = 0; c < in.nchunks; ++c) {
        const double* p = w->chunk + (int64_t)(in.chunk0 + c) * 8;
#pragma unroll
        for (int k = 0; k < NV; ++k) tot[k] += p[k];
    }
    double* o = out + (int64_t)i * 4 * SP;
    if (pass == 0) {
        for (int k = 0; k < 4; ++k) o[k] = 0.0;
        for (int s = 1; s < SP; ++s) {
            const double m = tot[3 * (s - 1)] / N;
            w->mstat[(int64_t)i * 8 + 3 * (s - 1)] = m;
            o[4 * s] = m;
            o[4 * s + 2] = tot[3 * (s - 1) + 1] / N;
            o[4 * s + 3] = tot[3 * (s - 1) + 2] / N;
        }
    } else {
        for (int s = 1; s < SP; ++s) o[4 * s + 1] = tot[3 * (s - 1)] / N;
    }
}

template <int SP>
void launch_mstat(const StepArgs& a, double* out, hipStream_t s)
{
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(fke_mstat_chunk_kernel<SP>, dim3(a.nchunks), dim3(256), 0, s, a, pass);
        hipLaunchKernelGGL(fke_mstat_fold_kernel<SP>, dim3((a.n_inst + 63) / 64), dim3(64), 0, s, a, pass, out);
    }
}

// 64 x 4 tiles (one cell per thread) for the final kernel's rho(t) partials.
const Variant kFKE2[] = {
    {64, 4, 256, 0, 1, "fke2_f64_t256", &launch_fke<2, 64, 4, 256>, 0, 0, nullptr, 1},
};
const Variant kFKE3[] = {
    {64, 4, 256, 0, 1, "fke3_f64_t256", &launch_fke<3, 64, 4, 256>, 0, 0, nullptr, 1},
};

// The reference's movement constants of consumer s of a point (FKE:195-250,
// :3764-3772), its reach W (2 W + 1 must fit the lattice), the largest shift
// reach S and its stencils appended to SD / SV.
int fke_move_of(const tm_fke_params& p, int s, bool drawn, std::vector<double>& SD, std::vector<double>& SV,
                MoveSpecies* out)
{
    tm_fke_move_species ms{};
    double sd[256], sv[256];
    if (tm_fke_move_constants(p.D_mov[s], p.v[s], p.sigma[s], p.dt_analytic, p.dt_adv[s], p.mv_invariance, p.dx,
                              p.CI, &ms, sd, 256, sv, 256) != 0)
        return TM_ERR_ARG;
    out->rD = ms.rD;
    out->rV = ms.rV;
    out->pref = ms.mu_prefactor;
    // the largest advection shift: |v cos|, |v sin| <= |v| where velocities are
    // drawn, none otherwise (as in move_args)
    const double shift = (drawn ? std::fabs(p.v[s]) : 0.0) * std::fabs(ms.mu_prefactor);
    if (!(shift < 1e9)) return TM_ERR_ARG;
    const int S = (int)std::floor(shift + 0.5) + 1;
    out->W = std::max(ms.rD, ms.rV + S);
    out->S = S;
    out->offD = (int)SD.size();
    out->offV = (int)SV.size();
    SD.insert(SD.end(), sd, sd + ms.rD + 1);
    SV.insert(SV.end(), sv, sv + ms.rV + 1);
    return TM_OK;
}

}  // namespace

const Variant* fke2_variants_f64(int* n) { *n = (int)(sizeof(kFKE2) / sizeof(kFKE2[0])); return kFKE2; }
const Variant* fke3_variants_f64(int* n) { *n = (int)(sizeof(kFKE3) / sizeof(kFKE3[0])); return kFKE3; }

void launch_fke_mov_stats(int model, const StepArgs& a, double* out, hipStream_t s)
{
    if (model == TM_MODEL_FKE3) launch_mstat<3>(a, out, s);
    else launch_mstat<2>(a, out, s);
}

int fke_setup(int model, const tm_fke_params* P, int n_params, const tm_instance_desc* inst, int n_inst,
              const int64_t* cell_off, int64_t total_cells, void** work, int* nchunks, int* nblocks,
              const char** err)
{
    const int SP = model == TM_MODEL_FKE3 ? 3 : 2, ns = SP - 1;
    *work = nullptr;
    std::vector<FkeModelConst> mc((size_t)n_params);
    std::vector<double> SD, SV;
    std::vector<int2> disk;
    std::vector<int> p_off((size_t)n_params), p_len((size_t)n_params);
    for (int g = 0; g < n_params; ++g) {
        const tm_fke_params& p = P[g];
        FkeModelConst& m = mc[(size_t)g];
        m = FkeModelConst{};
        tm_fke_rk_params rk{};
        rk.a = p.a;
        rk.c = p.c;
        rk.gmax = p.gmax;
        rk.alpha = p.alpha;
        rk.rW = p.rW;
        rk.W0 = p.W0;
        for (int k = 0; k < 3; ++k) {
            rk.K[k] = p.K[k];
            rk.E[k] = p.E[k];
            rk.M[k] = p.M[k];
            for (int l = 0; l < 3; ++l) {
                rk.A[k][l] = p.A[k][l];
                rk.H[k][l] = p.H[k][l];
            }
        }
        rk.D_W = p.D_W;
        rk.D_O = p.D_O;
        rk.dt = p.dt;
        rk.dt_rk = p.dt_analytic;
        rk.dx = p.dx;
        if (!fke_rk_args(ns, 1, &rk, &m.rk)) {
            *err = "tm_fke_params: dx and dt_analytic must be > 0 and dt finite";
            return TM_ERR_ARG;
        }
        for (int s = 1; s < SP; ++s)
            if (fke_move_of(p, s, SP == 3, SD, SV, &m.ms[s - 1]) != TM_OK) {
                *err = "tm_fke_params: bad movement constants (need D_mov, sigma, dt_analytic, dx > 0; CI >= 1)";
                return TM_ERR_ARG;
            }
        for (int s = 0; s < 3; ++s) m.v[s] = p.v[s];
        m.eps = p.flush_eps;
        // the perception table: radius (int)(max pR / dx) (FKE:3858-3864)
        double rmax = p.pR[0];
        for (int s = 1; s < SP; ++s) rmax = p.pR[s] > rmax ? p.pR[s] : rmax;
        rmax = rmax / p.dx;
        if (!(rmax > 0.0) || !(rmax < TM_DPC_MAX_RADIUS + 1.0)) {
            *err = "tm_fke_params: perception radius max pR / dx must be in (0, TM_DPC_MAX_RADIUS]";
            return TM_ERR_ARG;
        }
        const int R = (int)rmax;
        std::vector<int32_t> off(2 * (size_t)(2 * R + 1) * (2 * R + 1));
        const int len = tm_perception_offsets(R, off.data(), (int)off.size() / 2);
        p_off[(size_t)g] = (int)disk.size();
        p_len[(size_t)g] = len;
        for (int k = 0; k < len; ++k) disk.push_back(make_int2(off[2 * k], off[2 * k + 1]));
        if (!(p.sigma[0] > 0.0) || !(p.dt > 0.0) || (p.sampler != TM_SAMPLER_EXACT && p.sampler != TM_SAMPLER_COMPAT)) {
            *err = "tm_fke_params: sigma[0] and dt must be > 0 and sampler a tm_sampler";
            return TM_ERR_ARG;
        }
    }
    std::vector<DPCInstInfo> info((size_t)n_inst);
    std::vector<int2> chunks, blocks;
    for (int i = 0; i < n_inst; ++i) {
        const tm_instance_desc& d = inst[i];
        const tm_fke_params& p = P[d.params_index];
        const FkeModelConst& m = mc[(size_t)d.params_index];
        for (int s = 0; s < ns; ++s)
            if (2 * m.ms[s].W + 1 > d.L) {
                *err = "tm_fke_params: a consumer's movement reaches past its lattice (2 W + 1 > L)";
                return TM_ERR_ARG;
            }
        double rmax = p.pR[0];
        for (int s = 1; s < SP; ++s) rmax = p.pR[s] > rmax ? p.pR[s] : rmax;
        rmax = rmax / p.dx;
        DPCInstInfo& in = info[(size_t)i];
        in = DPCInstInfo{};
        in.R = (int)rmax;
        in.disk_off = p_off[(size_t)d.params_index];
        in.disk_len = p_len[(size_t)d.params_index];
        for (int s = 0; s < SP; ++s) in.fr[s] = p.pR[s] / (p.dx * rmax);
        const int64_t cells = (int64_t)d.L * d.L;
        in.chunk0 = (int)chunks.size();
        in.nchunks = (int)((cells + kFkeChunkCells - 1) / kFkeChunkCells);
        for (int k = 0; k < in.nchunks; ++k) chunks.push_back(make_int2(i, k));
        for (int64_t c0 = 0; c0 < cells; c0 += 256) blocks.push_back(make_int2(i, (int)c0));
    }
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nd = (size_t)(SP + 2) * total_cells, nm = (size_t)ns * total_cells;
    size_t o = up(sizeof(FkeWork));
    const size_t o_mc = o;     o += up(sizeof(FkeModelConst) * mc.size());
    const size_t o_info = o;   o += up(sizeof(DPCInstInfo) * info.size());
    const size_t o_st = o;     o += up(sizeof(DPCStepState) * info.size());
    const size_t o_chunk = o;  o += up(sizeof(double) * 8 * chunks.size());
    const size_t o_chunks = o; o += up(sizeof(int2) * chunks.size());
    const size_t o_disk = o;   o += up(sizeof(int2) * disk.size());
    const size_t o_blocks = o; o += up(sizeof(int2) * blocks.size());
    const size_t o_sd = o;     o += up(sizeof(double) * SD.size());
    const size_t o_sv = o;     o += up(sizeof(double) * SV.size());
    const size_t o_mstat = o;  o += up(sizeof(double) * 8 * info.size());
    const size_t o_tables = o;   // everything above is uploaded; the scratch below is not
    const size_t o_sa = o;     o += up(sizeof(double) * nd);
    const size_t o_sb = o;     o += up(sizeof(double) * nd);
    const size_t o_acc = o;    o += up(sizeof(double) * nd);
    const size_t o_coef = o;   o += up(sizeof(double4) * nm);
    const size_t o_mu = o;     o += up(sizeof(int2) * nm);
    const size_t o_mv = o;     o += up(sizeof(double) * 4 * nm);
    const size_t o_pq = o;     o += up(sizeof(double) * 2 * nm);
    void* dev = nullptr;
    if (hipMalloc(&dev, o) != hipSuccess) {
        *err = "Rietkerk consumer tables and scratch: allocation failed";
        return TM_ERR_NOMEM;
    }
    char* base = (char*)dev;
    FkeWork w{};
    w.mc = (const FkeModelConst*)(base + o_mc);
    w.info = (const DPCInstInfo*)(base + o_info);
    w.st = (DPCStepState*)(base + o_st);
    w.chunk = (double*)(base + o_chunk);
    w.chunks = (const int2*)(base + o_chunks);
    w.disk = (const int2*)(base + o_disk);
    w.blocks = (const int2*)(base + o_blocks);
    w.SD = (const double*)(base + o_sd);
    w.SV = (const double*)(base + o_sv);
    w.mstat = (double*)(base + o_mstat);
    w.sa = (double*)(base + o_sa);
    w.sb = (double*)(base + o_sb);
    w.acc = (double*)(base + o_acc);
    w.coef = (double4*)(base + o_coef);
    w.mu = (int2*)(base + o_mu);
    w.mv = (double*)(base + o_mv);
    w.pq = (double*)(base + o_pq);
    std::vector<char> host(o_tables, 0);
    auto put = [&](size_t off, const void* src, size_t n) { if (n) std::memcpy(host.data() + off, src, n); };
    put(0, &w, sizeof(w));
    put(o_mc, mc.data(), sizeof(FkeModelConst) * mc.size());
    put(o_info, info.data(), sizeof(DPCInstInfo) * info.size());
    put(o_chunks, chunks.data(), sizeof(int2) * chunks.size());
    put(o_disk, disk.data(), sizeof(int2) * disk.size());
    put(o_blocks, blocks.data(), sizeof(int2) * blocks.size());
    put(o_sd, SD.data(), sizeof(double) * SD.size());
    put(o_sv, SV.data(), sizeof(double) * SV.size());
    if (hipMemcpy(dev, host.data(), o_tables, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dev);
        *err = "Rietkerk consumer tables: upload failed";
        return TM_ERR_HIP;
    }
    *work = dev;
    *nchunks = (int)chunks.size();
    *nblocks = (int)blocks.size();
    return TM_OK;
}

}  // namespace tmk

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
    DevBuf d_coef, d_mu, d_mv, d_pq;
    if (hipMalloc(&d_coef.p, sizeof(double4) * (size_t)ns * L2) != hipSuccess ||
        hipMalloc(&d_mu.p, sizeof(int2) * (size_t)ns * L2) != hipSuccess ||
        hipMalloc(&d_mv.p, sizeof(double) * 4 * (size_t)ns * L2) != hipSuccess ||
        hipMalloc(&d_pq.p, sizeof(double) * 2 * (size_t)ns * L2) != hipSuccess)
        return TM_ERR_NOMEM;
    const dim3 grid((unsigned)((L2 + 255) / 256), (unsigned)ns);
    hipLaunchKernelGGL(fke_move_prep_kernel, grid, dim3(256), 0, st, a, rho, gamma, (const double2*)v_eff,
                       (double4*)d_coef.p, (int2*)d_mu.p);
    hipLaunchKernelGGL(fke_move_disp_kernel, grid, dim3(256), 0, st, a, (const double4*)d_coef.p,
                       (const int2*)d_mu.p, (double*)d_mv.p);
    hipLaunchKernelGGL(fke_move_row_kernel, grid, dim3(256), 0, st, a, (const double*)d_mv.p, stencil_D, stencil_V,
                       (double*)d_pq.p);
    hipLaunchKernelGGL(fke_move_col_kernel, grid, dim3(256), 0, st, a, (const double*)d_pq.p, stencil_D, stencil_V,
                       out);
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
    DevBuf d_rho, d_gam, d_v, d_out, d_sd, d_sv, d_coef, d_mu, d_mv, d_pq;
    const size_t fb = sizeof(double) * (size_t)ns * L2;
    if (hipMalloc(&d_rho.p, fb) != hipSuccess || hipMalloc(&d_gam.p, fb) != hipSuccess ||
        hipMalloc(&d_v.p, 2 * fb) != hipSuccess || hipMalloc(&d_out.p, fb) != hipSuccess ||
        hipMalloc(&d_sd.p, sizeof(double) * nD) != hipSuccess || hipMalloc(&d_sv.p, sizeof(double) * nV) != hipSuccess ||
        hipMalloc(&d_coef.p, 4 * fb) != hipSuccess || hipMalloc(&d_mu.p, sizeof(int2) * (size_t)ns * L2) != hipSuccess ||
        hipMalloc(&d_mv.p, 4 * fb) != hipSuccess || hipMalloc(&d_pq.p, 2 * fb) != hipSuccess)
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
    hipLaunchKernelGGL(fke_move_disp_kernel, grid, dim3(256), 0, nullptr, a, (const double4*)d_coef.p,
                       (const int2*)d_mu.p, (double*)d_mv.p);
    hipLaunchKernelGGL(fke_move_row_kernel, grid, dim3(256), 0, nullptr, a, (const double*)d_mv.p,
                       (const double*)d_sd.p, (const double*)d_sv.p, (double*)d_pq.p);
    hipLaunchKernelGGL(fke_move_col_kernel, grid, dim3(256), 0, nullptr, a, (const double*)d_pq.p,
                       (const double*)d_sd.p, (const double*)d_sv.p, (double*)d_out.p);
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
