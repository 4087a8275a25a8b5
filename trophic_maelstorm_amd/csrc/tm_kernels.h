// tm_kernels.h -- fused step kernels for gfx950 (templated on storage type and
// tile shape; instantiated in tm_kernels_f64.hip / tm_kernels_f32.hip).
//
// One workgroup owns a TX x TY tile of one instance.  It stages the tile and a
// periodic halo of every field in LDS (halo 1 for DP, 4 for Rietkerk, whose
// RK4 re-applies the W/O Laplacian in each of its four stages), computes the
// whole step there, and writes each field once: the HBM traffic is one read and
// one write of every field per step plus halo re-reads that hit in L2.  The
// tile is a unit of work only -- every cell draws from its own Philox stream
// keyed by its global index, so results do not depend on tile shape.
//
// Per-step rho(t): each workgroup writes FP64 (sum, count) partials of its
// tile; combine_rows later folds them in tile order with compensated sums.
//
// Steps run in blocks (one captured hipGraph, or one eager chunk):
//   prep_consts  -- per-(step, instance) constants of the block into a table
//   step k       -- k = 0 .. B-1: step number = device counter + k
//   combine_rows -- rho(t) rows of the block
//   bump_counter -- device counter += B
// so a graph replays with correct step numbers, and the step kernels carry no
// atomics, fences or serial per-tile set-up.
#pragma once
#include "tm_device.h"
#include "tm_fast32.h"
#include "tm/tm.h"

namespace tmk {

constexpr int kMaxFields = 8;   // fields per instance: 1 (DP), 3 (DP2, Rietkerk), 5 (DP3, FKE2), 7 (FKE3)

struct DevInst {
    int32_t L;
    int32_t pidx;
    uint32_t seed;
    uint32_t inst_id;
    int64_t step0;      // instance step at device counter 0
    int64_t cell_off;   // first cell of this instance in cell-indexed arrays
    int32_t tile0;      // first tile in the tile table
    int32_t ntiles;
    int32_t pad0, pad1;
};

// A tile carries its instance's side and cell offset, so a workgroup can
// issue its field loads after a single descriptor load.
struct DevTile {
    int32_t inst, x0, y0, L;
    int64_t cell_off;
};

struct DevState {
    int64_t counter;     // steps taken since configure
    int64_t run_base;    // counter at the start of the current tm_run (row index base)
    uint32_t done;       // workgroups finished in the current launch
    uint32_t pad;
    // First numeric fault, as the minimum over (launch counter, instance, cell);
    // all ones while none.  Launches run in stream order, so a later launch only
    // records when no earlier one did, and within a launch atomicMin picks the
    // same cell the reference's sequential sweep reports first.
    uint64_t fault_counter;
    uint64_t fault_key;  // (instance << 40) | cell
};

struct StepArgs {
    const DevInst* inst;
    const DevTile* tiles;
    const void* consts;  // [B][n_inst] DPConst / RKConstD of this block (prep_consts)
    const void* in;      // T*: fields, instance block [nf][L*L] at cell_off * nf
    void* out;
    double2* partials;   // [nslots][ntiles][nf]: (FP64 sum, count)
    DevState* state;
    int64_t* draws;      // nullable: Poisson integer per cell of the last step
    const DevTile* tiles_a;  // split Rietkerk step: tiles of the Dornic kernel
    void* aux;           // split Rietkerk step: T* P* field, instance block [L*L] at cell_off
    int32_t ntiles;
    int32_t nslots;
    int32_t k;           // position of this launch in its block
    int32_t n_inst;
    int32_t xcd;         // XCD-aware tile order
    int32_t ntiles_a;
    const void* work;    // consumer models: DPCWork / FkeWork (device) of the batch
    int32_t nchunks;     // consumer models: statistics chunks of the batch
    int32_t halo;        // consumer models: LDS halo (largest perception radius, >= 1)
    int32_t nblocks;     // Rietkerk consumer models: 256-cell blocks of the batch
    int32_t pad;
};

// Periodic index: v mod L in [0, L).
__device__ __forceinline__ int wrap(int v, int L)
{
    if (v >= -L && v < 2 * L) return v < 0 ? v + L : (v >= L ? v - L : v);   // the usual case
    const int m = v % L;   // lattices smaller than a tile plus its halo
    return m < 0 ? m + L : m;
}

// blockIdx -> tile.  Workgroups are dealt round-robin to the 8 XCDs; give each
// XCD a contiguous run of tiles so neighbouring tiles (which re-read each
// other's halo rows) share that XCD's L2.
__device__ __forceinline__ int tile_of(int t, int ntiles, int xcd)
{
    if (!xcd) return t;
    const int x = t & 7, j = t >> 3;
    const int q = ntiles >> 3, r = ntiles & 7;
    return x * q + (x < r ? x : r) + j;
}

// Key = (phase << 60) | (instance << 40) | cell, where phase is the position of
// the step within a multi-step launch (0 for single-step launches): a fault of
// an earlier step of the launch, recorded in any order, always has the smaller
// key, as it has the smaller counter.  The host drops the phase when it decodes
// the key.
__device__ __forceinline__ void record_fault(DevState* st, int inst, int64_t cell, int64_t counter,
                                             int phase = 0)
{
    if (st->fault_counter < (uint64_t)counter) return;
    atomicMin((unsigned long long*)&st->fault_counter, (unsigned long long)counter);
    atomicMin((unsigned long long*)&st->fault_key, ((unsigned long long)phase << 60) |
                                                       ((unsigned long long)inst << 40) | (unsigned long long)cell);
}

// Cross-lane moves of an FP64 value on the VALU (no LDS traffic, unlike
// ds_bpermute): gfx950's v_permlane32_swap / v_permlane16_swap exchange the
// upper half (odd rows) of x with the lower half (even rows) of y, and DPP
// row shifts under bank masks give lane l the value of lane l ^ 8 or l ^ 4.
__device__ __forceinline__ void lanes_swap32(double& x, double& y)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false, false);
    x = __hiloint2double(hi[0], lo[0]);
    y = __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ void lanes_swap16(double& x, double& y)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
    x = __hiloint2double(hi[0], lo[0]);
    y = __hiloint2double(hi[1], lo[1]);
}

// DPP move of both halves: lanes in BANK (groups of 4 lanes of each 16-lane
// row) take src from the lane CTRL names, the others keep old.
template <int CTRL, int BANK>
__device__ __forceinline__ double dpp_f64(double old, double src)
{
    return __hiloint2double(__builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(src), CTRL, 0xF, BANK, false),
                            __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(src), CTRL, 0xF, BANK, false));
}

// Wave totals of four FP64 values in one pass: on return every lane of row r
// (lanes 16 r .. 16 r + 15) holds the sum of v[r] over the 64 lanes.  Fixed
// order: v[0] and v[2] (v[1] and v[3]) fold lane l with l + 32 while the two
// halves trade values, the pair folds rows 2 k with 2 k + 1 the same way, and
// each row then folds lanes l ^ 8, l ^ 4, l ^ 2, l ^ 1.
__device__ __forceinline__ double wave_sum4(double v0, double v1, double v2, double v3)
{
    lanes_swap32(v0, v2);                 // v0 = (v0 lower, v2 lower), v2 = (v0 upper, v2 upper)
    lanes_swap32(v1, v3);
    double a = v0 + v2, b = v1 + v3;      // lanes < 32: v0, v1 folded; lanes >= 32: v2, v3
    lanes_swap16(a, b);
    double t = a + b;                     // row r: v[r] folded over lanes l, l + 16, l + 32, l + 48
    t += dpp_f64<0x118, 0xC>(dpp_f64<0x108, 0x3>(t, t), t);   // row_shl:8 | row_shr:8 = l ^ 8
    t += dpp_f64<0x114, 0xA>(dpp_f64<0x104, 0x5>(t, t), t);   // row_shl:4 | row_shr:4 = l ^ 4
    t += dpp_f64<0x4E, 0xF>(t, t);                            // quad_perm [2,3,0,1] = l ^ 2
    t += dpp_f64<0xB1, 0xF>(t, t);                            // quad_perm [1,0,3,2] = l ^ 1
    return t;
}

// Inclusive prefix sum over the 64 lanes of a wave on the VALU: shifts by 1, 2,
// 4 and 8 lanes within each 16-lane row, then each row adds the totals of the
// rows before it (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and
// 3).  Integer sums: exact in any order.  Every lane must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);   // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return x;
}

// Wave totals of two FP64 values: on return lanes 0..31 hold the sum of v0 over
// the 64 lanes and lanes 32..63 that of v1.  The same folds as wave_sum4: lane
// l with l + 32, then rows 2 k with 2 k + 1, then l ^ 8, l ^ 4, l ^ 2, l ^ 1.
__device__ __forceinline__ double wave_sum2(double v0, double v1)
{
    lanes_swap32(v0, v1);
    double a = v0 + v1, b = a;            // lanes < 32: v0 folded; lanes >= 32: v1
    lanes_swap16(a, b);
    double t = a + b;
    t += dpp_f64<0x118, 0xC>(dpp_f64<0x108, 0x3>(t, t), t);
    t += dpp_f64<0x114, 0xA>(dpp_f64<0x104, 0x5>(t, t), t);
    t += dpp_f64<0x4E, 0xF>(t, t);
    t += dpp_f64<0xB1, 0xF>(t, t);
    return t;
}

// Fixed-order block reduction of NF (sum, count) pairs: the 2 NF values of each
// wave through wave_sum4, four at a time, then threads 0 .. 2 NF - 1 add the
// NT / 64 wave values in wave order and write (sum, count) partials.  Every
// lane of the block must be active.
template <int NT, int NF>
__device__ __forceinline__ void block_partials(const double (&sum)[NF], const double (&cnt)[NF], double2* dst,
                                               double* red)
{
    constexpr int NW = NT / 64, NV = 2 * NF;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int g = 0; g < NV; g += 4) {
        double v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = g + r;
            v[r] = i < NV ? ((i & 1) ? cnt[i >> 1] : sum[i >> 1]) : 0.0;
        }
        const double t = wave_sum4(v[0], v[1], v[2], v[3]);
        const int i = g + (lane >> 4);
        if ((lane & 15) == 0 && i < NV) red[i * NW + w] = t;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NW; ++i) s += red[threadIdx.x * NW + i];
        ((double*)dst)[threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Arithmetic policy: FP64 keeps the reference's divisions and operation order
// (bit-exact with the CPU restatement); FP32 uses the hardware reciprocal.
// ---------------------------------------------------------------------------
template <typename T> struct Arith;
template <> struct Arith<double> {
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }
    static __device__ __forceinline__ double largest() { return 1.7976931348623157e308; }
};
template <> struct Arith<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ bool finite(float x) { return __builtin_isfinite(x); }
    static __device__ __forceinline__ float largest() { return 3.40282347e38f; }
};

// Dornic Poisson-Gamma transition of one live cell (rietkerk_bjork_FKE.cpp:4168-4196).
// ONE = true: as the FP32 dornic_cell_f<true> -- false where the transition needs
// PTRS (mean >= 12), the Gamma trial's logarithmic test (its squeeze failed) or
// a second Gamma trial, otherwise the same words and value.
template <bool ONE, class St>
__device__ __forceinline__ bool dornic_cell_d(St& s, double alpha_i, double cur, const DornicConst& k,
                                              bool compat, int64_t& n, double& y)
{
#pragma clang fp contract(off)
    const double mu = -1.0 + 2.0 * alpha_i * k.sigma2_1;
    const double ziggy = k.ll * cur;
    if (ONE && ziggy >= 12.0) return false;
    double gru;
    if (ziggy == 0.0) {
        n = 0;
        gru = mu + 1.0;
    } else {
        n = poisson_d(s, ziggy, compat);
        gru = mu + 1.0 + (double)n;
    }
    if (compat && gru > 100.0) {
        y = (norm_icdf_d(s.u64()) * dsqrt(gru) + gru) / k.lambda;
        return true;
    }
    double g;
    if (!gamma_d_try<ONE ? 1 : 64, St>(s, gru, g)) return false;
    y = g / k.lambda;
    return true;
}

__device__ __forceinline__ double dornic_cell(Stream& s, double alpha_i, double cur, const DornicConst& k,
                                              bool compat, int64_t& n)
{
    double y;
    dornic_cell_d<false, Stream>(s, alpha_i, cur, k, compat, n, y);
    return y;
}

// FP32 storage: the Poisson mean z = lambda e^{beta dt} rho in FP64 (it decides
// the sampler and is resolved as rho -> 0), the Gamma shape mu + 1 + n in FP32
// from shape0 = mu + 1 = 2 alpha_i / sigma^2 (formed by the caller in FP32), and
// the Poisson and Gamma deviates from the FP32 samplers of tm_fast32.h.
// ONE = true is the first pass of the deferring step (dp_tile_deferred): it
// returns false, leaving n and y unspecified, where the transition needs more
// inversion terms than the counted search holds, the PTRS trial's logarithmic
// test (outside its box) or a second PTRS trial, the Gamma trial's logarithmic
// test (its squeeze failed) or a second Gamma trial, or a fifth word; otherwise
// it draws the same words and computes the same value as ONE = false.
template <bool ONE, class St>
__device__ __forceinline__ bool dornic_cell_f(St& s, float shape0, double zd, const DornicConst& k, bool compat,
                                              int64_t& n, float& y)
{
    bool big = false;   // the Poisson draw took a PTRS trial's two words
    float nf;           // the count in FP32, formed from its int or FP64 source: an
                        // int64 -> float conversion costs a dozen instructions
    if (ONE) {
        // at most four words: the inversion's one or one PTRS trial's two, then one
        // Gamma trial's two, and the boost's where the inversion left room for it
        if (compat && zd > 100.0) return false;
        if (zd < 12.0) {
            int ni = 0;
            if (zd > 0.0) {
                float S, P;
                if (!poisson_count((float)zd, s.v32(), ni, S, P)) return false;
            }
            n = ni;
            nf = (float)ni;
        } else {
            // one PTRS trial, settled by its box alone: a trial outside it (about a
            // fifth at large means) is finished by the full sampler in the second
            // pass, so no lane of the first pass waits on its logarithms
            const PtrsF pt = ptrs_setup(zd);
            const float U = s.u32() - 0.5f;
            const float V = s.u32();
            const float us = 0.5f - __builtin_fabsf(U);
            if (!ptrs_in_box(pt, us, V)) return false;
            const double kd = ptrs_k(pt, U, us);
            n = (int64_t)kd;
            nf = (float)kd;
            big = true;
        }
    } else {
        const double nd = poisson_fd(s, zd, compat);
        n = (int64_t)nd;
        nf = (float)nd;
    }
    const float gru = shape0 + nf;
    float g;
    if (compat && gru > 100.0f) {
        g = norm_icdf_f(s.u32()) * __builtin_sqrtf(gru) + gru;
    } else {
        if (ONE && big && gru < 1.0f) return false;
        if (!gamma_f_try<ONE ? 1 : 64, St>(s, gru, g)) return false;
    }
    // g / lambda with 1/lambda carried as hi + lo: a single FP32 rounding of 1/lambda
    // would bias every cell by the same relative amount every step.
    y = __builtin_fmaf(g, k.inv_lambda_hi, g * k.inv_lambda_lo);
    return true;
}

__device__ __forceinline__ float dornic_cell_f32(Stream& s, float shape0, double zd, const DornicConst& k,
                                                 bool compat, int64_t& n)
{
    float y;
    dornic_cell_f<false, Stream>(s, shape0, zd, k, compat, n, y);
    return y;
}

// The Dornic transition of a cell that is not absorbing, from its value and its
// four neighbours (old field), in the storage type's arithmetic: FP64 keeps
// alpha_i and every Dornic quantity in FP64 in the reference's order; FP32 sums
// the neighbours in FP32 (same order) and forms the Poisson mean in FP64.
// Returns cur unchanged where alpha_i = 0 and cur = 0 (no draw, n = -1).
template <typename T> struct DornicLive;
template <> struct DornicLive<double> {
    static __device__ __forceinline__ double step(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step,
                                                  double cur, double up, double dn, double lf, double rt,
                                                  const DornicConst& k, bool compat, int64_t& n);
};
template <> struct DornicLive<float> {
    static __device__ __forceinline__ float step(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step, float cur,
                                                 float up, float dn, float lf, float rt, const DornicConst& k,
                                                 bool compat, int64_t& n)
    {
        const float nsum = ((up + dn) + lf) + rt;
        if (nsum == 0.0f && cur == 0.0f) return cur;
        Stream st(seed, inst, cell, step);
        return dornic_cell_f32(st, k.shape_f * nsum, k.ll * (double)cur, k, compat, n);
    }
};

// alpha_i = (D/dx^2)(up + down + left + right) in FP64, in the reference's order
// (rietkerk_bjork_FKE.cpp:4158-4159).
template <typename T>
__device__ __forceinline__ double nbr_alpha(double dc, T up, T dn, T lf, T rt)
{
#pragma clang fp contract(off)
    return dc * ((double)up + (double)dn + (double)lf + (double)rt);
}

__device__ __forceinline__ double DornicLive<double>::step(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step,
                                                           double cur, double up, double dn, double lf, double rt,
                                                           const DornicConst& k, bool compat, int64_t& n)
{
    const double alpha_i = nbr_alpha(k.dc, up, dn, lf, rt);
    if (alpha_i == 0.0 && cur == 0.0) return cur;
    Stream st(seed, inst, cell, step);
    return dornic_cell(st, alpha_i, cur, k, compat, n);
}

// ---------------------------------------------------------------------------
// DP nonlinear sub-step (f = -b rho^2, multiSPDP.cpp:1662).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T dp_f(T b, T y) { return -b * y * y; }

template <typename T>
__device__ __forceinline__ T dp_nonlinear(int scheme, T b, T dt, T y)
{
    if (scheme == TM_DP_EXACT) return Arith<T>::div(y, (T)1 + b * y * dt);
    if (scheme == TM_DP_RK4) {
        const T dt2 = dt / (T)2.0, dt6 = dt / (T)6.0;
        const T k1 = dp_f(b, y);
        T ys = y + dt2 * k1;
        const T k2 = dp_f(b, ys);
        ys = y + dt2 * k2;
        const T k3 = dp_f(b, ys);
        ys = y + dt * k3;
        const T k4 = dp_f(b, ys);
        return y + dt6 * (k1 + (T)2.0 * k2 + (T)2.0 * k3 + k4);
    }
    return y + dt * dp_f(b, y);
}

struct DPConst {
    DornicConst k;
    double b, dt, eps, neg;
    int scheme, compat;
    int64_t step;
};

__device__ __forceinline__ DPConst dp_constants(const tm_dp_params& p, int64_t step)
{
#pragma clang fp contract(off)
    double a = p.a;
    if (p.ramp_steps > 0) {
        const int64_t m = step < p.ramp_steps ? step : p.ramp_steps;
        a = p.a + (p.a_end - p.a) * ((double)m / (double)p.ramp_steps);
    }
    DPConst c;
    c.k = dornic_constants(a, p.D, p.sigma, p.dt, p.dx);
    c.b = p.b;
    c.dt = p.dt;
    c.eps = p.flush_eps;
    c.neg = p.flush_neg;
    c.scheme = p.scheme;
    c.compat = p.sampler == TM_SAMPLER_COMPAT;
    c.step = step;
    return c;
}

// ---------------------------------------------------------------------------
// DP step: Dornic + nonlinear sub-step + checks + flush + rho(t) partials.
// multiSPDP.cpp:2901-2958 and :3222-3262.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t as_bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ uint32_t as_bits(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (uint32_t)b | (uint32_t)(b >> 32);
}

// Wave-level (sum, count) reduction through wave_sum2; lanes 0 and 32 write the
// wave's (sum, count) partial.  Every lane of the wave must be active.
__device__ __forceinline__ void wave_partial(double sum, double cnt, double2* dst)
{
    const double t = wave_sum2(sum, cnt);
    const int lane = threadIdx.x & 63;
    if ((lane & 31) == 0) ((double*)dst)[lane >> 5] = t;
}

// Wave partial (sum, count) of one step into its slot; an all-zero wave writes zeros.
template <int NT>
__device__ __forceinline__ void step_partial(const StepArgs& a, int tile, int64_t counter, double s,
                                             uint32_t nlive)
{
    const int64_t slot = counter & (int64_t)(a.nslots - 1);
    double2* dst = a.partials + (slot * a.ntiles + tile) * (NT / 64) + (threadIdx.x >> 6);
    if (__ballot(nlive != 0 || s != 0.0) == 0) {
        if ((threadIdx.x & 63) == 0) *dst = make_double2(0.0, 0.0);
    } else {
        wave_partial(s, (double)nlive, dst);
    }
}

// Asynchronous copy of a tile and its halo into LDS with global_load_lds: each
// wave instruction lands 64 lanes' data at consecutive LDS addresses, so the
// staging costs no registers and overlaps whatever the workgroup does until
// its next s_waitcnt.  The LDS image is row-major: RH rows [y0 - H, ...) of
// RWV = TX + 2 HC elements covering columns [x0 - HC, x0 + TX + HC), periodic,
// where HC is the halo H rounded up to whole 16-byte chunks of V = 16 / sizeof(T)
// elements.  When L is a multiple of V every 16-byte chunk of a row is
// contiguous in memory and one global_load_lds_dwordx4 moves it; otherwise the
// image is filled dword by dword.
template <typename T, int H>
struct LdsFrame {
    static constexpr int V = 16 / (int)sizeof(T);
    static constexpr int HC = ((H + V - 1) / V) * V;
};

template <typename T, int TX, int RH, int NT, int H = 1>
__device__ __forceinline__ void tile_to_lds(const T* fin, int L, int x0, int y0, uint32_t* lds)
{
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr int V = LdsFrame<T, H>::V;
    constexpr int HC = LdsFrame<T, H>::HC;
    constexpr int RWV = TX + 2 * HC;
    const int tid = threadIdx.x;
    const int wave_base = tid & ~63;
    if (L % V == 0) {
        constexpr int CPR = RWV / V;                  // 16-byte chunks per row
        constexpr int NC = RH * CPR;
        constexpr int NLD = (NC + NT - 1) / NT;
        // not unrolled: the copies are fire-and-forget, nothing stays live
#pragma unroll 1
        for (int j = 0; j < NLD; ++j) {
            const int q = j * NT + tid;
            if (q < NC) {   // lanes past the image copy nothing, so the buffer can end with it
                const int ry = q / CPR, cx = q - (q / CPR) * CPR;
                int gy = y0 + ry - H, gx = x0 - HC + cx * V;
                gy = gy < 0 ? gy + L : (gy >= L ? gy - L : gy);   // one wrap: L >= RH, L >= RWV
                gx = gx < 0 ? gx + L : (gx >= L ? gx - L : gx);
                if (L < RWV || L < RH) {
                    gy = wrap(y0 + ry - H, L);
                    gx = wrap(x0 - HC + cx * V, L);
                }
                const uint32_t off = (uint32_t)gy * (uint32_t)L + (uint32_t)gx;
                __builtin_amdgcn_global_load_lds(fin + off, lds + 4 * (j * NT + wave_base), 16, 0, 0);
            }
        }
    } else {
        constexpr int DW = sizeof(T) / 4;             // dwords per element
        constexpr int NQ = RH * RWV * DW;
        constexpr int NLD = (NQ + NT - 1) / NT;
        const uint32_t* src0 = (const uint32_t*)fin;
#pragma unroll 1
        for (int j = 0; j < NLD; ++j) {
            const int q = j * NT + tid;
            if (q < NQ) {
                const int e = q / DW, w = q - (q / DW) * DW;
                const int ry = e / RWV, rx = e - (e / RWV) * RWV;
                const int gy = wrap(y0 + ry - H, L), gx = wrap(x0 - HC + rx, L);
                const uint32_t off = ((uint32_t)gy * (uint32_t)L + (uint32_t)gx) * DW + w;
                __builtin_amdgcn_global_load_lds(src0 + off, lds + j * NT + wave_base, 4, 0, 0);
            }
        }
    }
#endif
}

// As tile_to_lds for NFR fields of one tile at once (same L, same tile): each
// chunk's address is formed once and serves every field's copy.
template <typename T, int TX, int RH, int NT, int H, int NFR>
__device__ __forceinline__ void tiles_to_lds(const T* const (&fin)[NFR], int L, int x0, int y0,
                                             uint32_t* const (&lds)[NFR])
{
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr int V = LdsFrame<T, H>::V;
    constexpr int HC = LdsFrame<T, H>::HC;
    constexpr int RWV = TX + 2 * HC;
    if (L % V != 0 || L < RWV || L < RH) {
#pragma unroll
        for (int f = 0; f < NFR; ++f) tile_to_lds<T, TX, RH, NT, H>(fin[f], L, x0, y0, lds[f]);
        return;
    }
    const int tid = threadIdx.x;
    const int wave_base = tid & ~63;
    constexpr int CPR = RWV / V;                  // 16-byte chunks per row
    constexpr int NC = RH * CPR;
    constexpr int NLD = (NC + NT - 1) / NT;
#pragma unroll 1
    for (int j = 0; j < NLD; ++j) {
        const int q = j * NT + tid;
        uint32_t off = 0;
        if (q < NC) {
            const int ry = q / CPR, cx = q - (q / CPR) * CPR;
            int gy = y0 + ry - H, gx = x0 - HC + cx * V;
            gy = gy < 0 ? gy + L : (gy >= L ? gy - L : gy);   // one wrap: L >= RH, L >= RWV
            gx = gx < 0 ? gx + L : (gx >= L ? gx - L : gx);
            off = (uint32_t)gy * (uint32_t)L + (uint32_t)gx;
        }
#pragma unroll
        for (int f = 0; f < NFR; ++f)
            __builtin_amdgcn_global_load_lds(fin[f] + off, lds[f] + 4 * (j * NT + wave_base), 16, 0, 0);
    }
#endif
}

template <typename T> struct Quad { T v[4]; };

template <typename T>
__device__ __forceinline__ T sel4(const Quad<T>& q, int i)
{
    return i == 0 ? q.v[0] : (i == 1 ? q.v[1] : (i == 2 ? q.v[2] : q.v[3]));
}


// Store of the group of 4 cells at row[gx0 ..]: one 16-byte store (two for
// FP64) when the instance's rows are 16-byte aligned and the group is whole,
// else per cell.  The per-cell form is a rolled loop on purpose: written as
// straight-line single stores next to the 16-byte form, the two were merged by
// the optimiser into four single-word stores on every path.
template <typename T>
__device__ __forceinline__ void store_quad(T* row, int gx0, int nvalid, bool vec, const Quad<T>& q)
{
    if (vec && nvalid == 4) {
        if constexpr (sizeof(T) == 4) {
            *(float4*)(row + gx0) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
        } else {
            *(double2*)(row + gx0) = make_double2(q.v[0], q.v[1]);
            *(double2*)(row + gx0 + 2) = make_double2(q.v[2], q.v[3]);
        }
    } else {
#pragma unroll 1
        for (int i = 0; i < nvalid; ++i) row[gx0 + i] = sel4(q, i);
    }
}

// MODE 0: the DP step.  MODE 1: the Dornic sub-step of the Rietkerk P field
// alone (rietkerk_bjork_FKE.cpp:4154-4211), writing P* to a.aux for the RK4
// kernel -- same staging, absorbing-group skip and sampler, on its own tiles.
struct RKConstD;
template <int MODE> struct StepConst;
template <> struct StepConst<0> { typedef DPConst type; };
template <> struct StepConst<1> { typedef RKConstD type; };

template <typename T, int TX, int TY, int NT>
struct DPShape {
    static constexpr int V = 16 / (int)sizeof(T);
    static constexpr int RWV = TX + 2 * V, RH = TY + 2;      // LDS image (see tile_to_lds)
    static constexpr int CG = TX / 4;                        // 4-cell groups per tile row
    static constexpr int R4 = NT / CG;                       // tile rows covered per pass
    static constexpr int G = TY / R4;                        // groups per thread per tile
    static constexpr int NW = NT / 64;                       // waves = partials per tile
    static constexpr int NQ = RH * RWV * (int)(sizeof(T) / 4);
    // dwords per image: the image itself (tile_to_lds copies nothing past it).
    // Rounded up to whole DMA rounds it cost the two-pass 128 x 24 FP32 kernel a
    // sixth workgroup per CU (28.7 KB of LDS against 26.5 KB).
    static constexpr int NBUF = NQ;
    static_assert(NQ % 4 == 0, "whole 16-byte chunks");
    static_assert(NT % CG == 0 && TY % R4 == 0 && TX % 64 == 0, "tile shape");
};

// The step of one cell from its value and its four neighbours' (old field):
// the absorbing skip (FKE:4161-4164), the Dornic transition, and for MODE 0 the
// DP nonlinear sub-step and the flush.  n gets the Poisson integer (-1 where
// skipped); bad is set when MODE 0 produced a negative or non-finite value.
template <typename T, int MODE>
__device__ __forceinline__ T cell_update(const typename StepConst<MODE>::type& c, const DevInst& di,
                                         uint32_t cell, T cur, T up, T dn, T lf, T rt, int64_t& n, bool& bad)
{
    T y = (T)0;
    n = -1;
    bad = false;
    if (!(cur == (T)0 && up == (T)0 && dn == (T)0 && lf == (T)0 && rt == (T)0)) {
        y = DornicLive<T>::step(di.seed, di.inst_id, cell, c.step, cur, up, dn, lf, rt, c.k, c.compat, n);
        if constexpr (MODE == 0) {
            y = dp_nonlinear<T>(c.scheme, (T)c.b, (T)c.dt, y);
            bad = y < (T)0 || !Arith<T>::finite(y);
            if (y < (T)c.eps && y >= (T)c.neg) y = (T)0;
        }
    }
    return y;
}

// First pass of the deferring step: cell_update<T, MODE> where the one-trial
// transition succeeds (returns true, same y, n and bad); false where the cell
// has to be computed again by cell_update.
template <typename T, int MODE>
__device__ __forceinline__ bool cell_update_first(const typename StepConst<MODE>::type& c, const DevInst& di,
                                                  uint32_t cell, T cur, T up, T dn, T lf, T rt, T& y, int64_t& n,
                                                  bool& bad)
{
    static_assert(MODE == 0 || sizeof(T) == 4, "the Rietkerk Dornic first pass is FP32");
    y = (T)0;
    n = -1;
    bad = false;
    if (!(cur == (T)0 && up == (T)0 && dn == (T)0 && lf == (T)0 && rt == (T)0)) {
        y = cur;
        if constexpr (sizeof(T) == 4) {
            const float nsum = ((up + dn) + lf) + rt;
            if (!(nsum == 0.0f && cur == 0.0f)) {
                Stream4 st(di.seed, di.inst_id, cell, c.step);
                if (!dornic_cell_f<true, Stream4>(st, c.k.shape_f * nsum, c.k.ll * (double)cur, c.k, c.compat != 0,
                                                  n, y))
                    return false;
            }
        } else {
            const double alpha_i = nbr_alpha(c.k.dc, up, dn, lf, rt);
            if (!(alpha_i == 0.0 && cur == (T)0)) {
                Stream4 st(di.seed, di.inst_id, cell, c.step);
                if (!dornic_cell_d<true, Stream4>(st, alpha_i, cur, c.k, c.compat != 0, n, y)) return false;
            }
        }
        if constexpr (MODE == 0) {
            y = dp_nonlinear<T>(c.scheme, (T)c.b, (T)c.dt, y);
            bad = y < (T)0 || !Arith<T>::finite(y);
            if (y < (T)c.eps && y >= (T)c.neg) y = (T)0;
        }
    }
    return true;
}

// Whether the first NQ dwords of an LDS image are all zero (every thread of the
// workgroup calls it; one barrier).  An image of +0 only is an absorbing tile:
// the step leaves every cell at 0 and draws nothing.
template <int NT, int NQ>
__device__ __forceinline__ bool image_all_zero(const uint32_t* img, uint32_t* s_flag)
{
    static_assert(NQ % 4 == 0, "16-byte reads");
    uint32_t any = 0;
    const uint4* v = (const uint4*)img;
    for (int i = threadIdx.x; i < NQ / 4; i += NT) {
        const uint4 x = v[i];
        any |= x.x | x.y | x.z | x.w;
    }
    const uint64_t b = __ballot(any != 0);
    if ((threadIdx.x & 63) == 0) s_flag[threadIdx.x >> 6] = b != 0 ? 1u : 0u;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r |= s_flag[w];
    return r == 0;
}

// The absorbing tile's output: zeros in its valid cells (16-byte stores where
// aligned), draws -1, and zero rho(t) partials in each of the nsl given slots.
template <typename T, int TX, int TY, int NT>
__device__ __forceinline__ void zero_tile(const StepArgs& a, int tile, const DevTile& tl, T* fout,
                                          const int64_t* slots, int nsl)
{
    constexpr int CG = TX / 4;
    const int L = tl.L;
    const bool vec = (L % 4 == 0) && (tl.cell_off % 4 == 0);
    Quad<T> z;
    z.v[0] = z.v[1] = z.v[2] = z.v[3] = (T)0;
    for (int g = threadIdx.x; g < TY * CG; g += NT) {
        const int gy = tl.y0 + g / CG, gx0 = tl.x0 + 4 * (g % CG);
        if (gy >= L || gx0 >= L) continue;
        const int nvalid = min(4, L - gx0);
        store_quad(fout + (int64_t)gy * L, gx0, nvalid, vec, z);
        if (a.draws) {
            const int64_t c0 = tl.cell_off + (int64_t)gy * L + gx0;
#pragma unroll 1
            for (int q = 0; q < nvalid; ++q) a.draws[c0 + q] = -1;
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < nsl; ++k)
            a.partials[(slots[k] * a.ntiles + tile) * (NT / 64) + (threadIdx.x >> 6)] = make_double2(0.0, 0.0);
}

// One tile from its LDS image: the per-cell step on each of this thread's G
// 4-cell groups (rows ly0 + j R4), then (MODE 0) the wave's rho(t) partial.
template <typename T, int TX, int TY, int NT, int MODE>
__device__ __forceinline__ void dp_tile(const StepArgs& a, int tile, const DevTile& tl, const DevInst& di,
                                        const typename StepConst<MODE>::type& c, const T* s_rho,
                                        int64_t counter, int64_t slot)
{
    typedef DPShape<T, TX, TY, NT> S;
    constexpr int V = S::V, RWV = S::RWV, R4 = S::R4;
    const int tid = threadIdx.x;
    const int cg = tid % S::CG, ly0 = tid / S::CG;
    const int L = tl.L;
    T* fout = (MODE == 1 ? (T*)a.aux : (T*)a.out) + tl.cell_off;
    const bool vec = (L % 4 == 0) && (tl.cell_off % 4 == 0);
    double sum[1] = {0.0};
    uint32_t nlive = 0;
    // A group whose 14-value stencil is all +0 is absorbing (the reference
    // skips each of its cells, FKE:4161-4164) and only stores zeros; other
    // groups run the per-cell step.
    for (int j = 0; j < S::G; ++j) {
        const int ly = ly0 + j * R4;
        const int gy = tl.y0 + ly, gx0 = tl.x0 + 4 * cg;
        if (gy >= L || gx0 >= L) continue;
        const int r = (ly + 1) * RWV + V + 4 * cg;
        const uint32_t cell0 = (uint32_t)gy * (uint32_t)L + (uint32_t)gx0;
        const bool full = gx0 + 4 <= L;
        uint32_t bits = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bits |= as_bits(s_rho[r + q]) | as_bits(s_rho[r - RWV + q]) | as_bits(s_rho[r + RWV + q]);
        bits |= as_bits(s_rho[r - 1]) | as_bits(s_rho[r + 4]);
        const int nvalid = min(4, L - gx0);
        Quad<T> yq;
        yq.v[0] = yq.v[1] = yq.v[2] = yq.v[3] = (T)0;
        if (bits == 0 && full) {
            if (a.draws) {
#pragma unroll
                for (int q = 0; q < 4; ++q) a.draws[tl.cell_off + cell0 + q] = -1;
            }
        } else {
#pragma unroll 1
            for (int q = 0; q < nvalid; ++q) {
                const int rq = r + q;
                const uint32_t cell = cell0 + q;
                int64_t n;
                bool bad;
                const T y = cell_update<T, MODE>(c, di, cell, s_rho[rq], s_rho[rq - RWV], s_rho[rq + RWV],
                                                 s_rho[rq - 1], s_rho[rq + 1], n, bad);
                if (bad) record_fault(a.state, tl.inst, cell, counter);
                if (a.draws) a.draws[tl.cell_off + cell] = n;
                if constexpr (MODE == 0) {
                    sum[0] += (double)y;
                    nlive += y > (T)0 ? 1u : 0u;
                }
                yq.v[0] = q == 0 ? y : yq.v[0];
                yq.v[1] = q == 1 ? y : yq.v[1];
                yq.v[2] = q == 2 ? y : yq.v[2];
                yq.v[3] = q == 3 ? y : yq.v[3];
            }
        }
        store_quad(fout + (int64_t)gy * L, gx0, nvalid, vec, yq);
    }
    if constexpr (MODE == 0) {
        double2* dst = a.partials + (slot * a.ntiles + tile) * S::NW + (tid >> 6);
        if (__ballot(nlive != 0 || sum[0] != 0.0) == 0) {   // an all-zero wave: no reduction
            if ((tid & 63) == 0) *dst = make_double2(0.0, 0.0);
        } else {
            wave_partial(sum[0], (double)nlive, dst);
        }
    }
}

// The DP step of one tile in two passes.  A Gamma trial is rejected in a few
// percent of the cells, but a wavefront runs the next trial, and the second
// Philox block it needs, whenever any of its 64 lanes needs it, which is nearly
// always.  So the first pass stops each cell's transition after one trial: a
// cell whose first PTRS or Gamma trial was rejected, or that needs more words
// than the first Philox block, goes to a queue in LDS.  After a barrier the
// second pass runs the full step (cell_update) on the queued cells, packed
// densely into the lanes.  The first pass is the first stage of the full step
// (same words, same arithmetic), so the field and the draws are those of
// dp_tile.  The queue is filled in thread order (a bit mask per thread and a
// workgroup prefix sum); it holds every cell of the tile, so it cannot
// overflow.  A queue slot holds the cell's tile offset, then its new value,
// which the cell's owner thread reads back after a second barrier.  Each
// thread then stores its own 4-cell groups once (a patch store of single
// deferred cells would fetch their whole cache lines again) and sums its own
// cells in dp_tile's order, so the rho(t) partials are dp_tile's bit for bit.
template <typename T, int TX, int TY, int NT, int MODE>
__device__ __forceinline__ void dp_tile_deferred(const StepArgs& a, int tile, const DevTile& tl, const DevInst& di,
                                                 const typename StepConst<MODE>::type& c, const T* s_rho,
                                                 int64_t counter, T* s_q, uint32_t* s_wn)
{
    typedef DPShape<T, TX, TY, NT> S;
    static_assert(S::G * 4 <= 32, "one mask bit per cell of a thread");
    constexpr int V = S::V, RWV = S::RWV, R4 = S::R4, G = S::G;
    const int tid = threadIdx.x;
    const int cg = tid % S::CG, ly0 = tid / S::CG;
    const int L = tl.L;
    T* fout = (MODE == 1 ? (T*)a.aux : (T*)a.out) + tl.cell_off;
    const bool vec = (L % 4 == 0) && (tl.cell_off % 4 == 0);
    uint32_t dmask = 0;                                       // bit 4 j + q: deferred
    Quad<T> own[G];                                           // first-pass values (0 where deferred)
#pragma unroll
    for (int j = 0; j < G; ++j) {
        own[j].v[0] = own[j].v[1] = own[j].v[2] = own[j].v[3] = (T)0;
        const int ly = ly0 + j * R4;
        const int gy = tl.y0 + ly, gx0 = tl.x0 + 4 * cg;
        if (gy >= L || gx0 >= L) continue;
        const int r = (ly + 1) * RWV + V + 4 * cg;
        const uint32_t cell0 = (uint32_t)gy * (uint32_t)L + (uint32_t)gx0;
        const bool full = gx0 + 4 <= L;
        uint32_t bits = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bits |= as_bits(s_rho[r + q]) | as_bits(s_rho[r - RWV + q]) | as_bits(s_rho[r + RWV + q]);
        bits |= as_bits(s_rho[r - 1]) | as_bits(s_rho[r + 4]);
        const int nvalid = min(4, L - gx0);
        Quad<T> yq;
        yq.v[0] = yq.v[1] = yq.v[2] = yq.v[3] = (T)0;
        if (bits == 0 && full) {
            if (a.draws) {
#pragma unroll
                for (int q = 0; q < 4; ++q) a.draws[tl.cell_off + cell0 + q] = -1;
            }
        } else {
#pragma unroll 1
            for (int q = 0; q < nvalid; ++q) {
                const int rq = r + q;
                const uint32_t cell = cell0 + q;
                T y;
                int64_t n;
                bool bad;
                const bool done = cell_update_first<T, MODE>(c, di, cell, s_rho[rq], s_rho[rq - RWV],
                                                             s_rho[rq + RWV], s_rho[rq - 1], s_rho[rq + 1], y, n,
                                                             bad);
                if (!done) {
                    dmask |= 1u << (4 * j + q);
                    y = (T)0;
                    n = -1;
                } else if (bad) {
                    record_fault(a.state, tl.inst, cell, counter);
                }
                if (a.draws) a.draws[tl.cell_off + cell] = n;
                yq.v[0] = q == 0 ? y : yq.v[0];
                yq.v[1] = q == 1 ? y : yq.v[1];
                yq.v[2] = q == 2 ? y : yq.v[2];
                yq.v[3] = q == 3 ? y : yq.v[3];
            }
        }
        own[j] = yq;
    }
    // Queue offsets: an inclusive scan of the per-thread counts within the wave,
    // then the totals of the waves before it.
    const uint32_t ndef = (uint32_t)__builtin_popcount(dmask);
    const int lane = tid & 63;
    const uint32_t incl = wave_incl_scan(ndef);
    if (lane == 63) s_wn[tid >> 6] = incl;
    __syncthreads();
    uint32_t base = incl - ndef, nq = 0;
#pragma unroll
    for (int w = 0; w < S::NW; ++w) {
        base += w < (tid >> 6) ? s_wn[w] : 0u;
        nq += s_wn[w];
    }
    uint32_t* const s_qo = (uint32_t*)s_q;                    // slot i: offset at word i * DW
    constexpr int DW = (int)(sizeof(T) / 4);
    {
        uint32_t b = base;
#pragma unroll 1
        for (uint32_t m = dmask; m != 0; m &= m - 1) {
            const int k = __builtin_ctz(m);
            s_qo[(b++) * DW] = (uint32_t)((ly0 + (k >> 2) * R4) * TX + 4 * cg + (k & 3));
        }
    }
    __syncthreads();
#pragma unroll 1
    for (uint32_t i = tid; i < nq; i += NT) {
        const uint32_t e = s_qo[i * DW];
        const int ly = (int)(e / TX), lx = (int)(e % TX);
        const int r = (ly + 1) * RWV + V + lx;
        const uint32_t cell = (uint32_t)(tl.y0 + ly) * (uint32_t)L + (uint32_t)(tl.x0 + lx);
        int64_t n;
        bool bad;
        const T y = cell_update<T, MODE>(c, di, cell, s_rho[r], s_rho[r - RWV], s_rho[r + RWV], s_rho[r - 1],
                                         s_rho[r + 1], n, bad);
        if (bad) record_fault(a.state, tl.inst, cell, counter);
        if (a.draws) a.draws[tl.cell_off + cell] = n;
        s_q[i] = y;                                           // the slot's thread read its offset above
    }
    __syncthreads();
    // Each thread stores its own groups once, the deferred cells read back from
    // their queue slots (in bit order, as they were queued): whole 16-byte
    // stores, no second write of a cell.  Then (MODE 0) the rho(t) partials of
    // the same values in dp_tile's order.
    double sum = 0.0;
    uint32_t nlive = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const int ly = ly0 + j * R4;
        const int gy = tl.y0 + ly, gx0 = tl.x0 + 4 * cg;
        if (gy >= L || gx0 >= L) continue;
        const int nvalid = min(4, L - gx0);
        Quad<T> yq = own[j];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q >= nvalid) break;
            if ((dmask >> (4 * j + q)) & 1u) yq.v[q] = s_q[base++];
            if constexpr (MODE == 0) {
                sum += (double)yq.v[q];
                nlive += yq.v[q] > (T)0 ? 1u : 0u;
            }
        }
        store_quad(fout + (int64_t)gy * L, gx0, nvalid, vec, yq);
    }
    if constexpr (MODE == 0) step_partial<NT>(a, tile, counter, sum, nlive);
}

// MODE 0: the DP step.  MODE 1: the Dornic sub-step of the Rietkerk P field
// alone (rietkerk_bjork_FKE.cpp:4154-4211), writing P* to a.aux for the RK4
// kernel -- same staging, absorbing-group skip and sampler, on its own tiles.
// Q = 1: the two-pass tile (dp_tile_deferred) with its queue of deferred
// cells -- the DP step in either precision, the Rietkerk Dornic sub-step in
// FP32.  One tile per workgroup: its descriptor, instance and constants are
// read before the kernel writes anything, so they are scalar loads issued
// together, and the image copy follows at once.
template <typename T, int TX, int TY, int NT, int MODE, int MINW = 1, int Q = 0>
__global__ __launch_bounds__(NT, MINW) void dp_step_kernel(StepArgs a)
{
    typedef DPShape<T, TX, TY, NT> S;
    typedef typename StepConst<MODE>::type Const;
    constexpr int NFIN = MODE == 1 ? 3 : 1;           // fields per instance block of `in`
    static_assert(Q == 0 || MODE == 0 || sizeof(T) == 4, "deferring Rietkerk Dornic step: FP32 only");
    __shared__ __align__(16) uint32_t s_img[S::NBUF];
    __shared__ uint32_t s_flag[NT / 64];
    __shared__ __align__(8) T s_q[Q ? TX * TY : 1];   // deferred cells: tile offsets, then values
    __shared__ uint32_t s_wn[NT / 64];                // ... and their count per wave

    const int64_t counter = a.state->counter + a.k;
    const int64_t slot = counter & (int64_t)(a.nslots - 1);   // nslots: a power of two
    const DevTile* tiles = MODE == 1 ? a.tiles_a : a.tiles;
    const int ntiles = MODE == 1 ? a.ntiles_a : a.ntiles;
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    const DevTile tl = tiles[tile];
    const DevInst di = a.inst[tl.inst];
    const Const c = ((const Const*)a.consts)[(int64_t)a.k * a.n_inst + tl.inst];
    tile_to_lds<T, TX, S::RH, NT>((const T*)a.in + NFIN * tl.cell_off, tl.L, tl.x0, tl.y0, s_img);
    __builtin_amdgcn_s_waitcnt(0);   // this thread's copies have landed
    __syncthreads();                 // ... everyone's
    if (image_all_zero<NT, S::NQ>(s_img, s_flag)) {
        zero_tile<T, TX, TY, NT>(a, tile, tl, (MODE == 1 ? (T*)a.aux : (T*)a.out) + tl.cell_off, &slot,
                                 MODE == 0 ? 1 : 0);
        return;
    }
    if constexpr (Q != 0)
        dp_tile_deferred<T, TX, TY, NT, MODE>(a, tile, tl, di, c, (const T*)s_img, counter, s_q, s_wn);
    else
        dp_tile<T, TX, TY, NT, MODE>(a, tile, tl, di, c, (const T*)s_img, counter, slot);
}

// ---------------------------------------------------------------------------
// K DP steps per launch (temporal blocking): steps k = a.k .. k + K - 1 of one
// tile from one radius-K image of the field.  Phase j (j = 1 .. K - 1) computes
// step k + j - 1 on the tile and its radius-(K - j) ring into an LDS frame (two
// frames, used in turn); the ring's cells belong to the neighbouring tiles,
// which compute the same values from the same inputs and Philox counters, so
// the result is the single-step result bit for bit.  The last phase computes
// step k + K - 1 of the tile from the last frame and writes it.  The field is
// read and written once per K steps.  Only the tile's own cells count in the
// rho(t) partials (one set per step) and record faults; draws are recorded for
// the last step (the last step of a run that ends on this launch).  The image
// and the frames share one layout: row width RW, tile column lx at HC + lx,
// tile row ly at (ly + radius) rows from the top.
// ---------------------------------------------------------------------------
template <typename T, int TX, int TY, int NT, int K>
struct DPKShape {
    static constexpr int HC = LdsFrame<T, K>::HC;            // column halo (whole 16-byte chunks)
    static constexpr int RW = TX + 2 * HC;                   // row width of the image and the frames
    static constexpr int RH = TY + 2 * K;                    // image rows: tile rows -K .. TY + K - 1
    static constexpr int FH = TY + 2 * (K - 1);              // frame rows (the largest frame)
    static constexpr int CG = TX / 4;                        // 4-cell groups per row
    static constexpr int NW = NT / 64;
    static constexpr int NQ = RH * RW * (int)(sizeof(T) / 4);
    static constexpr int NBUF = ((NQ + 4 * NT - 1) / (4 * NT)) * 4 * NT;   // image dwords (DMA rounds)
    static_assert(TX % 64 == 0 && NT % 64 == 0 && K >= 2 && K <= HC, "tile shape");
};

// The (sum, count) of this thread's own cells in a frame of the tile's step
// result (row width RW, tile column lx at HC + lx, tile row ly at row ly + R),
// in the thread -> cell order of dp_tile: groups j = 0 .. G-1 at tile row
// ly0 + j R4, cells q = 0 .. 3 of each, valid cells only.  Every kernel forms
// its rho(t) partials in this order, so the single-step, two-pass and K-step
// launches of one tile shape write the same partials bit for bit.
template <typename T, int TX, int TY, int NT>
__device__ __forceinline__ void own_partial(const T* fr, int RW, int HC, int R, const DevTile& tl, double& s,
                                            uint32_t& nlive)
{
    typedef DPShape<T, TX, TY, NT> S;
    const int cg = threadIdx.x % S::CG, ly0 = threadIdx.x / S::CG;
    const int L = tl.L;
    s = 0.0;
    nlive = 0;
    for (int j = 0; j < S::G; ++j) {
        const int ly = ly0 + j * S::R4;
        const int gy = tl.y0 + ly, gx0 = tl.x0 + 4 * cg;
        if (gy >= L || gx0 >= L) continue;
        const int nvalid = min(4, L - gx0);
        const T* row = fr + (ly + R) * RW + HC + 4 * cg;
#pragma unroll 1
        for (int q = 0; q < nvalid; ++q) {
            const T y = row[q];
            s += (double)y;
            nlive += y > (T)0 ? 1u : 0u;
        }
    }
}

template <typename T, int TX, int TY, int NT, int K, int MINW = 1>
__global__ __launch_bounds__(NT, MINW) void dp_stepk_kernel(StepArgs a)
{
    typedef DPKShape<T, TX, TY, NT, K> S;
    constexpr int HC = S::HC, RW = S::RW, CG = S::CG;
    __shared__ __align__(16) uint32_t s_img[S::NBUF];
    // One frame buffer; from phase 2 on, each phase writes into the buffer its
    // source is not in, so the image's buffer serves as the second frame.
    __shared__ __align__(16) T s_fr[S::FH * RW];
    static_assert(S::FH * RW * (int)sizeof(T) <= S::NBUF * 4, "a frame fits the image buffer");
    __shared__ uint32_t s_flag[NT / 64];

    const int64_t counter = a.state->counter + a.k;            // step k
    const int tile = blockIdx.x;
    if (tile >= a.ntiles) return;
    const DevTile tl = a.tiles[tile];
    const DevInst di = a.inst[tl.inst];
    const int L = tl.L;
    tile_to_lds<T, TX, S::RH, NT, K>((const T*)a.in + tl.cell_off, L, tl.x0, tl.y0, s_img);
    __builtin_amdgcn_s_waitcnt(0);   // this thread's copies have landed
    __syncthreads();                 // ... everyone's
    if (image_all_zero<NT, S::NQ>(s_img, s_flag)) {   // absorbing for all K steps
        int64_t slots[K];
#pragma unroll
        for (int j = 0; j < K; ++j) slots[j] = (counter + j) & (int64_t)(a.nslots - 1);
        zero_tile<T, TX, TY, NT>(a, tile, tl, (T*)a.out + tl.cell_off, slots, K);
        return;
    }

    const int tid = threadIdx.x;
    const T* src = (const T*)s_img;
    int srad = K;                                              // radius of the source frame
#pragma unroll 1
    for (int j = 1; j < K; ++j) {
        // ---- phase j: step k + j - 1 on tile rows and columns -R .. T + R - 1 ----
        const int R = K - j;
        const DPConst c = ((const DPConst*)a.consts)[(int64_t)(a.k + j - 1) * a.n_inst + tl.inst];
        T* dstf = (j & 1) ? s_fr : (T*)s_img;                   // never the buffer `src` is in
        const int ngroups = (TY + 2 * R) * CG;
        for (int g = tid; g < ngroups; g += NT) {
            const int ly = g / CG - R, cg = g % CG;
            const int r = (ly + srad) * RW + HC + 4 * cg;          // source index of the group's first cell
            const int m = (ly + R) * RW + HC + 4 * cg;             // destination index
            uint32_t bits = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                bits |= as_bits(src[r + q]) | as_bits(src[r - RW + q]) | as_bits(src[r + RW + q]);
            bits |= as_bits(src[r - 1]) | as_bits(src[r + 4]);
            Quad<T> yq;
            yq.v[0] = yq.v[1] = yq.v[2] = yq.v[3] = (T)0;
            if (bits != 0) {
                const int gy = wrap(tl.y0 + ly, L);
                const bool own_row = ly >= 0 && ly < TY && tl.y0 + ly < L;
#pragma unroll 1
                for (int q = 0; q < 4; ++q) {
                    const int rq = r + q;
                    const int lx = 4 * cg + q;
                    const uint32_t cell = (uint32_t)gy * (uint32_t)L + (uint32_t)wrap(tl.x0 + lx, L);
                    int64_t n;
                    bool bad;
                    const T y = cell_update<T, 0>(c, di, cell, src[rq], src[rq - RW], src[rq + RW], src[rq - 1],
                                                  src[rq + 1], n, bad);
                    if (bad && own_row && tl.x0 + lx < L) record_fault(a.state, tl.inst, cell, counter + j - 1, j - 1);
                    yq.v[0] = q == 0 ? y : yq.v[0];
                    yq.v[1] = q == 1 ? y : yq.v[1];
                    yq.v[2] = q == 2 ? y : yq.v[2];
                    yq.v[3] = q == 3 ? y : yq.v[3];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) dstf[m + q] = yq.v[q];
        }
        // the ring's columns -R .. -1 and TX .. TX + R - 1, one cell per thread
        const int nedge = 2 * R * (TY + 2 * R);
        for (int e = tid; e < nedge; e += NT) {
            const int ly = e / (2 * R) - R, cc = e % (2 * R);
            const int lx = cc < R ? cc - R : TX + (cc - R);
            const int rq = (ly + srad) * RW + HC + lx;
            const uint32_t cell = (uint32_t)wrap(tl.y0 + ly, L) * (uint32_t)L + (uint32_t)wrap(tl.x0 + lx, L);
            int64_t n;
            bool bad;
            dstf[(ly + R) * RW + HC + lx] = cell_update<T, 0>(c, di, cell, src[rq], src[rq - RW], src[rq + RW],
                                                              src[rq - 1], src[rq + 1], n, bad);
        }
        __syncthreads();
        {   // this step's partials from the finished frame, in the single-step order
            double sum;
            uint32_t nlive;
            own_partial<T, TX, TY, NT>(dstf, RW, HC, R, tl, sum, nlive);
            step_partial<NT>(a, tile, counter + j - 1, sum, nlive);
        }
        src = dstf;
        srad = R;
    }

    // ---- last phase: step k + K - 1 on the tile, from the last frame (radius 1) ----
    const DPConst c = ((const DPConst*)a.consts)[(int64_t)(a.k + K - 1) * a.n_inst + tl.inst];
    double sum = 0.0;
    uint32_t nlive = 0;
    T* fout = (T*)a.out + tl.cell_off;
    const bool vec = (L % 4 == 0) && (tl.cell_off % 4 == 0);
    for (int g = tid; g < TY * CG; g += NT) {
        const int ly = g / CG, cg = g % CG;
        const int gy = tl.y0 + ly, gx0 = tl.x0 + 4 * cg;
        if (gy >= L || gx0 >= L) continue;
        const int r = (ly + 1) * RW + HC + 4 * cg;
        const uint32_t cell0 = (uint32_t)gy * (uint32_t)L + (uint32_t)gx0;
        const bool full = gx0 + 4 <= L;
        uint32_t bits = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bits |= as_bits(src[r + q]) | as_bits(src[r - RW + q]) | as_bits(src[r + RW + q]);
        bits |= as_bits(src[r - 1]) | as_bits(src[r + 4]);
        const int nvalid = min(4, L - gx0);
        Quad<T> yq;
        yq.v[0] = yq.v[1] = yq.v[2] = yq.v[3] = (T)0;
        if (bits == 0 && full) {
            if (a.draws) {
#pragma unroll
                for (int q = 0; q < 4; ++q) a.draws[tl.cell_off + cell0 + q] = -1;
            }
        } else {
#pragma unroll 1
            for (int q = 0; q < nvalid; ++q) {
                const int rq = r + q;
                const uint32_t cell = cell0 + q;
                int64_t n;
                bool bad;
                const T y = cell_update<T, 0>(c, di, cell, src[rq], src[rq - RW], src[rq + RW], src[rq - 1],
                                              src[rq + 1], n, bad);
                if (bad) record_fault(a.state, tl.inst, cell, counter + K - 1, K - 1);
                if (a.draws) a.draws[tl.cell_off + cell] = n;
                sum += (double)y;
                nlive += y > (T)0 ? 1u : 0u;
                yq.v[0] = q == 0 ? y : yq.v[0];
                yq.v[1] = q == 1 ? y : yq.v[1];
                yq.v[2] = q == 2 ? y : yq.v[2];
                yq.v[3] = q == 3 ? y : yq.v[3];
            }
        }
        store_quad(fout + (int64_t)gy * L, gx0, nvalid, vec, yq);
    }
    step_partial<NT>(a, tile, counter + K - 1, sum, nlive);
}

// ---------------------------------------------------------------------------
// Rietkerk (SPB = 1): Dornic on P, then RK4 of f_2Dor on (P*, W, O), checks,
// flush, rho(t) partials -- rietkerk_bjork_FKE.cpp:4154-4465, :2363-2444.
// ---------------------------------------------------------------------------
template <typename T>
struct RKConst {
    DornicConst k;          // P Dornic constants (beta = -d - 4 D_P/dx^2)
    T R, cgmax, gmax, alpha, rW, K1, K2, K2W0, dw, dO, dt, dt2, dt6, eps;
    T aK2W0;                // alpha k2 W0 (the FP32 form of f_2Dor)
    int compat;
    int64_t step;
};

// FP64 constants as tabled by prep_consts, with their FP32 roundings beside them
// (the scalar unit has no FP64 -> FP32 conversion: formed once per (step,
// instance) here rather than by every wave); rk_cast gives the stepping type's copy.
struct RKConstD {
    DornicConst k;
    double R, cgmax, gmax, alpha, rW, K1, K2, K2W0, dw, dO, dt, dt2, dt6, eps;
    int compat, pad;
    int64_t step;
    float fR, fcgmax, fgmax, falpha, frW, fK1, fK2, fK2W0, fdw, fdO, fdt, fdt2, fdt6, feps, faK2W0;
};

__device__ __forceinline__ RKConstD rk_constants(const tm_rietkerk_params& p, int64_t step)
{
#pragma clang fp contract(off)
    RKConstD c;
    const double dx2 = p.dx * p.dx;
    c.k = dornic_constants(-p.d, p.D[0], p.sigma, p.dt, p.dx);
    c.R = p.R;
    c.cgmax = p.c * p.gmax;
    c.gmax = p.gmax;
    c.alpha = p.alpha;
    c.rW = p.rW;
    c.K1 = p.K[1];
    c.K2 = p.K[2];
    c.K2W0 = p.K[2] * p.W0;
    c.dw = p.D[1] / (dx2);
    c.dO = p.D[2] / (dx2);
    c.dt = p.dt;
    c.dt2 = p.dt / 2.0;
    c.dt6 = p.dt / 6.0;
    c.eps = p.flush_eps;
    c.compat = p.sampler == TM_SAMPLER_COMPAT;
    c.pad = 0;
    c.step = step;
    c.fR = (float)c.R; c.fcgmax = (float)c.cgmax; c.fgmax = (float)c.gmax; c.falpha = (float)c.alpha;
    c.frW = (float)c.rW; c.fK1 = (float)c.K1; c.fK2 = (float)c.K2; c.fK2W0 = (float)c.K2W0;
    c.fdw = (float)c.dw; c.fdO = (float)c.dO; c.fdt = (float)c.dt; c.fdt2 = (float)c.dt2;
    c.fdt6 = (float)c.dt6; c.feps = (float)c.eps;
    c.faK2W0 = (float)(c.alpha * c.K2W0);
    return c;
}

template <typename T>
__device__ __forceinline__ RKConst<T> rk_cast(const RKConstD& d)
{
    RKConst<T> c;
    c.k = d.k;
    c.R = d.R; c.cgmax = d.cgmax; c.gmax = d.gmax; c.alpha = d.alpha; c.rW = d.rW;
    c.K1 = d.K1; c.K2 = d.K2; c.K2W0 = d.K2W0; c.dw = d.dw; c.dO = d.dO;
    c.dt = d.dt; c.dt2 = d.dt2; c.dt6 = d.dt6; c.eps = d.eps;
    c.aK2W0 = d.alpha * d.K2W0;
    c.compat = d.compat;
    c.step = d.step;
    return c;
}

template <>
__device__ __forceinline__ RKConst<float> rk_cast<float>(const RKConstD& d)
{
    RKConst<float> c;
    c.k = d.k;
    c.R = d.fR; c.cgmax = d.fcgmax; c.gmax = d.fgmax; c.alpha = d.falpha; c.rW = d.frW;
    c.K1 = d.fK1; c.K2 = d.fK2; c.K2W0 = d.fK2W0; c.dw = d.fdw; c.dO = d.fdO;
    c.dt = d.fdt; c.dt2 = d.fdt2; c.dt6 = d.fdt6; c.eps = d.feps;
    c.aK2W0 = d.faK2W0;
    c.compat = d.compat;
    c.step = d.step;
    return c;
}

// f_2Dor at one cell (rietkerk_bjork_FKE.cpp:2380-2386), with nW / nO the sums
// up + down + left + right of the W / O neighbours.
template <typename T>
__device__ __forceinline__ void rk_rhs(const RKConst<T>& c, T P, T W, T O, T nW, T nO, T& fP, T& fW,
                                       T& fO)
{
#pragma clang fp contract(off)
    fP = Arith<T>::div(c.cgmax * W * P, W + c.K1);
    const T A = Arith<T>::div(c.alpha * (P + c.K2W0), P + c.K2) * O;
    fW = A - Arith<T>::div(c.gmax * W * P, W + c.K1) - c.rW * W + c.dw * (nW - (T)4 * W);
    fO = c.R - A + c.dO * (nO - (T)4 * O);
}

// FP32: the same terms with the hardware reciprocal, the quotient t = W P / (W + k1)
// shared by the two uptake terms, alpha (P + k2 W0) as one fused multiply-add
// and explicit fused multiply-adds elsewhere (no rounding order to keep: the
// FP32 path is checked against FP64 statistically, and every FP32 kernel uses
// this one form).  The diffusion terms keep d (n - 4 x), whose difference is
// formed before the large factor d_O multiplies it.
template <>
__device__ __forceinline__ void rk_rhs<float>(const RKConst<float>& c, float P, float W, float O, float nW,
                                              float nO, float& fP, float& fW, float& fO)
{
    const float t = (W * P) * __builtin_amdgcn_rcpf(W + c.K1);
    fP = c.cgmax * t;
    const float A = __builtin_fmaf(c.alpha, P, c.aK2W0) * __builtin_amdgcn_rcpf(P + c.K2) * O;
    fW = __builtin_fmaf(c.dw, __builtin_fmaf(-4.0f, W, nW), __builtin_fmaf(-c.rW, W, __builtin_fmaf(-c.gmax, t, A)));
    fO = __builtin_fmaf(c.dO, __builtin_fmaf(-4.0f, O, nO), c.R - A);
}

// Ring of radius r around a TX x TY tile: index k -> tile-relative (ly, lx).
// constexpr: also evaluated at compile time for the ring offset tables below.
template <int TX, int TY, int R>
constexpr void ring_cell(int k, int& ly, int& lx)
{
    constexpr int W = TX + 2 * R;
    if (k < R * W) {
        ly = -R + k / W;
        lx = -R + k % W;
    } else if (k < 2 * R * W) {
        k -= R * W;
        ly = TY + k / W;
        lx = -R + k % W;
    } else {
        k -= 2 * R * W;
        ly = k / (2 * R);
        const int cc = k % (2 * R);
        lx = cc < R ? cc - R : TX + (cc - R);
    }
}

// (W, O) of one cell, stored together so that one 2-element LDS read fetches both.
template <typename T>
struct alignas(2 * sizeof(T)) WOPair {
    T w, o;
};

template <int TX, int TY, int R>
struct Ring {
    static constexpr int N = (TX + 2 * R) * (TY + 2 * R) - TX * TY;
};

// The ring cells each thread of an NT-thread workgroup of rk_step_kernel owns,
// as LDS offsets formed at compile time.  Entry e of thread t's row is
// (i4 << 16) | i3 for its e-th ring cell -- the radius-3 ring's entries t + j NT
// (j < NR3), then the radius-2 ring's and the radius-1 ring's -- with i4 / i3
// the cell's index in the radius-4 / radius-3 frame, and 0xFFFF0000 past a
// ring's end (i4 reads as -1 after an arithmetic shift).  A thread reads its row
// with one 16-byte load instead of forming the offsets from divisions and
// selects (about a hundred instructions per thread).
template <int TX, int TY, int NT>
struct RingOffsets {
    static constexpr int NR3 = (Ring<TX, TY, 3>::N + NT - 1) / NT;
    static constexpr int NR2 = (Ring<TX, TY, 2>::N + NT - 1) / NT;
    static constexpr int NR1 = (Ring<TX, TY, 1>::N + NT - 1) / NT;
    static constexpr int NE = NR3 + NR2 + NR1;
    static constexpr int NE4 = (NE + 3) / 4 * 4;
    static_assert((TY + 8) * (TX + 8) < 0x8000, "frame indices fit 15 bits");
    struct alignas(16) Row { uint32_t v[NE4]; };
    Row row[NT];
};

template <int TX, int TY, int R, int NT, int NR>
constexpr void ring_offsets_fill(RingOffsets<TX, TY, NT>& t, int e0)
{
    for (int tid = 0; tid < NT; ++tid) {
        for (int j = 0; j < NR; ++j) {
            const int k = tid + j * NT;
            uint32_t v = 0xFFFF0000u;
            if (k < Ring<TX, TY, R>::N) {
                int ly = 0, lx = 0;
                ring_cell<TX, TY, R>(k, ly, lx);
                const int i4 = (ly + 4) * (TX + 8) + (lx + 4), i3 = (ly + 3) * (TX + 6) + (lx + 3);
                v = ((uint32_t)i4 << 16) | (uint32_t)i3;
            }
            t.row[tid].v[e0 + j] = v;
        }
    }
}

template <int TX, int TY, int NT>
constexpr RingOffsets<TX, TY, NT> make_ring_offsets()
{
    RingOffsets<TX, TY, NT> t{};
    using RO = RingOffsets<TX, TY, NT>;
    ring_offsets_fill<TX, TY, 3, NT, RO::NR3>(t, 0);
    ring_offsets_fill<TX, TY, 2, NT, RO::NR2>(t, RO::NR3);
    ring_offsets_fill<TX, TY, 1, NT, RO::NR1>(t, RO::NR3 + RO::NR2);
    return t;
}

template <int TX, int TY, int NT>
__device__ const RingOffsets<TX, TY, NT> kRingOffsets = make_ring_offsets<TX, TY, NT>();

// Fixed-order block reduction of three (FP64 sum, count) pairs, each wave's
// sums through wave_sum4 and the counts already wave totals (ballots); threads
// 0..5 add the NT / 64 wave values in wave order and write (sum, count)
// partials.  Every lane of the block must be active.
template <int NT>
__device__ __forceinline__ void block_partials3(const double (&sum)[3], const uint32_t (&wcnt)[3], double2* dst,
                                                double* red)
{
    constexpr int NW = NT / 64;
    const double t = wave_sum4(sum[0], sum[1], sum[2], 0.0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if ((lane & 15) == 0 && lane < 48) red[(2 * (lane >> 4)) * NW + w] = t;
    if (lane == 0) {
#pragma unroll
        for (int f = 0; f < 3; ++f) red[(2 * f + 1) * NW + w] = (double)wcnt[f];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NW; ++i) s += red[threadIdx.x * NW + i];
        ((double*)dst)[threadIdx.x] = s;
    }
}

// SPLIT = false: the whole step in one kernel, the Dornic P* computed on the
// radius-3 frame of every tile (ring cells are sampled by both neighbouring
// tiles).  SPLIT = true: P* comes from the Dornic kernel (dp_step_kernel
// MODE 1) through a.aux, and this kernel runs the RK4 part only.
// Each thread owns NOWN tile cells and a few cells of each ring (radius 3, 2,
// 1); their LDS offsets in the radius-4 and radius-3 frames are computed once
// at entry and reused by every stage.
template <typename T, int TX, int TY, int NT, bool SPLIT>
__global__ __launch_bounds__(NT) void rk_step_kernel(StepArgs a)
{
    constexpr int W4 = TX + 8, H4 = TY + 8;        // radius-4 frame (loaded fields)
    constexpr int W3 = TX + 6, H3 = TY + 6;        // radius-3 frame (P*, stage states)
    constexpr int NOWN = TX * TY / NT;             // owned cells per thread
    static_assert((TX * TY) % NT == 0, "tile shape");
    constexpr int NR3 = (Ring<TX, TY, 3>::N + NT - 1) / NT;
    constexpr int NR2 = (Ring<TX, TY, 2>::N + NT - 1) / NT;
    constexpr int NR1 = (Ring<TX, TY, 1>::N + NT - 1) / NT;
    static_assert(LdsFrame<T, 4>::HC == 4, "radius-4 frames are W4 = TX + 8 wide");
    constexpr int NBUF4 = ((H4 * W4 * (int)sizeof(T) / 4 + 4 * NT - 1) / (4 * NT)) * 4 * NT;
    // One LDS object, so that its members sit at fixed addresses: the (W, O)
    // stage states first, at address 0, where a cell and its four neighbours are
    // read at constant offsets (all under 256 pairs) from one base -- two paired
    // reads and one single in FP32, no address arithmetic per read.
    struct Lds {
        WOPair<T> wo[H3 * W3];              // stage states W, O on the radius-3 frame
        T p[H3 * W3];                       // stage state P
        T ps[SPLIT ? 1 : H3 * W3];          // fused step: P* on the radius-3 frame
        alignas(16) uint32_t fr[3][NBUF4];  // radius-4 frames of P (P* for the split step), W, O (DMA)
        double red[6 * (NT / 64)];
    };
    __shared__ Lds sm;
    T* const sP0 = (T*)sm.fr[0];
    T* const sW0 = (T*)sm.fr[1];
    T* const sO0 = (T*)sm.fr[2];

    const int64_t counter = a.state->counter + a.k;
    const int64_t slot = counter & (int64_t)(a.nslots - 1);   // nslots: a power of two
    const int tid = threadIdx.x;
    const T* in = (const T*)a.in;
    T* out = (T*)a.out;
    const RKConstD* consts = (const RKConstD*)a.consts + (int64_t)a.k * a.n_inst;

    auto i4 = [](int ly, int lx) { return (ly + 4) * W4 + (lx + 4); };
    auto i3 = [](int ly, int lx) { return (ly + 3) * W3 + (lx + 3); };
    auto own = [&](int j, int& ly, int& lx) {
        const int k = tid + j * NT;
        ly = k / TX;
        lx = k % TX;
    };

    if ((int)blockIdx.x >= a.ntiles) return;
    // one tile per workgroup; its descriptor, instance and constants are read
    // before the kernel writes anything (scalar loads), then the three frames
    // are staged by DMA
    const int tile = tile_of(blockIdx.x, a.ntiles, a.xcd);
    const DevTile tl = a.tiles[tile];
    const DevInst di = a.inst[tl.inst];
    const RKConstD cd = consts[tl.inst];
    const int L = tl.L;
    const int64_t N = (int64_t)L * L;
    const T* fin = in + 3 * tl.cell_off;
    T* fout = out + 3 * tl.cell_off;
    const T* p_src = SPLIT ? (const T*)a.aux + tl.cell_off : fin;
    {
        const T* const src[3] = {p_src, fin + N, fin + 2 * N};
        uint32_t* const dst[3] = {sm.fr[0], sm.fr[1], sm.fr[2]};
        tiles_to_lds<T, TX, H4, NT, 4, 3>(src, L, tl.x0, tl.y0, dst);
    }
    const RKConst<T> c = rk_cast<T>(cd);

    // LDS offsets (radius-4 frame, radius-3 frame) of this thread's cells; ring
    // entries past the ring's end hold -1 (from the compile-time table)
    int o4[NOWN], o3[NOWN], q4[NR3], q3[NR3], p4[NR2], p3[NR2], m4[NR1], m3[NR1];
#pragma unroll
    for (int j = 0; j < NOWN; ++j) {
        int ly, lx;
        own(j, ly, lx);
        o4[j] = i4(ly, lx);
        o3[j] = i3(ly, lx);
    }
    {
        using RO = RingOffsets<TX, TY, NT>;
        static_assert(RO::NR3 == NR3 && RO::NR2 == NR2 && RO::NR1 == NR1, "ring table shape");
        const typename RO::Row row = kRingOffsets<TX, TY, NT>.row[tid];
        auto hi = [](uint32_t v) { return (int)v >> 16; };
        auto lo = [](uint32_t v) { return (int)(v & 0xFFFFu); };
#pragma unroll
        for (int j = 0; j < NR3; ++j) { q4[j] = hi(row.v[j]); q3[j] = lo(row.v[j]); }
#pragma unroll
        for (int j = 0; j < NR2; ++j) { p4[j] = hi(row.v[NR3 + j]); p3[j] = lo(row.v[NR3 + j]); }
#pragma unroll
        for (int j = 0; j < NR1; ++j) {
            m4[j] = hi(row.v[NR3 + NR2 + j]);
            m3[j] = lo(row.v[NR3 + NR2 + j]);
        }
    }
    auto ps = [&](int r4, int r3) { return SPLIT ? sP0[r4] : sm.ps[r3]; };   // P*

    __builtin_amdgcn_s_waitcnt(0);   // this thread's copies have landed
    __syncthreads();                 // ... everyone's

    // ---- stage 0 (fused step): Dornic P* on the radius-3 frame (FKE:4154-4211) ----
    if constexpr (!SPLIT) {
        auto dornic_at = [&](int ly, int lx, int64_t* n_out) {
            const int r = i4(ly, lx);
            const int gy = wrap(tl.y0 + ly, L), gx = wrap(tl.x0 + lx, L);
            int64_t n = -1;
            const T y = DornicLive<T>::step(di.seed, di.inst_id, (uint32_t)((int64_t)gy * L + gx), c.step, sP0[r],
                                            sP0[r - W4], sP0[r + W4], sP0[r - 1], sP0[r + 1], c.k, c.compat != 0, n);
            *n_out = n;
            sm.ps[i3(ly, lx)] = y;
        };
        for (int j = 0; j < NOWN; ++j) {
            int ly, lx;
            own(j, ly, lx);
            int64_t n;
            dornic_at(ly, lx, &n);
            const int gy = tl.y0 + ly, gx = tl.x0 + lx;
            if (a.draws && gy < L && gx < L) a.draws[di.cell_off + (int64_t)gy * L + gx] = n;
        }
        for (int j = 0; j < NR3; ++j) {
            const int k = tid + j * NT;
            if (k < Ring<TX, TY, 3>::N) {
                int ly, lx;
                ring_cell<TX, TY, 3>(k, ly, lx);
                int64_t n;
                dornic_at(ly, lx, &n);
            }
        }
        __syncthreads();
    }

    // ---- stage 1: K1 from (P*, W, O) on the radius-3 frame ----
    // The tile's own y values are read here once and kept in registers.
    T accP[NOWN], accW[NOWN], accO[NOWN], yP[NOWN], yW[NOWN], yO[NOWN];
    auto put = [&](int r3, T P, T W, T O) {
        sm.wo[r3] = WOPair<T>{W, O};
        sm.p[r3] = P;
    };
    auto stage1 = [&](int r4, int r3, T& P, T& W, T& O, T& kP, T& kW, T& kO) {
        P = ps(r4, r3);
        W = sW0[r4];
        O = sO0[r4];
        const T nW = sW0[r4 - W4] + sW0[r4 + W4] + sW0[r4 - 1] + sW0[r4 + 1];
        const T nO = sO0[r4 - W4] + sO0[r4 + W4] + sO0[r4 - 1] + sO0[r4 + 1];
        rk_rhs<T>(c, P, W, O, nW, nO, kP, kW, kO);
        put(r3, P + c.dt2 * kP, W + c.dt2 * kW, O + c.dt2 * kO);
    };
#pragma unroll
    for (int j = 0; j < NOWN; ++j) stage1(o4[j], o3[j], yP[j], yW[j], yO[j], accP[j], accW[j], accO[j]);
#pragma unroll
    for (int j = 0; j < NR3; ++j) {
        if (q4[j] >= 0) {
            T P, W, O, kP, kW, kO;
            stage1(q4[j], q3[j], P, W, O, kP, kW, kO);
        }
    }
    __syncthreads();

    // ---- stages 2 and 3: K from the stage state, new state held in registers
    //      until every thread has read the old one.  The cells evaluated here lie
    //      within radius 2, so b = r - W3 - 1 >= 0 and the reads are at b + 1 (up),
    //      b + W3 (left), b + W3 + 1 (the cell), b + W3 + 2 (right), b + 2 W3 + 1
    //      (down); sums in the order up, down, left, right ----
    auto kstage = [&](int r, T& kP, T& kW, T& kO) {
        const WOPair<T>* bp = sm.wo + (r - W3 - 1);
        const WOPair<T> up = bp[1], lf = bp[W3], ce = bp[W3 + 1], rt = bp[W3 + 2], dn = bp[2 * W3 + 1];
        const T P = sm.p[r];
        const T nW = up.w + dn.w + lf.w + rt.w;
        const T nO = up.o + dn.o + lf.o + rt.o;
        rk_rhs<T>(c, P, ce.w, ce.o, nW, nO, kP, kW, kO);
    };
    {   // stage 2: K2 on radius 2, ys3 = y + dt/2 K2
        T nP[NOWN], nW[NOWN], nO[NOWN];
        T rP[NR2], rW[NR2], rO[NR2];
#pragma unroll
        for (int j = 0; j < NOWN; ++j) {
            T kP, kW, kO;
            kstage(o3[j], kP, kW, kO);
            accP[j] = accP[j] + (T)2.0 * kP;
            accW[j] = accW[j] + (T)2.0 * kW;
            accO[j] = accO[j] + (T)2.0 * kO;
            nP[j] = yP[j] + c.dt2 * kP;
            nW[j] = yW[j] + c.dt2 * kW;
            nO[j] = yO[j] + c.dt2 * kO;
        }
#pragma unroll
        for (int j = 0; j < NR2; ++j) {
            if (p4[j] >= 0) {
                T kP, kW, kO;
                kstage(p3[j], kP, kW, kO);
                const T P = ps(p4[j], p3[j]), W = sW0[p4[j]], O = sO0[p4[j]];
                rP[j] = P + c.dt2 * kP;
                rW[j] = W + c.dt2 * kW;
                rO[j] = O + c.dt2 * kO;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NOWN; ++j) put(o3[j], nP[j], nW[j], nO[j]);
#pragma unroll
        for (int j = 0; j < NR2; ++j) {
            if (p4[j] >= 0) put(p3[j], rP[j], rW[j], rO[j]);
        }
        __syncthreads();
    }
    {   // stage 3: K3 on radius 1, ys4 = y + dt K3
        T nP[NOWN], nW[NOWN], nO[NOWN];
        T rP[NR1], rW[NR1], rO[NR1];
#pragma unroll
        for (int j = 0; j < NOWN; ++j) {
            T kP, kW, kO;
            kstage(o3[j], kP, kW, kO);
            accP[j] = accP[j] + (T)2.0 * kP;
            accW[j] = accW[j] + (T)2.0 * kW;
            accO[j] = accO[j] + (T)2.0 * kO;
            nP[j] = yP[j] + c.dt * kP;
            nW[j] = yW[j] + c.dt * kW;
            nO[j] = yO[j] + c.dt * kO;
        }
#pragma unroll
        for (int j = 0; j < NR1; ++j) {
            if (m4[j] >= 0) {
                T kP, kW, kO;
                kstage(m3[j], kP, kW, kO);
                const T P = ps(m4[j], m3[j]), W = sW0[m4[j]], O = sO0[m4[j]];
                rP[j] = P + c.dt * kP;
                rW[j] = W + c.dt * kW;
                rO[j] = O + c.dt * kO;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NOWN; ++j) put(o3[j], nP[j], nW[j], nO[j]);
#pragma unroll
        for (int j = 0; j < NR1; ++j) {
            if (m4[j] >= 0) put(m3[j], rP[j], rW[j], rO[j]);
        }
        __syncthreads();
    }

    // ---- stage 4: K4 on the tile, combine, checks, flush, write ----
    double sum[3] = {0.0, 0.0, 0.0};
    uint32_t wcnt[3] = {0u, 0u, 0u};           // wave totals: ballots of the occupied cells
#pragma unroll
    for (int j = 0; j < NOWN; ++j) {
        int ly, lx;
        own(j, ly, lx);
        T kP, kW, kO;
        kstage(o3[j], kP, kW, kO);
        T y[3];
        y[0] = yP[j] + c.dt6 * (accP[j] + kP);
        y[1] = yW[j] + c.dt6 * (accW[j] + kW);
        y[2] = yO[j] + c.dt6 * (accO[j] + kO);
        const int gy = tl.y0 + ly, gx = tl.x0 + lx;
        const bool in = gy < L && gx < L;
        const int64_t cell = (int64_t)gy * L + gx;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            // ok fails for negative, NaN or infinite values (finite >= 0 values pass
            // both comparisons); a flushed value (0 <= v < eps) is ok.  Formed in
            // every lane, so the ballot below needs no value carried out of a branch.
            const bool ok = y[f] >= (T)0 && y[f] <= Arith<T>::largest();
            const T v = ok && y[f] < c.eps ? (T)0 : y[f];
            if (in) {
                if (!ok) record_fault(a.state, tl.inst, f * N + cell, counter);
                fout[f * N + cell] = v;
            }
            sum[f] += (double)(in ? v : (T)0);       // + 0 for cells past the lattice
            wcnt[f] += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(in && v > (T)0));
        }
    }
    double2* dst = a.partials + (slot * a.ntiles + tile) * 3;
    block_partials3<NT>(sum, wcnt, dst, sm.red);
}

// ---------------------------------------------------------------------------
// Block prologue and epilogue.
// prep_consts: table[k][i] = constants of instance i at step counter + k.
// bump_counter: counter += n (one thread; runs after the block's rows fold).
// ---------------------------------------------------------------------------
// Vegetation + consumers (multiSPDP.cpp with SPB = 2, 3): the per-step
// constants of the vegetation's Dornic step (its linear rate a ramped as in
// the DP model); the consumers' constants depend on each cell's perception and
// draw and are formed in the step kernel.
struct DPCConst {
    DornicConst k;
    double eps, neg;
    int compat;
    int64_t step;
};

__device__ __forceinline__ DPCConst dpc_constants(const tm_dpc_params& p, int64_t step)
{
#pragma clang fp contract(off)
    double a = p.a;
    if (p.ramp_steps > 0) {
        const int64_t m = step < p.ramp_steps ? step : p.ramp_steps;
        a = p.a + (p.a_end - p.a) * ((double)m / (double)p.ramp_steps);
    }
    DPCConst c;
    c.k = dornic_constants(a, p.D[0], p.sigma[0], p.dt, p.dx);
    c.eps = p.flush_eps;
    c.neg = p.flush_neg;
    c.compat = p.sampler == TM_SAMPLER_COMPAT;
    c.step = step;
    return c;
}

// The vegetation's Dornic step of the Rietkerk consumer models as a Rietkerk
// parameter set: d = M[0], D_P, sigma[0], dt, dx and the sampler (FKE:3847-3849,
// :4150-4211); the RK4 fields are not used by that step.
__host__ __device__ inline tm_rietkerk_params fke_veg_params(const tm_fke_params& p)
{
    tm_rietkerk_params r{};
    r.R = p.a;
    r.c = p.c;
    r.gmax = p.gmax;
    r.alpha = p.alpha;
    r.rW = p.rW;
    r.W0 = p.W0;
    r.d = p.M[0];
    r.D[0] = p.D_P;
    r.D[1] = p.D_W;
    r.D[2] = p.D_O;
    r.K[0] = p.K[0];
    r.K[1] = p.K[1];
    r.K[2] = p.K[2];
    r.sigma = p.sigma[0];
    r.dt = p.dt;
    r.dx = p.dx;
    r.flush_eps = p.flush_eps;
    r.sampler = p.sampler;
    return r;
}

template <int MODEL>
__global__ __launch_bounds__(256) void prep_consts_kernel(const DevInst* inst, const void* params,
                                                          const DevState* state, int n_inst, int nsteps,
                                                          void* table)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)n_inst * nsteps) return;
    const int k = (int)(idx / n_inst), i = (int)(idx % n_inst);
    const DevInst di = inst[i];
    const int64_t step = di.step0 + state->counter + k;
    if (MODEL == TM_MODEL_DP)
        ((DPConst*)table)[idx] = dp_constants(((const tm_dp_params*)params)[di.pidx], step);
    else if (MODEL == TM_MODEL_DP2)   // both consumer models
        ((DPCConst*)table)[idx] = dpc_constants(((const tm_dpc_params*)params)[di.pidx], step);
    else if (MODEL == TM_MODEL_FKE2)  // both Rietkerk consumer models: the vegetation's Dornic step
        ((RKConstD*)table)[idx] = rk_constants(fke_veg_params(((const tm_fke_params*)params)[di.pidx]), step);
    else
        ((RKConstD*)table)[idx] = rk_constants(((const tm_rietkerk_params*)params)[di.pidx], step);
}

#ifdef TM_DEFINE_COMBINE   // defined by exactly one translation unit
__global__ void bump_counter_kernel(DevState* state, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) state->counter += n;
}
#endif

// ---------------------------------------------------------------------------
// Partials of the current fields only (tm_measure).
// ---------------------------------------------------------------------------
template <typename T, int NF>
__global__ __launch_bounds__(256) void measure_kernel(const DevInst* inst, const DevTile* tiles, int ntiles,
                                                      int tx, int ty, const T* fields, double2* partials)
{
    __shared__ double s_red[2 * NF * 4];
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    const DevTile tl = tiles[tile];
    const DevInst di = inst[tl.inst];
    const int L = di.L;
    const int64_t N = (int64_t)L * L;
    const T* f0 = fields + NF * di.cell_off;
    double sum[NF], cnt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) { sum[f] = 0.0; cnt[f] = 0.0; }
    for (int k = threadIdx.x; k < tx * ty; k += 256) {
        const int gy = tl.y0 + k / tx, gx = tl.x0 + k % tx;
        if (gy >= L || gx >= L) continue;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const T v = f0[f * N + (int64_t)gy * L + gx];
            sum[f] += (double)v;
            cnt[f] += v > (T)0 ? 1.0 : 0.0;
        }
    }
    block_partials<256, NF>(sum, cnt, partials + (int64_t)tile * NF, s_red);
}

// ---------------------------------------------------------------------------
// Fold tile partials into rho(t) rows: one workgroup per (instance, step).
// Per-thread Kahan sums over a fixed tile stride, then a fixed pairwise tree
// of compensated (TwoSum) additions -- deterministic run to run.
// rows: [n_inst][rows_cap][2 nf]; row = step counter - run_base.
// ---------------------------------------------------------------------------
struct CombineArgs {
    const DevInst* inst;
    const double2* partials;
    const DevState* state;
    double* rows;
    int32_t ntiles, nslots, nf, nsteps;  // nsteps: steps folded by this launch
    int64_t rows_cap;
    int32_t measure;                     // 1: slot 0, row 0 (tm_measure)
    int32_t pw;                          // partials per tile (waves of the step kernel)
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

#ifdef TM_DEFINE_COMBINE   // defined by exactly one translation unit
__global__ __launch_bounds__(256) void combine_rows_kernel(CombineArgs a)
{
#pragma clang fp contract(off)
    __shared__ double s_s[256], s_e[256], s_n[256];
    const int i = blockIdx.x, j = blockIdx.y;
    const DevInst di = a.inst[i];
    int64_t slot = 0, row = 0;
    if (!a.measure) {
        const int64_t c = a.state->counter + j;   // step counter of this row (before the bump)
        slot = c & (int64_t)(a.nslots - 1);
        row = c - a.state->run_base;
    }
    const double2* p = a.partials + ((slot * a.ntiles + di.tile0) * a.pw) * a.nf;
    const int nsub = di.ntiles * a.pw;
    const int tid = threadIdx.x;
    const double cells = (double)di.L * (double)di.L;
    double* out = a.rows + ((int64_t)i * a.rows_cap + row) * (2 * a.nf);
    // one field at a time: the fold of each field is independent of the others
    for (int f = 0; f < a.nf; ++f) {
        double s = 0.0, e = 0.0, n = 0.0;
        for (int t = tid; t < nsub; t += 256) {
            const double2 v = p[(int64_t)t * a.nf + f];
            const double y = v.x - e;     // Kahan
            const double tt = s + y;
            e = (tt - s) - y;
            s = tt;
            n += v.y;
        }
        s_s[tid] = s;
        s_e[tid] = -e;
        s_n[tid] = n;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) {
                double ss, ee;
                two_sum(s_s[tid], s_s[tid + w], ss, ee);
                s_s[tid] = ss;
                s_e[tid] = s_e[tid] + s_e[tid + w] + ee;
                s_n[tid] += s_n[tid + w];
            }
            __syncthreads();
        }
        if (tid == 0) {
            out[2 * f] = s_n[0];
            out[2 * f + 1] = (s_s[0] + s_e[0]) / cells;
        }
        __syncthreads();   // the next field reuses the arrays
    }
}
#endif

}  // namespace tmk
