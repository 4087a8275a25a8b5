// tm_device.h -- counter-based randomness, FP64 transcendentals and the
// Dornic Poisson-Gamma samplers, shared by the FP64 and FP32 step kernels.
//
// Everything that can move an integer draw (Philox words, the FP64 uniforms,
// exp/log, the Poisson sampler, the FP64 Gamma sampler) uses only IEEE basic
// operations in a fixed order with contraction off, so the FP64 kernels draw
// exactly what a CPU restatement of the same steps draws.
//
// Reference semantics (paths relative to the reference's simulations/):
//   Dornic transition   Rietkerk_FPE/rietkerk_bjork_FKE.cpp:4157-4196
//                       (Percolation_FD/multiSPDP.cpp:2904-2943 for DP)
//   constants           rietkerk_bjork_FKE.cpp:3847-3849, multiSPDP.cpp:2485-2486
// The reference samples with libstdc++ <random> on one std::mt19937_64 stream
// per replicate (rietkerk_bjork_FKE.cpp:3896-3899); here every (seed, instance,
// cell, step) owns an independent Philox4x32-10 stream, so results do not
// depend on how the lattice is tiled or ordered.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TM_HD __host__ __device__ __forceinline__

namespace tmk {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Key = (seed, instance id); counter =
// (cell, step lo, step hi, block).  Initial fields use step words 0xFFFFFFFF.
// ---------------------------------------------------------------------------
struct Philox4 {
    uint32_t v0, v1, v2, v3;
};

// Each round's two three-way xors are one v_bitop3_b32 each (truth table 0x96 =
// a ^ b ^ c, new in gfx950); the compiler does not form them from ^ chains.
__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{c0, c1, c2, c3};
}

// FP32 uniform ((w >> 9) + 1/2) 2^-23 of a word: exact, never 0 or 1.  Formed
// from bits in two instructions: v_alignbit_b32 puts w >> 9 under the exponent
// of 1.0 (the float 1 + (w >> 9) 2^-23), and subtracting 1 - 2^-24 leaves
// (w >> 9) 2^-23 + 2^-24 exactly (Sterbenz: both operands lie in [1/2, 2]).
__device__ __forceinline__ float u32_of(uint32_t w)
{
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, w, 9u)) - 0x1.fffffep-1f;
}

// A per-(cell, step) stream of 32-bit words, four per Philox block.  Words are
// handed out by rotation (no dynamic register indexing).
struct Stream {
    uint32_t k0, k1, c0, c1, c2, blk;
    uint32_t b0, b1, b2, b3;
    int left;

    __device__ Stream(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step)
        : k0(seed), k1(inst), c0(cell), c1((uint32_t)(uint64_t)step),
          c2((uint32_t)((uint64_t)step >> 32)), blk(0), b0(0), b1(0), b2(0), b3(0), left(0)
    {
    }

    __device__ uint32_t word()
    {
        if (left == 0) {
            const Philox4 r = philox4x32_10(c0, c1, c2, blk, k0, k1);
            blk += 1u;
            b0 = r.v0; b1 = r.v1; b2 = r.v2; b3 = r.v3;
            left = 4;
        }
        const uint32_t w = b0;
        b0 = b1; b1 = b2; b2 = b3;
        left -= 1;
        return w;
    }
    // (w + 1/2) 2^-32: exact in double, never 0 or 1.
    __device__ double u64() { return ((double)word() + 0.5) * 0x1p-32; }
    // FP32 uniform ((w >> 9) + 1/2) 2^-23: exact, never 0 or 1.
    __device__ float u32() { return u32_of(word()); }
    // 1 - u from the same kind of word with full FP32 resolution near 0:
    // ((2^32 - 1 - w) + 1/2) 2^-32, in (0, 1] (1 only through rounding at the top),
    // rounded once: the fused multiply-add rounds the exact (x + 1/2) 2^-32.
    __device__ float v32() { return __builtin_fmaf((float)(~word()), 0x1p-32f, 0x1p-33f); }
};

// The first four words of a cell's stream, computed at once: block 0 of the
// same Philox stream, handed out in the same order, with no refill check.  For
// callers that provably draw at most four words (the first pass of the
// deferring FP32 step); a fifth word would repeat the fourth.
struct Stream4 {
    uint32_t b0, b1, b2, b3;

    __device__ Stream4(uint32_t seed, uint32_t inst, uint32_t cell, int64_t step)
    {
        const Philox4 r = philox4x32_10(cell, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), 0u, seed,
                                        inst);
        b0 = r.v0; b1 = r.v1; b2 = r.v2; b3 = r.v3;
    }
    __device__ uint32_t word()
    {
        const uint32_t w = b0;
        b0 = b1; b1 = b2; b2 = b3;
        return w;
    }
    __device__ double u64() { return ((double)word() + 0.5) * 0x1p-32; }
    __device__ float u32() { return u32_of(word()); }
    __device__ float v32() { return __builtin_fmaf((float)(~word()), 0x1p-32f, 0x1p-33f); }
};

// ---------------------------------------------------------------------------
// FP64 exp / log written from the mathematics: Cody-Waite reduction by ln 2 and
// a degree-13 Taylor polynomial; log via m = 2^-e x in [sqrt(1/2), sqrt 2) and
// the atanh series of s = (m-1)/(m+1) to s^23.  (<= 1 and <= 2 ulp.)
// ---------------------------------------------------------------------------
TM_HD double bits_to_double(uint64_t b)
{
    union { uint64_t u; double d; } x;
    x.u = b;
    return x.d;
}
TM_HD uint64_t double_to_bits(double d)
{
    union { uint64_t u; double d; } x;
    x.d = d;
    return x.u;
}

TM_HD double dexp(double x)
{
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x > 0x1.62e42fefa39efp+9) return __builtin_inf();
    if (x < -745.13321910194122) return 0.0;
    const double kd = __builtin_floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;
    p = p * r + 0x1.1eed8eff8d898p-29;
    p = p * r + 0x1.ae64567f544e4p-26;
    p = p * r + 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    if (k < -1020) { p = p * 0x1p-60; k += 60; }
    if (k > 1023) { p = p * 2.0; k -= 1; }
    return p * bits_to_double((uint64_t)(k + 1023) << 52);
}

TM_HD double dlog(double x)
{
#pragma clang fp contract(off)
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int e = 0;
    if (x < 0x1p-1022) { x = x * 0x1p54; e = -54; }
    const uint64_t b = double_to_bits(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double q = 0x1.642c8590b2164p-5;
    q = q * z + 0x1.8618618618618p-5;
    q = q * z + 0x1.af286bca1af28p-5;
    q = q * z + 0x1.e1e1e1e1e1e1ep-5;
    q = q * z + 0x1.1111111111111p-4;
    q = q * z + 0x1.3b13b13b13b14p-4;
    q = q * z + 0x1.745d1745d1746p-4;
    q = q * z + 0x1.c71c71c71c71cp-4;
    q = q * z + 0x1.2492492492492p-3;
    q = q * z + 0x1.999999999999ap-3;
    q = q * z + 0x1.5555555555555p-2;
    const double s2 = 2.0 * s;
    const double lm = s2 + s2 * (z * q);
    const double ed = (double)e;
    return ed * 0x1.62e42fee00000p-1 + (lm + ed * 0x1.a39ef35793c76p-33);
}

TM_HD double dsqrt(double x) { return __builtin_sqrt(x); }

// sin and cos of 2 pi u, for a consumer's advection direction: t = 4u (exact),
// q = nearest integer to t, x = (t - q) pi/2 in [-pi/4, pi/4], Taylor
// polynomials of sin x to x^17 and cos x to x^18 in x^2 (the next terms are
// below 1e-19), then the quarter-turn rotation by q.
TM_HD void dsincos2pi(double u, double& sn, double& cs)
{
#pragma clang fp contract(off)
    const double t = 4.0 * u;
    const double q = __builtin_floor(t + 0.5);
    const double x = (t - q) * 0x1.921fb54442d18p+0;
    const double z = x * x;
    double s = 0x1.952c77030ad4ap-49;
    s = s * z + -0x1.ae7f3e733b81fp-41;
    s = s * z + 0x1.6124613a86d09p-33;
    s = s * z + -0x1.ae64567f544e4p-26;
    s = s * z + 0x1.71de3a556c734p-19;
    s = s * z + -0x1.a01a01a01a01ap-13;
    s = s * z + 0x1.1111111111111p-7;
    s = s * z + -0x1.5555555555555p-3;
    s = x + x * (z * s);
    double c = -0x1.6827863b97d97p-53;
    c = c * z + 0x1.ae7f3e733b81fp-45;
    c = c * z + -0x1.93974a8c07c9dp-37;
    c = c * z + 0x1.1eed8eff8d898p-29;
    c = c * z + -0x1.27e4fb7789f5cp-22;
    c = c * z + 0x1.a01a01a01a01ap-16;
    c = c * z + -0x1.6c16c16c16c17p-10;
    c = c * z + 0x1.5555555555555p-5;
    c = c * z + -0x1.0000000000000p-1;
    c = 1.0 + z * c;
    const int k = ((int)q) & 3;
    sn = k == 0 ? s : (k == 1 ? c : (k == 2 ? -s : -c));
    cs = k == 0 ? c : (k == 1 ? -s : (k == 2 ? -c : s));
}

// Acklam's rational approximation of the standard normal quantile (|rel err|
// < 1.15e-9); one uniform in, one deviate out, no rejection.
TM_HD double norm_icdf_d(double p)
{
#pragma clang fp contract(off)
    const double plow = 0.02425;
    double q, num, den;
    if (p < plow || p > 1.0 - plow) {
        q = dsqrt(-2.0 * dlog(p < plow ? p : 1.0 - p));
        num = ((((-7.784894002430293e-03 * q - 3.223964580411365e-01) * q - 2.400758277161838e+00) * q
                - 2.549732539343734e+00) * q + 4.374664141464968e+00) * q + 2.938163982698783e+00;
        den = (((7.784695709041462e-03 * q + 3.224671290700398e-01) * q + 2.445134137142996e+00) * q
               + 3.754408661907416e+00) * q + 1.0;
        const double x = num / den;
        return p < plow ? x : -x;
    }
    q = p - 0.5;
    const double r = q * q;
    num = (((((-3.969683028665376e+01 * r + 2.209460984245205e+02) * r - 2.759285104469687e+02) * r
             + 1.383577518672690e+02) * r - 3.066479806614716e+01) * r + 2.506628277459239e+00) * q;
    den = ((((-5.447609879822406e+01 * r + 1.615858368580409e+02) * r - 1.556989798598866e+02) * r
            + 6.680131188771972e+01) * r - 1.328068155288572e+01) * r + 1.0;
    return num / den;
}

// The standard normal quantile in FP32, for the FP32 samplers: with t = 2p - 1 and
// w = -ln(4p(1 - p)) (= -ln(1 - t^2), formed without cancellation),
//   Phi^{-1}(p) = t * A(w - 5/2)         for w < 5 (|t| < 0.99663),
//               = t * B(sqrt(w) - 3)     above,
// the functional form of M. Giles, "Approximating the erfinv function" (GPU Computing
// Gems Jade Edition, 2011), scaled by sqrt(2).  A and B are degree-9 polynomials whose
// coefficients are this library's own weighted least-squares fit to Phi^{-1}
// (trophic_maelstorm_amd/csrc/fit_normal_quantile.py): 6e-9 relative in FP64, 2.4e-7
// (median 3e-8) evaluated in FP32 over every p that Stream::u32() yields.  One
// hardware log2 and nine FMAs per deviate; the square root and the tail polynomial
// run only in the waves that hold a tail lane.  p comes from Stream::u32(), so
// 4p(1 - p) >= 2^-22 and w is finite.
__device__ __forceinline__ float norm_icdf_f(float p)
{
    const float t = 2.0f * p - 1.0f;
    const float w = -0.693147181f * __builtin_amdgcn_logf((2.0f * p) * (2.0f * (1.0f - p)));
    float r;
    if (w < 5.0f) {
        const float z = w - 2.5f;
        r = -1.243243997e-08f;
        r = __builtin_fmaf(r, z, 3.285592243e-08f);
        r = __builtin_fmaf(r, z, 6.592649697e-07f);
        r = __builtin_fmaf(r, z, -4.897004146e-06f);
        r = __builtin_fmaf(r, z, -7.019234545e-06f);
        r = __builtin_fmaf(r, z, 3.087856749e-04f);
        r = __builtin_fmaf(r, z, -1.771644573e-03f);
        r = __builtin_fmaf(r, z, -5.907717161e-03f);
        r = __builtin_fmaf(r, z, 3.488020301e-01f);
        r = __builtin_fmaf(r, z, 2.123313427e+00f);
    } else {
        const float z = __builtin_amdgcn_sqrtf(w) - 3.0f;
        r = 1.839189063e-04f;
        r = __builtin_fmaf(r, z, -4.515237524e-04f);
        r = __builtin_fmaf(r, z, -1.106743875e-04f);
        r = __builtin_fmaf(r, z, 2.116327872e-03f);
        r = __builtin_fmaf(r, z, -5.071566440e-03f);
        r = __builtin_fmaf(r, z, 8.036936633e-03f);
        r = __builtin_fmaf(r, z, -1.080399193e-02f);
        r = __builtin_fmaf(r, z, 1.335824095e-02f);
        r = __builtin_fmaf(r, z, 1.416582465e+00f);
        r = __builtin_fmaf(r, z, 4.006434441e+00f);
    }
    return r * t;
}

// log(k!): exact table below 16, Stirling series (4 terms) above.
TM_HD double logfact(int64_t k)
{
#pragma clang fp contract(off)
    if (k < 16) {
        switch (k < 0 ? 0 : (int)k) {
        case 2: return 0x1.62e42fefa39efp-1;
        case 3: return 0x1.cab0bfa2a2002p+0;
        case 4: return 0x1.96ca77c922cf9p+1;
        case 5: return 0x1.326643c4479c9p+2;
        case 6: return 0x1.a51273acf01cap+2;
        case 7: return 0x1.10ce1f32dcc30p+3;
        case 8: return 0x1.5358e82fcb70dp+3;
        case 9: return 0x1.99a8921a7f7cfp+3;
        case 10: return 0x1.e357590954d15p+3;
        case 11: return 0x1.180973f3a8d74p+4;
        case 12: return 0x1.3fcba16d50143p+4;
        case 13: return 0x1.68d5a9c3b32cep+4;
        case 14: return 0x1.930f3df162a42p+4;
        case 15: return 0x1.be636a63fd346p+4;
        default: return 0.0;
        }
    }
    const double n = (double)k + 1.0;
    const double in = 1.0 / n;
    const double in2 = in * in;
    const double corr = in * (0x1.5555555555555p-4 - in2 * (0x1.6c16c16c16c17p-9
                        - in2 * (0x1.a01a01a01a01ap-11 - in2 * 0x1.3813813813814p-11)));
    return (n - 0.5) * dlog(n) - n + 0x1.d67f1c864beb5p-1 + corr;
}

// ---------------------------------------------------------------------------
// Poisson(z) in FP64 for both storage precisions (z = lambda e^{beta dt} rho is
// always formed in FP64).  z < 12: sequential-search inversion; z >= 12: PTRS
// transformed rejection (Hoermann 1993).  compat: the reference's Gaussian
// shortcut trunc(N sqrt z + z) above 100 (rietkerk_bjork_FKE.cpp:4175-4180).
// ---------------------------------------------------------------------------
template <class St>
TM_HD int64_t poisson_d(St& s, double z, bool compat)
{
#pragma clang fp contract(off)
    if (!(z > 0.0)) return 0;
    if (compat && z > 100.0) return (int64_t)(norm_icdf_d(s.u64()) * dsqrt(z) + z);
    if (z < 12.0) {
        const double u = s.u64();
        double p = dexp(-z);
        double F = p;
        int64_t k = 0;
        while (u > F && k < 200) {
            k += 1;
            p = p * z / (double)k;
            F = F + p;
        }
        return k;
    }
    const double slam = dsqrt(z);
    const double loglam = dlog(z);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double linv = dlog(invalpha);
    for (int it = 0; it < 64; ++it) {
        const double U = s.u64() - 0.5;
        const double V = s.u64();
        const double us = 0.5 - __builtin_fabs(U);
        const double kd = __builtin_floor((2.0 * a / us + b) * U + z + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kd;
        if (kd < 0.0 || (us < 0.013 && V > us)) continue;
        const double lhs = dlog(V) + linv - dlog(a / (us * us) + b);
        const double rhs = -z + kd * loglam - logfact((int64_t)kd);
        if (lhs <= rhs) return (int64_t)kd;
    }
    return (int64_t)__builtin_floor(z);
}

// Gamma(shape, 1), Marsaglia & Tsang (2000) with the U^(1/shape) boost below
// shape 1; shape <= 0 gives 0 (the Dornic shape-0 case).  FP64 version, at most
// TRIALS trials: with TRIALS = 64 it always returns true (64 rejections leave
// g = d); with fewer it returns false, leaving s and g unspecified, where the
// sampler would draw one more trial, or (TRIALS = 1) where the squeeze does not
// settle the trial: the two software logarithms of the test are left to the
// full sampler, which the caller runs from the transition's first word.
template <int TRIALS, class St>
TM_HD bool gamma_d_try(St& s, double shape, double& g)
{
#pragma clang fp contract(off)
    if (!(shape > 0.0)) { g = 0.0; return true; }
    const double a = shape < 1.0 ? shape + 1.0 : shape;
    const double d = a - 0x1.5555555555555p-2;
    const double c = 1.0 / dsqrt(9.0 * d);
    g = d;
    bool ok = false;
    for (int it = 0; it < TRIALS; ++it) {
        const double x = norm_icdf_d(s.u64());
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        const double u = s.u64();
        const double x2 = x * x;
        if (u < 1.0 - 0.0331 * x2 * x2) { g = d * v; ok = true; break; }
        if (TRIALS == 1) break;
        if (dlog(u) < 0.5 * x2 + d * (1.0 - v + dlog(v))) { g = d * v; ok = true; break; }
    }
    if (TRIALS < 64 && !ok) return false;
    if (shape < 1.0) g = g * dexp(dlog(s.u64()) / shape);
    return true;
}

TM_HD double gamma_d(Stream& s, double shape)
{
    double g;
    gamma_d_try<64, Stream>(s, shape, g);
    return g;
}

// ---------------------------------------------------------------------------
// Dornic constants for one parameter point (rietkerk_bjork_FKE.cpp:3847-3849,
// multiSPDP.cpp:2485-2486): beta = linear - 4 D/dx^2, e^{beta dt},
// lambda = 2 beta / (sigma^2 (e^{beta dt} - 1)).
// ---------------------------------------------------------------------------
struct DornicConst {
    double lambda;     // lambda
    double ll;         // lambda * e^{beta dt}   (Poisson mean per unit rho)
    double dc;         // D / dx^2
    double sigma2_1;   // 1 / sigma^2
    double inv_lambda; // 1 / lambda
    float inv_lambda_hi, inv_lambda_lo;   // 1 / lambda as an FP32 pair (the FP32 path)
    float shape_f;     // 2 (D/dx^2) / sigma^2 in FP32: mu + 1 = shape_f (up + down + left + right)
};

TM_HD DornicConst dornic_constants(double linear, double D, double sigma, double dt, double dx)
{
#pragma clang fp contract(off)
    const double dx2 = dx * dx;
    const double beta = linear - 4 * D / (dx2);
    const double lexp = dexp((beta) * dt);
    const double lambda = 2 * (beta) / (sigma * sigma * (lexp - 1.0));
    DornicConst c;
    c.lambda = lambda;
    c.ll = lambda * lexp;
    c.inv_lambda = 1.0 / lambda;
    c.inv_lambda_hi = (float)c.inv_lambda;
    c.inv_lambda_lo = (float)(c.inv_lambda - (double)c.inv_lambda_hi);
    c.dc = D / (dx2);
    c.sigma2_1 = 1 / (sigma * sigma);
    c.shape_f = (float)(2.0 * c.dc * c.sigma2_1);
    return c;
}

}  // namespace tmk
