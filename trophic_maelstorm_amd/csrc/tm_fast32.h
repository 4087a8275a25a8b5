// tm_fast32.h -- FP32 Poisson and Gamma samplers for the FP32-storage step.
//
// Same methods as the FP64 samplers in tm_device.h (inversion below 12, PTRS
// above; Marsaglia-Tsang Gamma), same Philox words in the same order, but
// evaluated in FP32 with the hardware transcendentals.  Where FP32 would lose
// the answer the formulation changes, not the precision:
//   * inversion runs on the upper tail S_k = P(N > k) against v = 1 - u, with
//     S_0 = -expm1(-z) from its series below 1/2, so P(N >= 1) = z stays
//     resolved as z -> 0 (the near-absorbing regime);
//   * PTRS forms k with the large mean added in FP64, and its log-pmf as
//     -stirlerr(k) - bd0(k, z) - log(2 pi k)/2 (Loader 2000), whose bd0 series
//     has no cancellation at k ~ z ~ 10^5;
//   * Marsaglia-Tsang's log test uses the series of d(1 - v + log v) + x^2/2 in
//     y = c x, which also has no cancellation at large shape.
#pragma once
#include "tm_device.h"

namespace tmk {

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float log2_hw(float x) { return __builtin_amdgcn_logf(x); }   // v_log_f32
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }  // v_exp_f32
__device__ __forceinline__ float logf_hw(float x) { return log2_hw(x) * 0.693147181f; }

// -expm1(-z) for z >= 0 in FP32, given e = e^{-z}: the series below 1/2, where
// 1 - e would cancel, and 1 - e above (both formed, one selected).
__device__ __forceinline__ float neg_expm1_neg(float z, float e)
{
    // z - z^2/2 + z^3/6 - ... to z^8 (truncation < 3e-8 relative at z = 1/2)
    float p = -2.48015873e-5f;
    p = p * z + 1.98412698e-4f;
    p = p * z - 1.38888889e-3f;
    p = p * z + 8.33333333e-3f;
    p = p * z - 4.16666667e-2f;
    p = p * z + 1.66666667e-1f;
    p = p * z - 0.5f;
    p = p * z + 1.0f;
    return z < 0.5f ? p * z : 1.0f - e;
}

// log(k!) - [(k + 1/2) log k - k + log(2 pi)/2]  (Stirling's error), k >= 1.
__device__ __forceinline__ float stirlerr_f(float k)
{
    if (k < 16.0f) {
        // exact values: lgamma(k+1) - ((k+0.5) log k - k + 0.5 log 2pi)
        const int i = (int)k;
        switch (i) {
        case 1: return 0.0810614667953272f;
        case 2: return 0.0413406959554092f;
        case 3: return 0.0276779256849983f;
        case 4: return 0.0207906721037650f;
        case 5: return 0.0166446911898211f;
        case 6: return 0.0138761288230707f;
        case 7: return 0.0118967099458917f;
        case 8: return 0.0104112652619720f;
        case 9: return 0.00925546218271273f;
        case 10: return 0.00833056343336287f;
        case 11: return 0.00757367548795184f;
        case 12: return 0.00694284010720953f;
        case 13: return 0.00640899418800420f;
        case 14: return 0.00595137011275885f;
        case 15: return 0.00555473355196280f;
        default: return 0.0f;
        }
    }
    const float r = rcpf(k), r2 = r * r;
    return r * (0.0833333333f - r2 * (0.00277777778f - r2 * 0.000793650794f));
}

// 1/k! in FP32, k = 0 .. 32 (all normal numbers; 1/34! would be subnormal).
static constexpr float kInvFact[33] = {
    1.0f, 1.0f, 0.5f, 0.1666666716337204f, 0.0416666679084301f, 0.008333333767950535f, 0.0013888889225199819f,
    0.00019841270113829523f, 2.4801587642286904e-05f, 2.7557318844628753e-06f, 2.755731998149713e-07f,
    2.5052107943679403e-08f, 2.0876755879584152e-09f, 1.6059044372074283e-10f, 1.147074536050896e-11f,
    7.647163609812713e-13f, 4.7794772561329454e-14f, 2.8114573589663704e-15f, 1.5619206814541513e-16f,
    8.220635078476521e-18f, 4.110317590937049e-19f, 1.9572941524685808e-20f, 8.896790959566848e-22f,
    3.868170297964731e-23f, 1.6117375912827668e-24f, 6.446950340479164e-26f, 2.4795962847996784e-27f,
    9.183689809957378e-29f, 3.279889164105036e-30f, 1.13099624727078e-31f, 3.769987809265653e-33f,
    1.2161250145930974e-34f, 3.800390670603429e-36f};

constexpr int kPoisTerms = 32;   // terms of the counted inversion

// Poisson(z), z < 12, by inversion on the upper tail S_k = P(N > k) against
// v = 1 - u: N = #{k >= 0 : v < S_j for every j <= k}.  S_0 = -expm1(-z) stays
// resolved as z -> 0 (P(N >= 1) = z near the absorbing state).  Every lane of a
// wavefront evaluates term k at the same time, so 1/k! is a literal and a term
// costs a multiply, a fused multiply-add, a compare and an add, with no branch
// but a wave-wide exit test every four terms: S_k = S_{k-1} - (e^{-z} z^k) / k!.
// Returns true when the count is final within kPoisTerms terms (every z < 12
// but with probability below 1e-6); otherwise n = kPoisTerms + 1, and S and P
// (= e^{-z} z^kPoisTerms) hold the state for poisson_tail.
__device__ __forceinline__ bool poisson_count(float z, float v, int& n, float& S, float& P)
{
    P = exp2_hw(z * -1.44269504f);          // e^{-z} = P(N = 0)
    S = neg_expm1_neg(z, P);                // P(N > 0)
    bool alive = v < S;
    n = alive ? 1 : 0;
#pragma unroll
    for (int k = 1; k <= kPoisTerms; ++k) {
        if (k > 1 && ((k - 1) & 3) == 0 && __builtin_amdgcn_ballot_w64(alive) == 0) break;
        P = P * z;
        S = __builtin_fmaf(-P, kInvFact[k], S);
        alive = alive && v < S;
        n += alive ? 1 : 0;
    }
    return !alive;
}

// The inversion past kPoisTerms terms (the sequential search, one term at a
// time).  It stops once the next term can no longer move S at FP32 resolution:
// a v in the rounding gap left below S maps to that extreme quantile.
__device__ __forceinline__ int poisson_tail(float z, float v, float S, float P)
{
    int k = kPoisTerms;
    float p = P * kInvFact[kPoisTerms];      // P(N = k)
    while (k < 200) {
        const float p1 = p * z * rcpf((float)(k + 1));
        if (p1 <= S * 0x1p-24f) return k + 1;
        S = S - p1;
        k += 1;
        if (!(v < S)) return k;
        p = p1;
    }
    return k;
}

// PTRS (Hoermann 1993) for z >= 12: the constants of a mean, and one trial
// from its two uniforms (U - 1/2 and V, drawn in that order).  k is formed with
// the large mean added in FP64; the log-pmf is -stirlerr(k) - bd0(k, z) -
// log(2 pi k)/2 (Loader 2000).  Returns true with kd set when the trial accepts.
struct PtrsF {
    double zd;
    float z, b, a, invalpha, vr;
};

__device__ __forceinline__ PtrsF ptrs_setup(double zd)
{
    PtrsF p;
    p.zd = zd;
    p.z = (float)zd;
    const float slam = __builtin_sqrtf(p.z);
    p.b = 0.931f + 2.53f * slam;
    p.a = -0.059f + 0.02483f * p.b;
    p.invalpha = 1.1239f + 1.1328f * rcpf(p.b - 3.4f);
    p.vr = 0.9277f - 3.6224f * rcpf(p.b - 2.0f);
    return p;
}

// A trial's candidate k = floor((2a/us + b) U + z + 0.43), us = 0.5 - |U|, and
// its fast acceptance (Hormann's box, us >= 0.07 and V <= v_r), which needs no
// logarithm.
__device__ __forceinline__ double ptrs_k(const PtrsF& p, float U, float us)
{
    return __builtin_floor((double)((2.0f * p.a * rcpf(us) + p.b) * U) + p.zd + 0.43);
}

__device__ __forceinline__ bool ptrs_in_box(const PtrsF& p, float us, float V) { return us >= 0.07f && V <= p.vr; }

__device__ __forceinline__ bool ptrs_trial(const PtrsF& p, float U, float V, double& kd)
{
    const float us = 0.5f - __builtin_fabsf(U);
    kd = ptrs_k(p, U, us);
    if (ptrs_in_box(p, us, V)) return true;
    if (kd < 0.0 || (us < 0.013f && V > us)) return false;
    const float k = (float)kd, z = p.z;
    // log p(k; z) = -stirlerr(k) - bd0(k, z) - log(2 pi k)/2, bd0 = k log(k/z) + z - k
    const float dk = (float)(kd - p.zd);
    float bd0;
    if (__builtin_fabsf(dk) < 0.1f * (k + z)) {
        const float w = dk * rcpf(k + z), w2 = w * w;
        bd0 = dk * w + 2.0f * k * w * w2 * (0.333333333f + w2 * (0.2f + w2 * (0.142857143f + w2 * 0.111111111f)));
    } else {
        bd0 = k * logf_hw(k * rcpf(z)) + z - k;
    }
    const float logp = k > 0.0f ? -stirlerr_f(k) - bd0 - 0.5f * logf_hw(6.28318531f * k) : -z;
    const float lhs = logf_hw(V * p.invalpha * rcpf(p.a * rcpf(us * us) + p.b));
    return lhs <= logp;
}

// Poisson(z), z formed in FP64 by the caller, returned as an integer-valued
// double (the callers want it as int64 and as float, each one conversion from
// FP64).  St: Stream, or Stream4 where at most four words are drawn.  z < 12:
// the counted inversion; z >= 12: PTRS.
template <class St>
__device__ __forceinline__ double poisson_fd(St& s, double zd, bool compat)
{
    if (!(zd > 0.0)) return 0.0;
    if (compat && zd > 100.0) {
        const float zf = (float)zd;
        return __builtin_trunc((double)(norm_icdf_f(s.u32()) * __builtin_sqrtf(zf)) + zd);
    }
    if (zd < 12.0) {
        const float z = (float)zd;
        const float v = s.v32();
        int n;
        float S, P;
        if (poisson_count(z, v, n, S, P)) return (double)n;
        return (double)poisson_tail(z, v, S, P);
    }
    const PtrsF pt = ptrs_setup(zd);
    for (int it = 0; it < 64; ++it) {
        const float U = s.u32() - 0.5f;
        const float V = s.u32();
        double kd;
        if (ptrs_trial(pt, U, V, kd)) return kd;
    }
    return __builtin_floor(zd);
}


// Gamma(shape, 1), Marsaglia-Tsang in FP32, at most TRIALS trials.  With
// TRIALS = 64 it always returns true (64 rejections leave g = d, as the FP64
// sampler does).  With fewer it returns false, and leaves s and g unspecified,
// where the sampler would have drawn one more trial, or (TRIALS = 1) where the
// squeeze does not settle the trial: the caller then starts the transition
// again from its first word with the full sampler.
template <int TRIALS, class St>
__device__ __forceinline__ bool gamma_f_try(St& s, float shape, float& g)
{
    if (!(shape > 0.0f)) { g = 0.0f; return true; }
    const float a = shape < 1.0f ? shape + 1.0f : shape;
    const float d = a - 0.333333343f;
    const float c = __builtin_amdgcn_rsqf(9.0f * d);
    const float k = c * c * 0.0833333358f;   // 1 / (108 d)
    g = d;
    bool ok = false;
    for (int it = 0; it < TRIALS; ++it) {
        const float x = norm_icdf_f(s.u32());
        const float y = c * x;
        float v = 1.0f + y;
        if (v <= 0.0f) continue;
        const float m = __builtin_fminf(v, 1.0f);
        v = v * v * v;
        const float u = s.u32();
        const float x2 = x * x;
        // Squeeze: u < 1 - x^4 / (108 d m), m = 1 + min(y, 0).  The test below is
        // log u < Q = x^2/2 + d(1 - v + log v) = 3d R(y), R(y) = log(1 + y) - y +
        // y^2/2 - y^3/3, and R(y) >= -y^4 / (4m) (the Taylor remainder of
        // log(1 + y) after y^4 is >= 0 for y >= 0; for y < 0, R = -sum_{k>=4}
        // |y|^k / k >= -y^4 / (4(1 + y))).  With y = c x, c^2 = 1/(9d):
        // Q >= -x^4/(108 d m), and e^Q >= 1 + Q, so every accepted u passes the
        // test too.  It accepts more than Marsaglia-Tsang's 1 - 0.0331 x^4 at
        // every d >= 2/3 (misses 6% of trials at d = 2/3, 1% at d = 4, under
        // 0.2% above d = 25), so a wave skips the logarithms more often.
        if (u * m < m - k * x2 * x2) { g = d * v; ok = true; break; }
        // One trial (the first pass of the deferring step): a trial the squeeze
        // leaves open is finished by the full sampler in the second pass, so
        // no lane of the first pass waits on the logarithms.
        if (TRIALS == 1) break;
        // x^2/2 + d(1 - v + log v); below |y| = 0.1 its series in y, which does not
        // cancel: d y^4 (-3/4 + 3/5 y - 1/2 y^2 + 3/7 y^3 - 3/8 y^4).  Both are
        // formed and one selected (v > 0 here, so the logarithm is finite).
        const float y2 = y * y;
        const float ser = d * y2 * y2 * (-0.75f + y * (0.6f + y * (-0.5f + y * (0.428571429f - 0.375f * y))));
        const float dir = 0.5f * x2 + d * (1.0f - v + logf_hw(v));
        const float acc = __builtin_fabsf(y) < 0.1f ? ser : dir;
        if (logf_hw(u) < acc) { g = d * v; ok = true; break; }
    }
    if (TRIALS < 64 && !ok) return false;
    if (shape < 1.0f) g = g * exp2_hw(log2_hw(s.u32()) * rcpf(shape));
    return true;
}

__device__ __forceinline__ float gamma_f(Stream& s, float shape)
{
    float g;
    gamma_f_try<64, Stream>(s, shape, g);
    return g;
}

}  // namespace tmk
