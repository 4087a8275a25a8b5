/*
 * tm_oracle.c -- CPU oracle (checker) for the Dornic / Rietkerk step.
 * TEST INFRASTRUCTURE ONLY; see tm_oracle.h for scope, citations and pins.
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math (oracle/Makefile).
 * Reference paths are relative to the reference's simulations/ directory.
 */
#include "tm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random   */
/* numbers: as easy as 1, 2, 3"): 10 rounds of the 4x32 S-P network.  */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void tmo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One uniform stream: key (seed, instance), counter (cell, step lo, step hi, block). */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int pos;
} stream_t;

static void st_init(stream_t* s, uint32_t seed, uint32_t inst, uint32_t cell, int64_t step)
{
    s->key[0] = seed; s->key[1] = inst;
    s->ctr[0] = cell;
    s->ctr[1] = (uint32_t)((uint64_t)step);
    s->ctr[2] = (uint32_t)((uint64_t)step >> 32);
    s->ctr[3] = 0;
    s->pos = 4;
}

static uint32_t st_word(stream_t* s)
{
    if (s->pos == 4) {
        tmo_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[3] += 1u;
        s->pos = 0;
    }
    return s->buf[s->pos++];
}

/* u in (0,1): (w + 1/2) 2^-32, exact in double. */
static double st_u(stream_t* s) { return ((double)st_word(s) + 0.5) * 0x1p-32; }

/* ------------------------------------------------------------------ */
/* exp / log from the mathematics (Cody-Waite reduction + Taylor /     */
/* atanh series), IEEE basic operations only.                          */
/* ------------------------------------------------------------------ */
static double as_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t as_u64(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

#define LN2_HI   0x1.62e42fee00000p-1
#define LN2_LO   0x1.a39ef35793c76p-33
#define INV_LN2  0x1.71547652b82fep+0
#define SQRT2    0x1.6a09e667f3bcdp+0

double tmo_exp(double x)
{
    if (x != x) return x;
    if (x > 0x1.62e42fefa39efp+9) return INFINITY;
    if (x < -745.13321910194122) return 0.0;
    double kd = floor(x * INV_LN2 + 0.5);
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 0x1.6124613a86d09p-33;          /* 1/13! */
    p = p * r + 0x1.1eed8eff8d898p-29;          /* 1/12! */
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
    return p * as_double((uint64_t)(k + 1023) << 52);
}

double tmo_log(double x)
{
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 0x1p-1022) { x = x * 0x1p54; e = -54; }
    uint64_t b = as_u64(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double q = 0x1.642c8590b2164p-5;            /* 1/23 */
    q = q * z + 0x1.8618618618618p-5;            /* 1/21 */
    q = q * z + 0x1.af286bca1af28p-5;
    q = q * z + 0x1.e1e1e1e1e1e1ep-5;
    q = q * z + 0x1.1111111111111p-4;
    q = q * z + 0x1.3b13b13b13b14p-4;
    q = q * z + 0x1.745d1745d1746p-4;
    q = q * z + 0x1.c71c71c71c71cp-4;
    q = q * z + 0x1.2492492492492p-3;
    q = q * z + 0x1.999999999999ap-3;
    q = q * z + 0x1.5555555555555p-2;            /* 1/3 */
    double s2 = 2.0 * s;
    double lm = s2 + s2 * (z * q);
    double ed = (double)e;
    return ed * LN2_HI + (lm + ed * LN2_LO);
}

/* Standard normal quantile: P. J. Acklam's rational approximation
 * (relative error < 1.15e-9), central region |p - 1/2| <= 0.47575. */
double tmo_norm_icdf(double p)
{
    const double plow = 0.02425;
    if (p < plow) {
        double q = sqrt(-2.0 * tmo_log(p));
        double num = ((((-7.784894002430293e-03 * q - 3.223964580411365e-01) * q - 2.400758277161838e+00) * q
                        - 2.549732539343734e+00) * q + 4.374664141464968e+00) * q + 2.938163982698783e+00;
        double den = (((7.784695709041462e-03 * q + 3.224671290700398e-01) * q + 2.445134137142996e+00) * q
                       + 3.754408661907416e+00) * q + 1.0;
        return num / den;
    }
    if (p <= 1.0 - plow) {
        double q = p - 0.5;
        double r = q * q;
        double num = (((((-3.969683028665376e+01 * r + 2.209460984245205e+02) * r - 2.759285104469687e+02) * r
                         + 1.383577518672690e+02) * r - 3.066479806614716e+01) * r + 2.506628277459239e+00) * q;
        double den = ((((-5.447609879822406e+01 * r + 1.615858368580409e+02) * r - 1.556989798598866e+02) * r
                        + 6.680131188771972e+01) * r - 1.328068155288572e+01) * r + 1.0;
        return num / den;
    }
    double q = sqrt(-2.0 * tmo_log(1.0 - p));
    double num = ((((-7.784894002430293e-03 * q - 3.223964580411365e-01) * q - 2.400758277161838e+00) * q
                    - 2.549732539343734e+00) * q + 4.374664141464968e+00) * q + 2.938163982698783e+00;
    double den = (((7.784695709041462e-03 * q + 3.224671290700398e-01) * q + 2.445134137142996e+00) * q
                   + 3.754408661907416e+00) * q + 1.0;
    return -(num / den);
}

/* log(k!): table below 16, Stirling series with 4 correction terms above. */
static const double LOGFACT[16] = {
    0x0.0p+0, 0x0.0p+0, 0x1.62e42fefa39efp-1, 0x1.cab0bfa2a2002p+0,
    0x1.96ca77c922cf9p+1, 0x1.326643c4479c9p+2, 0x1.a51273acf01cap+2, 0x1.10ce1f32dcc30p+3,
    0x1.5358e82fcb70dp+3, 0x1.99a8921a7f7cfp+3, 0x1.e357590954d15p+3, 0x1.180973f3a8d74p+4,
    0x1.3fcba16d50143p+4, 0x1.68d5a9c3b32cep+4, 0x1.930f3df162a42p+4, 0x1.be636a63fd346p+4,
};

double tmo_logfact(int64_t k)
{
    if (k < 16) return LOGFACT[k < 0 ? 0 : k];
    double n = (double)k + 1.0;
    double in = 1.0 / n;
    double in2 = in * in;
    double corr = in * (0x1.5555555555555p-4 - in2 * (0x1.6c16c16c16c17p-9
                  - in2 * (0x1.a01a01a01a01ap-11 - in2 * 0x1.3813813813814p-11)));
    return (n - 0.5) * tmo_log(n) - n + 0x1.d67f1c864beb5p-1 + corr;
}

/* ------------------------------------------------------------------ */
/* Samplers                                                            */
/* ------------------------------------------------------------------ */

/* Poisson(z).  z < 12: sequential-search inversion (Devroye 1986, X.3).
 * z >= 12: PTRS transformed rejection with squeeze (Hoermann 1993,
 * Insurance: Mathematics and Economics 12, 39-45). */
static int64_t poisson_exact(stream_t* s, double z)
{
    if (!(z > 0.0)) return 0;
    if (z < 12.0) {
        double u = st_u(s);
        double p = tmo_exp(-z);
        double F = p;
        int64_t k = 0;
        while (u > F && k < 200) {
            k += 1;
            p = p * z / (double)k;
            F = F + p;
        }
        return k;
    }
    double slam = sqrt(z);
    double loglam = tmo_log(z);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    double linv = tmo_log(invalpha);
    for (int it = 0; it < 64; ++it) {
        double U = st_u(s) - 0.5;
        double V = st_u(s);
        double us = 0.5 - fabs(U);
        double kd = floor((2.0 * a / us + b) * U + z + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kd;
        if (kd < 0.0 || (us < 0.013 && V > us)) continue;
        double lhs = tmo_log(V) + linv - tmo_log(a / (us * us) + b);
        double rhs = -z + kd * loglam - tmo_logfact((int64_t)kd);
        if (lhs <= rhs) return (int64_t)kd;
    }
    return (int64_t)floor(z);
}

/* Gamma(shape, 1): Marsaglia & Tsang (2000, ACM TOMS 26, 363-372) with the
 * U^(1/shape) boost for shape < 1; shape <= 0 gives 0 (Dornic shape 0). */
static double gamma_mt(stream_t* s, double shape)
{
    if (!(shape > 0.0)) return 0.0;
    double a = shape < 1.0 ? shape + 1.0 : shape;
    double d = a - 0x1.5555555555555p-2;
    double c = 1.0 / sqrt(9.0 * d);
    double g = d;
    for (int it = 0; it < 64; ++it) {
        double x = tmo_norm_icdf(st_u(s));
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        double u = st_u(s);
        double x2 = x * x;
        if (u < 1.0 - 0.0331 * x2 * x2) { g = d * v; break; }
        if (tmo_log(u) < 0.5 * x2 + d * (1.0 - v + tmo_log(v))) { g = d * v; break; }
    }
    if (shape < 1.0) g = g * tmo_exp(tmo_log(st_u(s)) / shape);
    return g;
}

void tmo_sample(int kind, double param, int mode, uint32_t seed, uint32_t inst,
                uint32_t cell, int64_t step, int64_t n, double* out)
{
    stream_t s;
    st_init(&s, seed, inst, cell, step);
    for (int64_t i = 0; i < n; ++i) {
        switch (kind) {
        case 0: out[i] = st_u(&s); break;
        case 1: out[i] = tmo_norm_icdf(st_u(&s)); break;
        case 2:
            if (mode == 1 && param > 100.0)
                out[i] = (double)(int64_t)(tmo_norm_icdf(st_u(&s)) * sqrt(param) + param);
            else
                out[i] = (double)poisson_exact(&s, param);
            break;
        default: out[i] = gamma_mt(&s, param); break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Dornic Poisson-Gamma transition of one cell                         */
/* FKE:4157-4196 (MSP:2904-2943 is the same statement for DP).         */
/* ------------------------------------------------------------------ */
static double dornic_cell(stream_t* s, double alpha_i, double cur, double ll, double lambda,
                          double sigma2_1, int compat, int64_t* n_out)
{
    double mu = -1.0 + 2.0 * alpha_i * sigma2_1;          /* FKE:4168 */
    double ziggy = ll * cur;                              /* FKE:4169, ll = lambda*lambda_exp */
    double gru;
    int64_t n;
    if (ziggy == 0.0) {                                   /* FKE:4171-4174 */
        n = 0;
        gru = mu + 1.0;
    } else if (compat && ziggy > 100.0) {                 /* FKE:4175-4180 */
        n = (int64_t)(tmo_norm_icdf(st_u(s)) * sqrt(ziggy) + ziggy);
        gru = mu + 1.0 + (double)n;
    } else {                                              /* FKE:4181-4185 */
        n = poisson_exact(s, ziggy);
        gru = mu + 1.0 + (double)n;
    }
    *n_out = n;
    if (compat && gru > 100.0)                            /* FKE:4186-4191 */
        return (tmo_norm_icdf(st_u(s)) * sqrt(gru) + gru) / lambda;
    return gamma_mt(s, gru) / lambda;                     /* FKE:4192-4196 */
}

void tmo_dornic_constants(double linear, double Dcoef, double sigma, double dt, double dx, double out[3])
{
    /* beta = linear - 4 D/dx^2 (FKE:3847 with linear = -d; MSP:2485 with linear = a) */
    double dx2 = dx * dx;
    double beta = linear - 4 * Dcoef / (dx2);
    double lexp = tmo_exp((beta) * dt);
    double lambda = 2 * (beta) / (sigma * sigma * (lexp - 1.0));
    out[0] = beta; out[1] = lexp; out[2] = lambda;
}

double tmo_dornic_cell(double alpha_i, double rho, double lambda, double lambda_exp, double sigma,
                       int sampler, uint32_t seed, uint32_t inst, uint32_t cell, int64_t step,
                       int64_t* n_out)
{
    stream_t s;
    st_init(&s, seed, inst, cell, step);
    int64_t n = -1;
    double r = rho;
    if (!(alpha_i == 0 && rho == 0))
        r = dornic_cell(&s, alpha_i, rho, lambda * lambda_exp, lambda, 1 / (sigma * sigma), sampler == 1, &n);
    if (n_out) *n_out = n;
    return r;
}

void tmo_dornic_batch(double alpha_i, double rho, double lambda, double lambda_exp, double sigma,
                      int sampler, uint32_t seed, uint32_t inst, int64_t step, int64_t n,
                      double* out, int64_t* n_out)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = -1;
        out[i] = tmo_dornic_cell(alpha_i, rho, lambda, lambda_exp, sigma, sampler, seed, inst,
                                 (uint32_t)i, step, &k);
        if (n_out) n_out[i] = k;
    }
}

/* ------------------------------------------------------------------ */
/* Initial fields, reductions, Laplacian                               */
/* ------------------------------------------------------------------ */
void tmo_init_uniform(int L, int nf, uint32_t seed, uint32_t inst,
                      const double* lo, const double* hi, double* out)
{
    const int64_t N = (int64_t)L * L;
    uint32_t key[2] = {seed, inst};
    for (int64_t i = 0; i < N; ++i) {
        uint32_t w[4];
        for (int f = 0; f < nf; ++f) {     /* field f: word f % 4 of block f / 4 */
            if (f % 4 == 0) {
                uint32_t ctr[4] = {(uint32_t)i, 0xFFFFFFFFu, 0xFFFFFFFFu, (uint32_t)(f / 4)};
                tmo_philox4x32_10(ctr, key, w);
            }
            double u = ((double)w[f % 4] + 0.5) * 0x1p-32;
            out[f * N + i] = lo[f] + (hi[f] - lo[f]) * u;
        }
    }
}

/* Post-step flush: x < eps && x >= neg -> +0.0, anything else unchanged (NaN
 * included).  Rietkerk eps 1e-10, neg 0 (FKE:4459-4461); DP eps 1e-12,
 * neg -1e-80 (MSP:3257-3258). */
double tmo_flush(double x, double eps, double neg) { return (x < eps && x >= neg) ? 0.0 : x; }

/* mean_of_vector (FKE:1238-1246): naive sum / size; occupied_sites_of_vector
 * (FKE:1271-1280): count of entries > 0.  Row layout FKE:4024 (count, mean). */
void tmo_reduce(int L, int nf, const double* fields, double* row)
{
    const int64_t N = (int64_t)L * L;
    for (int f = 0; f < nf; ++f) {
        const double* x = fields + f * N;
        double sum = 0.0, cnt = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            sum += x[i];
            if (x[i] > 0) cnt += 1;
        }
        row[2 * f] = cnt;
        row[2 * f + 1] = sum / (double)N;
    }
}

/* Periodic 5-point neighbours (determine_neighbours_Sq4, FKE:1634-1653):
 * up = row p-1, down = row p+1, left = column q-1, right = column q+1. */
static inline int64_t nb_up(int L, int p, int q) { return (int64_t)((p - 1 + L) % L) * L + q; }
static inline int64_t nb_dn(int L, int p, int q) { return (int64_t)((p + 1) % L) * L + q; }
static inline int64_t nb_lf(int L, int p, int q) { return (int64_t)p * L + (q - 1 + L) % L; }
static inline int64_t nb_rt(int L, int p, int q) { return (int64_t)p * L + (q + 1) % L; }

static inline double nsum(const double* x, int L, int p, int q)
{
    return x[nb_up(L, p, q)] + x[nb_dn(L, p, q)] + x[nb_lf(L, p, q)] + x[nb_rt(L, p, q)];
}

void tmo_laplacian(int L, const double* in, double* out)
{
    for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q) {
            int64_t i = (int64_t)p * L + q;
            out[i] = nsum(in, L, p, q) - 4 * in[i];
        }
}

/* ------------------------------------------------------------------ */
/* DP                                                                   */
/* ------------------------------------------------------------------ */
static double dp_f(double b, double y) { return -b * y * y; }   /* f_DP_Dor_1Sp, MSP:1662 */

double tmo_dp_nonlinear(int scheme, double b, double dt, double y)
{
    if (scheme == 2)                                  /* stochasticSPDP.cpp:1066 */
        return (y) / (1 + b * y * dt);
    if (scheme == 1) {                                /* MSP:1835-1885 */
        double dt2 = dt / 2.0, dt6 = dt / 6.0;
        double k1 = dp_f(b, y);
        double ys = y + (dt2) * k1;
        double k2 = dp_f(b, ys);
        ys = y + (dt2) * k2;
        double k3 = dp_f(b, ys);
        ys = y + (dt) * k3;
        double k4 = dp_f(b, ys);
        return y + (dt6) * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
    }
    return y + dt * dp_f(b, y);                       /* explicit Euler */
}

static double ramp_a(const tmo_dp_params* p, int64_t step)
{
    if (p->ramp_steps <= 0) return p->a;
    int64_t m = step < p->ramp_steps ? step : p->ramp_steps;
    return p->a + (p->a_end - p->a) * ((double)m / (double)p->ramp_steps);
}

int tmo_dp_run(const tmo_dp_params* p, int L, uint32_t seed, uint32_t inst, int64_t step0,
               int64_t nsteps, double* rho, double* rows, int64_t* draws, int64_t* fault_cell)
{
    const int64_t N = (int64_t)L * L;
    double* old = (double*)malloc(sizeof(double) * N);
    const double dx2 = p->dx * p->dx;
    const double Dc = p->D / (dx2);                   /* MSP:2490 diff_coefficient */
    const double sigma2_1 = 1 / (p->sigma * p->sigma);
    const int compat = p->sampler == 1;
    int fault = 0;
    if (fault_cell) *fault_cell = -1;
    for (int64_t n = 0; n < nsteps; ++n) {
        const int64_t step = step0 + n;
        double cst[3];
        tmo_dornic_constants(ramp_a(p, step), p->D, p->sigma, p->dt, p->dx, cst);
        const double lambda = cst[2], ll = cst[2] * cst[1];
        memcpy(old, rho, sizeof(double) * N);          /* DRho (MSP:3259) */
        const int last = (n == nsteps - 1);
        for (int pp = 0; pp < L; ++pp)
            for (int q = 0; q < L; ++q) {
                const int64_t i = (int64_t)pp * L + q;
                double alpha_i = Dc * (nsum(old, L, pp, q));  /* MSP:2904-2905 */
                double y = old[i];
                int64_t draw = -1;
                if (!(alpha_i == 0 && old[i] == 0)) {          /* MSP:2907-2910 */
                    stream_t s;
                    st_init(&s, seed, inst, (uint32_t)i, step);
                    y = dornic_cell(&s, alpha_i, old[i], ll, lambda, sigma2_1, compat, &draw);
                }
                if (last && draws) draws[i] = draw;
                y = tmo_dp_nonlinear(p->scheme, p->b, p->dt, y);
                if (y < 0 || !isfinite(y)) {                   /* MSP:3236-3254 */
                    if (!fault && fault_cell) *fault_cell = i;
                    fault = 1;
                }
                rho[i] = tmo_flush(y, p->flush_eps, p->flush_neg);   /* MSP:3257-3258 */
            }
        if (rows) tmo_reduce(L, 1, rho, rows + 2 * n);
        if (fault) break;
    }
    free(old);
    return fault;
}

/* ------------------------------------------------------------------ */
/* Rietkerk                                                             */
/* ------------------------------------------------------------------ */
/* f_2Dor (FKE:2363-2390): reaction of (P, W, O) plus W/O diffusion.
 * The P linear and noise terms belong to the Dornic sub-step. */
static void rk_rhs(const tmo_rk_params* p, int L, const double* y, double* f)
{
    const int64_t N = (int64_t)L * L;
    const double dx2 = p->dx * p->dx;
    const double Dw = p->D[1] / (dx2), Do = p->D[2] / (dx2);   /* FKE:3866-3881 */
    const double cgmax = p->c * p->gmax;
    const double* P = y; const double* W = y + N; const double* O = y + 2 * N;
    for (int pp = 0; pp < L; ++pp)
        for (int q = 0; q < L; ++q) {
            const int64_t i = (int64_t)pp * L + q;
            f[i] = cgmax * W[i] * P[i] / (W[i] + p->K[1]);
            double A = p->alpha * (P[i] + p->K[2] * p->W0) / (P[i] + p->K[2]) * O[i];
            f[N + i] = A - p->gmax * W[i] * P[i] / (W[i] + p->K[1]) - p->rW * W[i]
                       + (Dw) * (nsum(W, L, pp, q) - 4 * W[i]);
            f[2 * N + i] = p->R - A + (Do) * (nsum(O, L, pp, q) - 4 * O[i]);
        }
}

void tmo_rk_rhs(const tmo_rk_params* p, int L, const double* y, double* f) { rk_rhs(p, L, y, f); }

/* RK4_Integrate_Stochastic_1Sp (FKE:2393-2429) on the full 3-field state. */
static void rk4_full(const tmo_rk_params* p, int L, double* y, double* ys, double* k1, double* k2,
                     double* k3, double* k4)
{
    const int64_t M = 3 * (int64_t)L * L;
    const double dt = p->dt, dt2 = dt / 2.0, dt6 = dt / 6.0;   /* FKE:191 */
    rk_rhs(p, L, y, k1);
    for (int64_t i = 0; i < M; ++i) ys[i] = y[i] + (dt2) * k1[i];
    rk_rhs(p, L, ys, k2);
    for (int64_t i = 0; i < M; ++i) ys[i] = y[i] + (dt2) * k2[i];
    rk_rhs(p, L, ys, k3);
    for (int64_t i = 0; i < M; ++i) ys[i] = y[i] + (dt) * k3[i];
    rk_rhs(p, L, ys, k4);
    for (int64_t i = 0; i < M; ++i) y[i] += (dt6) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

void tmo_rk4_deterministic(const tmo_rk_params* p, int L, int64_t nsteps, double* y)
{
    const int64_t M = 3 * (int64_t)L * L;
    double* buf = (double*)malloc(sizeof(double) * M * 5);
    for (int64_t n = 0; n < nsteps; ++n)
        rk4_full(p, L, y, buf, buf + M, buf + 2 * M, buf + 3 * M, buf + 4 * M);
    free(buf);
}

int tmo_rk_run(const tmo_rk_params* p, int L, uint32_t seed, uint32_t inst, int64_t step0,
               int64_t nsteps, double* fields, double* rows, int64_t* draws, int64_t* fault_cell)
{
    const int64_t N = (int64_t)L * L, M = 3 * N;
    double* buf = (double*)malloc(sizeof(double) * (M * 5 + N));
    double* old = buf + 5 * M;
    const double dx2 = p->dx * p->dx;
    double cst[3];
    tmo_dornic_constants(-p->d, p->D[0], p->sigma, p->dt, p->dx, cst);   /* FKE:3847-3849 */
    const double lambda = cst[2], ll = cst[2] * cst[1];
    const double Dp = p->D[0] / (dx2);
    const double sigma2_1 = 1 / (p->sigma * p->sigma);
    const int compat = p->sampler == 1;
    int fault = 0;
    if (fault_cell) *fault_cell = -1;
    for (int64_t n = 0; n < nsteps; ++n) {
        const int64_t step = step0 + n;
        const int last = (n == nsteps - 1);
        memcpy(old, fields, sizeof(double) * N);        /* DRho[0] */
        for (int pp = 0; pp < L; ++pp)                   /* FKE:4154-4211 */
            for (int q = 0; q < L; ++q) {
                const int64_t i = (int64_t)pp * L + q;
                double alpha_i = Dp * (nsum(old, L, pp, q));
                int64_t draw = -1;
                if (!(alpha_i == 0 && old[i] == 0)) {
                    stream_t s;
                    st_init(&s, seed, inst, (uint32_t)i, step);
                    fields[i] = dornic_cell(&s, alpha_i, old[i], ll, lambda, sigma2_1, compat, &draw);
                }
                if (last && draws) draws[i] = draw;
            }
        rk4_full(p, L, fields, buf, buf + M, buf + 2 * M, buf + 3 * M, buf + 4 * M);   /* FKE:4235 */
        for (int64_t i = 0; i < M; ++i) {                /* FKE:4352-4462 */
            double y = fields[i];
            if (y < 0 || !isfinite(y)) {
                if (!fault && fault_cell) *fault_cell = i;
                fault = 1;
            }
            fields[i] = tmo_flush(y, p->flush_eps, 0.0);   /* FKE:4459-4461 */
        }
        if (rows) tmo_reduce(L, 3, fields, rows + 6 * n);
        if (fault) break;
    }
    free(buf);
    return fault;
}

/* ---- across-replicate statistics ------------------------------------------
 * Restated replicate by replicate, as the reference folds them: the first
 * replicate of a parameter point initialises the table (FKE:4578-4597), every
 * later one updates it through the running-statistics recurrence
 * (FKE:1319-1369), whose time loop stops at the replicate's first measurement
 * with no occupied cell.  With flag 2 a replicate's rows after that measurement
 * are zero, as the reference's Rho_M rows are when the time loop has ended
 * there (FKE:3086-3135, 4068-4072).  Table row: [1..nf-1] other-field means,
 * then mean, var, surviving count, mean occupied count of field 0. */
void tmo_replicate_stats(const double* series, int n_inst, int n_meas, int nf,
                         const int32_t* params_index, int n_params, int flags, double* out)
{
    const int w = 2 * nf, nc = nf + 3;
    const int intdiv = (flags & 1) != 0, stop = (flags & 2) != 0;
    double* x = (double*)malloc(sizeof(double) * (size_t)(n_meas > 0 ? n_meas : 1) * w);
    memset(out, 0, sizeof(double) * (size_t)n_params * n_meas * nc);
    for (int g = 0; g < n_params; ++g) {
        double* tab = out + (int64_t)g * n_meas * nc;
        int j = 0;
        for (int k = 0; k < n_inst; ++k) {
            if (params_index[k] != g) continue;
            /* this replicate's logged rows (X_curr), zero after its end if asked */
            int ended = 0;
            for (int i = 0; i < n_meas; ++i) {
                const double* src = series + ((int64_t)k * n_meas + i) * w;
                for (int c = 0; c < w; ++c) x[(int64_t)i * w + c] = ended ? 0.0 : src[c];
                if (stop && src[0] == 0.0) ended = 1;
            }
            if (j == 0) {                                         /* FKE:4581-4596 */
                for (int i = 0; i < n_meas; ++i) {
                    double* row = tab + (int64_t)i * nc;
                    const double* xi = x + (int64_t)i * w;
                    row[nf - 1] = xi[1];
                    row[nf] = 0.0;
                    row[nf + 2] = xi[0];
                    if (xi[0] > 0) row[nf + 1] = 1;
                    else if (xi[0] == 0) row[nf + 1] = 0;
                    for (int f = 1; f < nf; ++f) row[f - 1] = xi[2 * f + 1];
                }
            } else {
                for (int i = 0; i < n_meas; ++i) {                /* FKE:1330-1358 */
                    double* row = tab + (int64_t)i * nc;
                    const double* xi = x + (int64_t)i * w;
                    const int r = row[nf + 1] + 1;
                    if (xi[0] == 0.0) break;
                    const double mean_prev = row[nf - 1];
                    if (row[nf + 1] == 0.0) {
                        row[nf - 1] = xi[1];
                        row[nf] = 0.0;
                        row[nf + 2] = xi[0];
                    } else {
                        row[nf - 1] = (xi[1] + (r - 1) * row[nf - 1]) / r;
                        row[nf + 2] = (xi[0] + (r - 1) * row[nf + 2]) / r;
                        const double inner = (r - 2) * row[nf] + (xi[1] - row[nf - 1]) * (xi[1] - row[nf - 1]);
                        if (intdiv)
                            row[nf] = (row[nf - 1] - mean_prev) * (row[nf - 1] - mean_prev) + (1 / (r - 1)) * inner;
                        else
                            row[nf] = (row[nf - 1] - mean_prev) * (row[nf - 1] - mean_prev) + inner / (r - 1);
                    }
                    row[nf + 1] = r;
                }
                for (int i = 0; i < n_meas; ++i)                  /* FKE:1361-1366 */
                    for (int f = 1; f < nf; ++f) {
                        double* row = tab + (int64_t)i * nc;
                        row[f - 1] = (x[(int64_t)i * w + 2 * f + 1] + j * row[f - 1]) / (j + 1);
                    }
            }
            ++j;
        }
    }
    free(x);
}

/* ---- initial frames -------------------------------------------------------- */
/* init_randconstframe (FKE:409-431): one uniform per cell decides high or low
 * for every field together. */
void tmo_init_speckle(int L, int nf, uint32_t seed, uint32_t inst, double perc,
                      const double* high, const double* low, double* out)
{
    const int64_t N = (int64_t)L * L;
    for (int64_t i = 0; i < N; ++i) {
        const uint32_t ctr[4] = {(uint32_t)i, 0xFFFFFFFEu, 0xFFFFFFFFu, 0u}, key[2] = {seed, inst};
        uint32_t w[4];
        tmo_philox4x32_10(ctr, key, w);
        const double r = ((double)w[0] + 0.5) * 0x1p-32;
        for (int s = 0; s < nf; ++s) out[s * N + i] = r <= perc ? high[s] : low[s];
    }
}

/* init_gaussframe (FKE:1089-1183) with init_gaussiantear (FKE:1185-1199):
 * centres cx = [g/8, 7g/8, 7g/8, g/8, ...], cy = [g/8, g/8, 7g/8, 7g/8, ...]
 * (integer division), widths g / sd_div, then amp exp(-(di^2 + dj^2) / (2 sd sd))
 * at cell i*g + j; the remaining fields are the constants amp[s]. */
void tmo_init_gauss(int L, int nf, int n_gauss, const double* sd_div, int n_sd,
                    const double* amp, double* out)
{
    const int g = L;
    for (int s = 0; s < nf; ++s) {
        double* a = out + (int64_t)s * g * g;
        if (s >= n_gauss) {
            for (int64_t i = 0; i < (int64_t)g * g; ++i) a[i] = amp[s];
            continue;
        }
        const double cx = (s % 4) % 3 == 0 ? g / 8 : 7 * g / 8;
        const double cy = (s % 4) < 2 ? g / 8 : 7 * g / 8;
        const double sd = g / sd_div[s < n_sd ? s : n_sd - 1];
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j)
                a[(int64_t)i * g + j] = amp[s] * tmo_exp(-((i - cx) * (i - cx) + (j - cy) * (j - cy)) / (2 * sd * sd));
    }
}

/* ------------------------------------------------------------------ */
/* Consumer species: vegetation + grazer (SPB = 2) and + predator      */
/* (SPB = 3), the time-loop body of dP_Dornic_2D_MultiSp               */
/* (MSP:2901-2958 vegetation Dornic, :2962-2978 spatial means and the  */
/* cover factor, :1699-1786 / :2166-2284 perception gamma, :3041-3190   */
/* consumer Dornic with upwind advection, :1670-1697 / :1789-1819 +     */
/* :1821-1898 RK4 of the trophic terms, :3228-3262 checks and flush).  */
/* ------------------------------------------------------------------ */

/* sin and cos of 2 pi u: t = 4u (exact), q = nearest integer to t, x = (t - q) pi/2
 * in [-pi/4, pi/4]; Taylor polynomials of sin x to x^17 and cos x to x^18 in x^2
 * (next terms below 1e-19), then the quarter-turn rotation by q. */
void tmo_sincos2pi(double u, double* s_out, double* c_out)
{
    const double t = 4.0 * u;
    const double q = floor(t + 0.5);
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
    if (k == 0) { *s_out = s; *c_out = c; }
    else if (k == 1) { *s_out = c; *c_out = -s; }
    else if (k == 2) { *s_out = -s; *c_out = -c; }
    else { *s_out = -c; *c_out = s; }
}

/* Fixed-order sum of x[0..n): chunks of 4096 cells; in a chunk lane t (0..255)
 * adds cells t, t + 256, ... from 0.0, the 256 lane sums are folded pairwise
 * (s[t] += s[t + o] for o = 128, 64, ..., 1), and the chunk sums are added in
 * chunk order from 0.0.  It replaces mean_of_vector's sequential sum
 * (MSP:997-1005) wherever the mean steers the dynamics, so that a parallel
 * reduction can reproduce it exactly. */
double tmo_fixed_sum(const double* x, int64_t n)
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

static int64_t wrapi(int64_t v, int L) { int64_t m = v % L; return m < 0 ? m + L : m; }

/* f_DP_Dor_2Sp (MSP:1670-1697) and f_DP_Dor_3Sp (MSP:1789-1819) at one cell. */
static void dpc_f(const tmo_dpc_params* p, int ns, const double* y, double* f)
{
    const double A01H01 = p->A[0][1] * p->H[0][1];
    const double EA1 = p->E[1] * p->A[0][1];
    f[0] = -p->b * y[0] * y[0] - (p->A[0][1] * y[0] / (1 + A01H01 * y[0])) * y[1];
    if (ns == 2) {
        f[1] = (EA1 * y[0] / (1 + A01H01 * y[0])) * y[1];
        return;
    }
    const double A12H12 = p->A[1][2] * p->H[1][2];
    const double EA2 = p->E[2] * p->A[1][2];
    f[1] = (EA1 * y[0] / (1 + A01H01 * y[0])) * y[1] - p->A[1][2] * y[1] / (1 + A12H12 * y[1]) * y[2];
    f[2] = (EA2 * y[1] / (1 + A12H12 * y[1])) * y[2];
}

/* RK4_Integrate_Stochastic_MultiSp (MSP:1821-1898) at one cell: the trophic
 * terms have no spatial coupling, so the four stages are pointwise. */
void tmo_dpc_rk4(const tmo_dpc_params* p, int ns, double* y)
{
    const double dt = p->dt, dt2 = dt / 2.0, dt6 = dt / 6.0;
    double k1[3], k2[3], k3[3], k4[3], ys[3] = {0.0, 0.0, 0.0};
    dpc_f(p, ns, y, k1);
    for (int s = 0; s < ns; ++s) ys[s] = y[s] + (dt2) * k1[s];
    dpc_f(p, ns, ys, k2);
    for (int s = 0; s < ns; ++s) ys[s] = y[s] + (dt2) * k2[s];
    dpc_f(p, ns, ys, k3);
    for (int s = 0; s < ns; ++s) ys[s] = y[s] + (dt) * k3[s];
    dpc_f(p, ns, ys, k4);
    for (int s = 0; s < ns; ++s) y[s] += (dt6) * (k1[s] + 2.0 * k2[s] + 2.0 * k3[s] + k4[s]);
}

static double dpc_ramp_a(const tmo_dpc_params* p, int64_t step)
{
    if (p->ramp_steps <= 0) return p->a;
    int64_t m = step < p->ramp_steps ? step : p->ramp_steps;
    return p->a + (p->a_end - p->a) * ((double)m / (double)p->ramp_steps);
}

int tmo_dpc_run(const tmo_dpc_params* p, int ns, int L, const int32_t* disk, int disk_len,
                uint32_t seed, uint32_t inst, int64_t step0, int64_t nsteps, double* fields,
                double* rows, int64_t* fault_index)
{
    const int nf = 2 * ns - 1;                 /* rho_0 .. rho_{ns-1}, gamma_1 .. gamma_{ns-1} */
    const int64_t N = (int64_t)L * L;
    double* old = (double*)malloc(sizeof(double) * N * nf);
    const double eps_g = 1.0e-12;              /* calc_gamma eps, MSP:1715 / :2182 */
    const double dx = p->dx, dt = p->dt;
    const double dx2 = dx * dx, dx1 = 1 / dx, dx1_2 = 1 / (dx2), dxb2 = dx / 2.0;   /* MSP:2470 */
    const double Dc0 = p->D[0] / (dx2);        /* diff_coefficient[0], MSP:2479 */
    double sigma2_1[3];
    for (int s = 0; s < ns; ++s) sigma2_1[s] = 1 / (p->sigma[s] * p->sigma[s]);
    double rmax = p->pR[0];                    /* r_max_effective, MSP:2429 */
    for (int s = 1; s < ns; ++s) if (p->pR[s] > rmax) rmax = p->pR[s];
    rmax = rmax / dx;
    double fr[3];
    for (int s = 0; s < ns; ++s) fr[s] = p->pR[s] / (dx * rmax);   /* MSP:2434-2435 */
    const int compat = p->sampler == 1;
    int fault = 0;
    if (fault_index) *fault_index = -1;
    for (int64_t n = 0; n < nsteps; ++n) {
        const int64_t step = step0 + n;
        memcpy(old, fields, sizeof(double) * N * nf);
        /* spatial means and vegetation cover of the step's start (MSP:2962-2975) */
        double mean[3], inv[3];
        for (int s = 0; s < ns; ++s) mean[s] = tmo_fixed_sum(old + s * N, N) / (double)N;
        double occ = 0.0;
        for (int64_t i = 0; i < N; ++i) if (old[i] > 0) occ += 1;
        double nfac = 1 - occ / (double)N;
        if (nfac < 0.35) nfac = 0.35;
        const int new_len = (int)(nfac * nfac * (double)disk_len);   /* MSP:1716 / :2183 */
        for (int s = 0; s < ns; ++s) inv[s] = mean[s] >= eps_g ? 1 / mean[s] : 1 / eps_g;
        int eff[3];
        for (int s = 0; s < ns; ++s) {
            eff[s] = (int)(fr[s] * fr[s] * (double)new_len);
            if (ns == 3 && eff[s] < 1 && fr[s] > 0) eff[s] = 1;    /* MSP:2215-2216 */
        }
        double cst[3];
        tmo_dornic_constants(dpc_ramp_a(p, step), p->D[0], p->sigma[0], dt, dx, cst);
        const double lam0 = cst[2], ll0 = cst[2] * cst[1];
        for (int pp = 0; pp < L; ++pp)
            for (int q = 0; q < L; ++q) {
                const int64_t i = (int64_t)pp * L + q;
                const int64_t up = nb_up(L, pp, q), dn = nb_dn(L, pp, q), lf = nb_lf(L, pp, q), rt = nb_rt(L, pp, q);
                double y[3];
                /* vegetation Dornic (MSP:2901-2945) */
                const double* r0 = old;
                const double alpha0 = Dc0 * (r0[up] + r0[dn] + r0[lf] + r0[rt]);
                y[0] = r0[i];
                if (!(alpha0 == 0 && r0[i] == 0)) {
                    stream_t st;
                    int64_t k;
                    st_init(&st, seed, inst, (uint32_t)i, step);
                    y[0] = dornic_cell(&st, alpha0, r0[i], ll0, lam0, sigma2_1[0], compat, &k);
                }
                /* perception gamma (calc_gamma_2Sp_NonRefugia / calc_gamma_3Sp_NonRefugia) */
                double gam[3] = {0.0, 0.0, 0.0};
                for (int s = 1; s < ns; ++s) {
                    double g = old[(ns + s - 1) * N + i];
                    if (fr[s] == 0.0) { gam[s] = g; continue; }
                    if (mean[s] < eps_g) { gam[s] = 0.0; continue; }
                    if (mean[s - 1] < eps_g) { gam[s] = 0.0; continue; }
                    const double* src = old + (s - 1) * N;
                    if (ns == 2) {                                 /* MSP:1756-1760: accumulates */
                        for (int k = 0; k < eff[s]; ++k)
                            g += src[wrapi(pp + disk[2 * k], L) * L + wrapi(q + disk[2 * k + 1], L)];
                        g = (g * inv[s - 1]) / (eff[s]);
                    } else if (s == 1) {                           /* MSP:2228-2249 */
                        double sum = 0.0;
                        const double* pr = old + 2 * N;
                        for (int k = 0; k < eff[s]; ++k) {
                            const int64_t j = wrapi(pp + disk[2 * k], L) * L + wrapi(q + disk[2 * k + 1], L);
                            if (mean[2] < eps_g) sum += src[j];
                            else sum += src[j] * (1 - pr[j] * inv[2]);
                        }
                        g = (sum * inv[s - 1]) / (eff[s]);
                    } else {                                       /* MSP:2251-2258 */
                        double sum = 0.0;
                        for (int k = 0; k < eff[s]; ++k)
                            sum += src[wrapi(pp + disk[2 * k], L) * L + wrapi(q + disk[2 * k + 1], L)];
                        g = (sum / (eff[s])) * inv[s - 1];
                    }
                    if (g > 1) g = 1.0;                            /* MSP:1763-1766 */
                    else if (g < 0) g = 0.0;
                    gam[s] = g;
                }
                /* consumer Dornic with advection (MSP:3041-3190) */
                for (int s = 1; s < ns; ++s) {
                    const double* rs = old + s * N;
                    y[s] = rs[i];
                    if (mean[s] == 0.0) continue;                  /* MSP:3044-3045 */
                    if (gam[s] > 1) gam[s] = 1;
                    stream_t st;
                    st_init(&st, seed, inst, (uint32_t)i, step);
                    st.ctr[3] = (uint32_t)s << 24;
                    double sn, cs;
                    tmo_sincos2pi(st_u(&st), &sn, &cs);
                    const double vx = p->v[s] * cs, vy = p->v[s] * sn;
                    const double vxa = vx * (double)((0 < vx) - (vx < 0));
                    const double vya = vy * (double)((0 < vy) - (vy < 0));
                    const double gp = (1 - gam[s]);
                    double Dx = (p->D[s] - (dxb2) * (gp * vxa * (1 - dt * dx1 * gp * vxa))) * dx1_2;
                    double Dy = (p->D[s] - (dxb2) * (gp * vya * (1 - dt * dx1 * gp * vya))) * dx1_2;
                    if (Dx < Dc0) Dx = Dc0;
                    if (Dy < Dc0) Dy = Dc0;
                    const int64_t ux = vx > 0 ? lf : rt, uy = vy > 0 ? up : dn;
                    const double alpha = gam[s] * (Dy * (rs[up] + rs[dn]) + Dx * (rs[lf] + rs[rt]))
                                         + dx1 * (gp) * (vxa * rs[ux] + vya * rs[uy]);
                    if (alpha == 0 && rs[i] == 0) continue;
                    const double beta = -p->M[s] - 2 * gam[s] * (Dx + Dy) - (1 - gam[s]) * (vxa + vya) * dx1;
                    const double lexp = tmo_exp((beta) * dt);
                    const double lam = 2 * (beta) * (sigma2_1[s]) / (lexp - 1.0);
                    int64_t k;
                    y[s] = dornic_cell(&st, alpha, rs[i], lam * lexp, lam, sigma2_1[s], compat, &k);
                }
                tmo_dpc_rk4(p, ns, y);
                for (int s = 0; s < ns; ++s) {
                    if (y[s] < 0 || !isfinite(y[s])) {
                        const int64_t key = (int64_t)s * N + i;
                        if (fault_index && (*fault_index < 0 || key < *fault_index)) *fault_index = key;
                        fault = 1;
                    }
                    fields[s * N + i] = tmo_flush(y[s], p->flush_eps, p->flush_neg);
                }
                for (int s = 1; s < ns; ++s) fields[(ns + s - 1) * N + i] = gam[s];
            }
        if (rows) tmo_reduce(L, nf, fields, rows + 2 * nf * n);
        if (fault) break;
    }
    free(old);
    return fault;
}

/* ---- multi-species replicate statistics -----------------------------------
 * multiSPDP.cpp:3376-3391 (first replicate: per species mean, 0, surviving
 * 1 or 0, active) and var_mean_incremental_surv_runs (MSP:1080-1123: per
 * species, the recurrence over the measurements until that species' first
 * empty one, where its loop breaks).  With flag 2 the replicate's rows after
 * the vegetation's first empty measurement are zero (its time loop ended).
 * out: [n_params][n_meas][4 nsp]. */
void tmo_species_stats(const double* series, int n_inst, int n_meas, int nf, int nsp,
                       const int32_t* params_index, int n_params, int flags, double* out)
{
    const int w = 2 * nf, nc = 4 * nsp;
    const int intdiv = (flags & 1) != 0, stop = (flags & 2) != 0;
    double* x = (double*)malloc(sizeof(double) * (size_t)(n_meas > 0 ? n_meas : 1) * w);
    memset(out, 0, sizeof(double) * (size_t)n_params * n_meas * nc);
    for (int g = 0; g < n_params; ++g) {
        double* tab = out + (int64_t)g * n_meas * nc;
        int j = 0;
        for (int k = 0; k < n_inst; ++k) {
            if (params_index[k] != g) continue;
            int ended = 0;
            for (int i = 0; i < n_meas; ++i) {
                const double* src = series + ((int64_t)k * n_meas + i) * w;
                for (int c = 0; c < w; ++c) x[(int64_t)i * w + c] = ended ? 0.0 : src[c];
                if (stop && src[0] == 0.0) ended = 1;
            }
            for (int s = 0; s < nsp; ++s) {
                if (j == 0) {                                     /* MSP:3380-3390 */
                    for (int i = 0; i < n_meas; ++i) {
                        double* row = tab + (int64_t)i * nc + 4 * s;
                        const double* xi = x + (int64_t)i * w + 2 * s;
                        row[0] = xi[1];
                        row[1] = 0.0;
                        row[3] = xi[0];
                        if (xi[0] > 0) row[2] = 1;
                        else if (xi[0] == 0) row[2] = 0;
                    }
                    continue;
                }
                for (int i = 0; i < n_meas; ++i) {                /* MSP:1087-1120 */
                    double* row = tab + (int64_t)i * nc + 4 * s;
                    const double* xi = x + (int64_t)i * w + 2 * s;
                    const int r = row[2] + 1;
                    if (xi[0] == 0.0) break;
                    const double mean_prev = row[0];
                    if (row[2] == 0.0) {
                        row[0] = xi[1];
                        row[1] = 0.0;
                        row[3] = xi[0];
                    } else {
                        row[0] = (xi[1] + (r - 1) * row[0]) / r;
                        row[3] = (xi[0] + (r - 1) * row[3]) / r;
                        const double inner = (r - 2) * row[1] + (xi[1] - row[0]) * (xi[1] - row[0]);
                        if (intdiv)
                            row[1] = (row[0] - mean_prev) * (row[0] - mean_prev) + (1 / (r - 1)) * inner;
                        else
                            row[1] = (row[0] - mean_prev) * (row[0] - mean_prev) + inner / (r - 1);
                    }
                    row[2] = r;
                }
            }
            ++j;
        }
    }
    free(x);
}

/* ------------------------------------------------------------------ */
/* Consumer movement of the Rietkerk model: the analytic Fokker-Planck  */
/* step of advdiff_FKE_MultiSp (FKE:3614-3695, the CPU twin of the      */
/* reference's CUDA kernel updateMovement, kernel_advdiff_FKE2D.cu:92). */
/* ------------------------------------------------------------------ */

/* precompute_gaussian (FKE:1735-1749): a normalised half-stencil of R + 1 values. */
static void fke_gauss(double sigma, int R, double* st)
{
    const double PI_ = 3.141592653589793;
    double sum = 0.0;
    const double factor = 1.0 / (sqrt(2 * PI_) * sigma);
    for (int x = 0; x <= R; ++x) {
        st[x] = factor * exp(-0.5 * (x * x) / (sigma * sigma));
        sum += (x > 0) ? 2 * st[x] : st[0];
    }
    for (int x = 0; x <= R; ++x) st[x] /= sum;
}

/* set_global_FKE_params (FKE:195-250) for one consumer species, and its two
 * stencils (FKE:3764-3772; both are Gaussians of the species' noise sigma). */
int tmo_fke_move_constants(double D, double v, double sigma, double dt_analytic, double dt_adv,
                           int mv_invariance, double dx, int CI, int32_t* rD, int32_t* rV,
                           double* mu_pref, double* stencil_D, int cap_D, double* stencil_V, int cap_V)
{
    const double Dt_ay = dt_analytic;
    const double beta_vel = mv_invariance == 1 ? 3.0 / dt_adv : 3.0 / dt_analytic;
    const double exp_beta_Dt = exp(-beta_vel * dt_analytic);
    const double exp_2beta_Dt = exp(-2 * beta_vel * dt_analytic);
    const double sigma_v = sqrt(2 * beta_vel * 1 * fabs(v));
    const double sigma_D = sqrt(2 * D * dt_analytic);
    const double sigma_vx = (sigma_v / beta_vel) * sqrt((Dt_ay + (1 / (2 * beta_vel)) * (1 - exp_2beta_Dt)
                                                         - (2 / beta_vel) * (1 - exp_beta_Dt)));
    *rD = (int)(CI * sigma_D / dx);
    *rV = (int)(CI * sigma_vx / dx);
    *mu_pref = (1 / dx) * (Dt_ay - ((1 - exp_beta_Dt) / beta_vel));
    if (*rD + 1 > cap_D || *rV + 1 > cap_V) return -1;
    fke_gauss(sigma, *rD, stencil_D);
    fke_gauss(sigma, *rV, stencil_V);
    return 0;
}

/* One species: out = the scatter of every cell with rho >= 1e-10 (FKE:3626-3628;
 * epsilon at rietkerk_bjork_basic.h:119) through the product of its row and
 * column kernels, gamma^2 (diffusion x diffusion), gamma - gamma^2 (each mixed
 * pair) and (1 - gamma)^2 (advection x advection), centred on the cell for
 * diffusion and on its advected mean for advection.  The reference's advection
 * bounds read .first below the mean and .second above (FKE:3640-3641); both
 * are int(CI sigma_v / dx) of the same sigma, one rV here. */
void tmo_fke_movement(int L, const double* rho, const double* gamma, const double* vx, const double* vy,
                      int rD, int rV, double mu_pref, const double* SD, const double* SV, double* out)
{
    const int64_t L2 = (int64_t)L * L;
    for (int64_t i = 0; i < L2; ++i) out[i] = 0.0;
    for (int64_t i = 0; i < L2; ++i) {
        if (rho[i] < 1.0e-10) continue;
        const int c_i = (int)(i / L), c_j = (int)(i % L);
        const double g = gamma[i], g2 = g * g, r = rho[i];
        const int min_p = c_i - rD, max_p = c_i + rD, min_q = c_j - rD, max_q = c_j + rD;
        const int mu_x = (int)(round(c_i + vx[i] * (mu_pref)) + L) % L;
        const int mu_y = (int)(round(c_j + vy[i] * (mu_pref)) + L) % L;
        const int min_p_vx = mu_x - rV, max_p_vx = mu_x + rV, min_q_vy = mu_y - rV, max_q_vy = mu_y + rV;
        for (int p = min_p; p <= max_p; ++p) {
            const int dxx = abs(p - c_i);
            for (int q = min_q; q <= max_q; ++q) {
                const int dyy = abs(q - c_j);
                out[(int64_t)((p + L) % L) * L + (q + L) % L] += (g2) * r * (SD[dxx] * SD[dyy]);
            }
            for (int q = min_q_vy; q <= max_q_vy; ++q) {
                const int dyy = abs(q - mu_y);
                out[(int64_t)((p + L) % L) * L + (q + L) % L] += (g - g2) * r * (SD[dxx] * SV[dyy]);
            }
        }
        for (int p = min_p_vx; p <= max_p_vx; ++p) {
            const int dxx = abs(p - mu_x);
            for (int q = min_q; q <= max_q; ++q) {
                const int dyy = abs(q - c_j);
                out[(int64_t)((p + L) % L) * L + (q + L) % L] += (g - g2) * r * (SV[dxx] * SD[dyy]);
            }
            for (int q = min_q_vy; q <= max_q_vy; ++q) {
                const int dyy = abs(q - mu_y);
                out[(int64_t)((p + L) % L) * L + (q + L) % L] += (1 - g) * (1 - g) * r * (SV[dxx] * SV[dyy]);
            }
        }
    }
}

/* Perception and velocities of the Rietkerk model's consumers:
 * calc_gamma_vel_2Sp_NonRefugia (FKE:2534-2606) and calc_gamma_vel_3Sp_NonRefugia
 * (FKE:3371-3482), from the old densities rho[ns][L*L] (vegetation first).
 * gamma[ns-1][L*L] (consumers 1 .. ns-1) is updated in place; both forms add onto
 * the previous gamma.  With three species every species s (vegetation
 * included) draws a velocity v[s] (cos, sin)(2 pi u) per cell, u = word 0 of
 * Philox(seed, inst; cell, step, block (s << 24) | 0x800000), into veff[s][cell];
 * the two-species form draws none.  Spatial means by tmo_fixed_sum; the cover
 * factor and neighbourhood counts as in the DP consumer step, with the
 * three-species form's floor of one (FKE:3386, :3416). */
void tmo_fke_perception(int ns, int L, const double* pR, double dx, const double* v, const int32_t* disk,
                        int disk_len, uint32_t seed, uint32_t inst, int64_t step, const double* rho,
                        double* gamma, double* veff)
{
    const int64_t N = (int64_t)L * L;
    const double eps_g = 1.0e-12;
    double mean[3], inv[3], fr[3];
    for (int s = 0; s < ns; ++s) mean[s] = tmo_fixed_sum(rho + s * N, N) / (double)N;
    double occ = 0.0;
    for (int64_t i = 0; i < N; ++i) if (rho[i] > 0) occ += 1;
    double nfac = 1 - occ / (double)N;
    if (nfac < 0.35) nfac = 0.35;
    int new_len = (int)(nfac * nfac * (double)disk_len);
    if (ns == 3 && new_len < 1) new_len = 1;
    for (int s = 0; s < ns; ++s) inv[s] = mean[s] >= eps_g ? 1 / mean[s] : 1 / eps_g;
    double rmax = pR[0];
    for (int s = 1; s < ns; ++s) if (pR[s] > rmax) rmax = pR[s];
    rmax = rmax / dx;
    for (int s = 0; s < ns; ++s) fr[s] = pR[s] / (dx * rmax);
    for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q) {
            const int64_t i = (int64_t)p * L + q;
            if (ns == 3)
                for (int s = 0; s < ns; ++s) {
                    stream_t st;
                    st_init(&st, seed, inst, (uint32_t)i, step);
                    st.ctr[3] = ((uint32_t)s << 24) | 0x800000u;
                    double sn, cs;
                    tmo_sincos2pi(st_u(&st), &sn, &cs);
                    veff[2 * (s * N + i)] = v[s] * cs;
                    veff[2 * (s * N + i) + 1] = v[s] * sn;
                }
            for (int s = 1; s < ns; ++s) {
                double* g = gamma + (s - 1) * N + i;
                if (fr[s] == 0.0) continue;
                int eff = (int)(fr[s] * fr[s] * (double)new_len);
                if (ns == 3 && eff < 1 && fr[s] > 0) eff = 1;
                if (mean[s] < eps_g) { *g = 1.0; continue; }
                if (mean[s - 1] < eps_g) { *g = 0.0; continue; }
                const double* src = rho + (s - 1) * N;
                double acc = *g;
                for (int k = 0; k < eff; ++k) {
                    const int64_t j = wrapi(p + disk[2 * k], L) * L + wrapi(q + disk[2 * k + 1], L);
                    if (ns == 3 && s == 1 && !(mean[2] < eps_g)) acc += src[j] * (1 - rho[2 * N + j] * inv[2]);
                    else acc += src[j];
                }
                if (ns == 2 || s == 1) acc = (acc * inv[s - 1]) / (eff);
                else acc = (acc / (eff)) * inv[s - 1];
                if (acc > 1) acc = 1.0;
                else if (acc < 0) acc = 0.0;
                *g = acc;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* The Rietkerk consumer model's RK4 (FKE:2449-2533, :2611-2720).            */
/* ------------------------------------------------------------------------ */

/* f_DorFKE_2Sp (ns = 1, fields P G W O) and f_DorFKE_3Sp (ns = 2, P G Pr W O). */
static void fke_rhs(const tmo_fke_rk_params* p, int ns, int L, const double* R, double* f)
{
    const int64_t L2 = (int64_t)L * L;
    const double dx2 = p->dx * p->dx;
    const double dW = p->D_W / (dx2), dO = p->D_O / (dx2);   /* diff_coefficient, FKE:3849-3852 */
    const double cgmax = p->c * p->gmax, K2W0 = p->K[2] * p->W0, A01H01 = p->A[0][1] * p->H[0][1];
    const double A12H12 = p->A[1][2] * p->H[1][2];
    const double EA1 = p->E[1] * p->A[0][1], EA2 = p->E[2] * p->A[1][2];
    const int w = ns + 1, o = ns + 2;
    const double* Pv = R;
    const double* Gv = R + L2;
    const double* Wv = R + w * L2;
    const double* Ov = R + o * L2;
    for (int64_t i = 0; i < L2; ++i) {
        const int y = (int)(i / L), x = (int)(i % L);
        /* nR2[i]: i - g, i + g, i - 1, i + 1 with periodic wrap (FKE:3821) */
        const int64_t n00 = (int64_t)((y + L - 1) % L) * L + x, n01 = (int64_t)((y + 1) % L) * L + x;
        const int64_t n10 = (int64_t)y * L + (x + L - 1) % L, n11 = (int64_t)y * L + (x + 1) % L;
        f[i] = cgmax * Wv[i] * Pv[i] / (Wv[i] + p->K[1]) - (p->A[0][1] * Pv[i] / (1 + A01H01 * Pv[i])) * Gv[i];
        if (ns == 1) {
            f[L2 + i] = (EA1 * Pv[i] / (1 + A01H01 * Pv[i])) * Gv[i] - p->M[1] * Gv[i];
        } else {
            const double* Qv = R + 2 * L2;
            f[L2 + i] = (EA1 * Pv[i] / (1 + A01H01 * Pv[i])) * Gv[i] - p->M[1] * Gv[i]
                        - p->A[1][2] * Gv[i] / (1 + A12H12 * Gv[i]) * Qv[i];
            f[2 * L2 + i] = (EA2 * Gv[i] / (1 + A12H12 * Gv[i])) * Qv[i] - p->M[2] * Qv[i];
        }
        f[w * L2 + i] = p->alpha * (Pv[i] + K2W0) / (Pv[i] + p->K[2]) * Ov[i]
                        - p->gmax * Wv[i] * Pv[i] / (Wv[i] + p->K[1]) - p->rW * Wv[i]
                        + (dW) * (Wv[n00] + Wv[n01] + Wv[n10] + Wv[n11] - 4 * Wv[i]);
        f[o * L2 + i] = p->a - p->alpha * (Pv[i] + K2W0) / (Pv[i] + p->K[2]) * Ov[i]
                        + (dO) * (Ov[n00] + Ov[n01] + Ov[n10] + Ov[n11] - 4 * Ov[i]);
    }
}

int64_t tmo_fke_rk4(const tmo_fke_rk_params* p, int ns, int L, double* y)
{
    const int64_t n = (int64_t)(ns + 3) * L * L;
    const double dt2 = p->dt_rk / 2.0, dt6 = p->dt_rk / 6.0, dt = p->dt;
    double* K1 = (double*)malloc(sizeof(double) * n);
    double* K2 = (double*)malloc(sizeof(double) * n);
    double* K3 = (double*)malloc(sizeof(double) * n);
    double* K4 = (double*)malloc(sizeof(double) * n);
    double* T = (double*)malloc(sizeof(double) * n);
    int64_t bad = 0;
    fke_rhs(p, ns, L, y, K1);
    for (int64_t j = 0; j < n; ++j) T[j] = y[j] + (dt2) * K1[j];
    fke_rhs(p, ns, L, T, K2);
    for (int64_t j = 0; j < n; ++j) T[j] = y[j] + (dt2) * K2[j];
    fke_rhs(p, ns, L, T, K3);
    for (int64_t j = 0; j < n; ++j) T[j] = y[j] + (dt) * K3[j];
    fke_rhs(p, ns, L, T, K4);
    for (int64_t j = 0; j < n; ++j) {
        y[j] += (dt6) * (K1[j] + 2.0 * K2[j] + 2.0 * K3[j] + K4[j]);
        if (y[j] < 0 || !isfinite(y[j])) ++bad;
    }
    free(K1);
    free(K2);
    free(K3);
    free(K4);
    free(T);
    return bad;
}

/* ------------------------------------------------------------------------ */
/* The Rietkerk consumer model as one step: its vegetation Dornic step, the */
/* movement in separable form and the movement statistics.               */
/* ------------------------------------------------------------------------ */

/* FKE:4150-4211: alpha_i from the old P (nR2 neighbours), the skip of empty
 * cells with empty neighbours, the transition of tmo_rk_run's P step. */
void tmo_fke_veg_dornic(const tmo_rk_params* p, int L, uint32_t seed, uint32_t inst, int64_t step,
                        const double* P, double* Pstar)
{
    const double dx2 = p->dx * p->dx;
    double cst[3];
    tmo_dornic_constants(-p->d, p->D[0], p->sigma, p->dt, p->dx, cst);   /* FKE:3847-3849 */
    const double lambda = cst[2], ll = cst[2] * cst[1];
    const double Dp = p->D[0] / (dx2);
    const double sigma2_1 = 1 / (p->sigma * p->sigma);
    for (int pp = 0; pp < L; ++pp)
        for (int q = 0; q < L; ++q) {
            const int64_t i = (int64_t)pp * L + q;
            const double alpha_i = Dp * (nsum(P, L, pp, q));
            Pstar[i] = P[i];
            if (!(alpha_i == 0 && P[i] == 0)) {
                stream_t s;
                st_init(&s, seed, inst, (uint32_t)i, step);
                int64_t draw = -1;
                Pstar[i] = dornic_cell(&s, alpha_i, P[i], ll, lambda, sigma2_1, p->sampler == 1, &draw);
            }
        }
}

void tmo_fke_movement_sep(int L, const double* rho, const double* gamma, const double* vx, const double* vy,
                          int rD, int rV, int S, double mu_pref, const double* SD, const double* SV, double* out)
{
    const int64_t L2 = (int64_t)L * L;
    double* cx = (double*)malloc(sizeof(double) * L2 * 3);
    int* mx = (int*)malloc(sizeof(int) * L2 * 2);
    double* y4 = (double*)malloc(sizeof(double) * L2 * 4);   /* cx, Y1, Y2, Z */
    double* pq = (double*)malloc(sizeof(double) * L2 * 2);   /* P, Q */
    for (int64_t i = 0; i < L2; ++i) {           /* each source's terms and advected mean (FKE:3626-3637) */
        if (rho[i] < 1.0e-10) {
            mx[2 * i] = -1;
            continue;
        }
        const int c_i = (int)(i / L), c_j = (int)(i % L);
        const double g = gamma[i], g2 = g * g, r = rho[i];
        mx[2 * i] = (int)(round(c_i + vx[i] * (mu_pref)) + L) % L;
        mx[2 * i + 1] = (int)(round(c_j + vy[i] * (mu_pref)) + L) % L;
        cx[3 * i] = (g2) * r;
        cx[3 * i + 1] = (g - g2) * r;
        cx[3 * i + 2] = (1 - g) * (1 - g) * r;
    }
    /* displacement sums over the (2S + 1)^2 sources around (p, q), rows then
     * columns ascending: Y1 over row p (sources whose my = q), Y2 over column q
     * (mx = p), Z over all (mx = p and my = q) */
    for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q) {
            const int64_t o = (int64_t)p * L + q;
            double c0 = 0.0, a1 = 0.0, a2 = 0.0, z = 0.0;
            if (mx[2 * o] >= 0) c0 = cx[3 * o];
            for (int di = -S; di <= S; ++di) {
                const int ci = wrapi(p + di, L);
                for (int dj = -S; dj <= S; ++dj) {
                    const int cj = wrapi(q + dj, L);
                    const int64_t i = (int64_t)ci * L + cj;
                    if (mx[2 * i] < 0) continue;
                    const int hx = mx[2 * i] == p, hy = mx[2 * i + 1] == q;
                    if (di == 0 && hy) a1 += cx[3 * i + 1];
                    if (dj == 0 && hx) a2 += cx[3 * i + 1];
                    if (hx && hy) z += cx[3 * i + 2];
                }
            }
            y4[4 * o] = c0;
            y4[4 * o + 1] = a1;
            y4[4 * o + 2] = a2;
            y4[4 * o + 3] = z;
        }
    /* along each row: P = SD * cx + SV * Y1, Q = SD * Y2 + SV * Z (taps ascending) */
    for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q) {
            double P = 0.0, Q = 0.0;
            for (int b = -rD; b <= rD; ++b) {
                const int64_t i = (int64_t)p * L + wrapi(q - b, L);
                P += SD[abs(b)] * y4[4 * i];
                Q += SD[abs(b)] * y4[4 * i + 2];
            }
            for (int b = -rV; b <= rV; ++b) {
                const int64_t i = (int64_t)p * L + wrapi(q - b, L);
                P += SV[abs(b)] * y4[4 * i + 1];
                Q += SV[abs(b)] * y4[4 * i + 3];
            }
            pq[2 * ((int64_t)p * L + q)] = P;
            pq[2 * ((int64_t)p * L + q) + 1] = Q;
        }
    /* along each column: out = SD * P + SV * Q (taps ascending) */
    for (int x = 0; x < L; ++x)
        for (int y = 0; y < L; ++y) {
            double acc = 0.0;
            for (int a = -rD; a <= rD; ++a) acc += SD[abs(a)] * pq[2 * ((int64_t)wrapi(x - a, L) * L + y)];
            for (int a = -rV; a <= rV; ++a) acc += SV[abs(a)] * pq[2 * ((int64_t)wrapi(x - a, L) * L + y) + 1];
            out[(int64_t)x * L + y] = acc;
        }
    free(cx);
    free(mx);
    free(y4);
    free(pq);
}

void tmo_fke_mov_stats(int L, const double* gamma, const double* vx, const double* vy, double out[4])
{
    const int64_t N = (int64_t)L * L;
    const double mean = tmo_fixed_sum(gamma, N) / (double)N;
    double* d = (double*)malloc(sizeof(double) * N);
    for (int64_t i = 0; i < N; ++i) d[i] = (gamma[i] - mean) * (gamma[i] - mean);
    out[0] = mean;
    out[1] = tmo_fixed_sum(d, N) / (double)N;
    out[2] = tmo_fixed_sum(vx, N) / (double)N;
    out[3] = tmo_fixed_sum(vy, N) / (double)N;
    free(d);
}
