// fke_loop.cpp -- a C++ caller of include/tm/tm.h, built with the host compiler and
// linked to libtm.so (no HIP headers): the Rietkerk model with grazer and predator
// (rietkerk_DorFPE_2D_MultiSp built with SPB = 3, loop body
// Rietkerk_FPE/rietkerk_bjork_FKE.cpp:4100-4462) for r replicates, with the driver's
// constants (order_3stoc_test_rietkerk.cpp:20-89), stepping n steps and reading the
// movement statistics (Rho_Mov, FKE:4034-4060) at the end.
//
//   fke_loop <fields.bin> <out.bin> <g> <r> <seed> <n_steps>
//
// fields.bin: the initial fields, FP64 [r][7][g*g] (P, G, Pr, W, O, gamma_1, gamma_2).
// out.bin: FP64 rows [r][n][14], Rho_Mov [r][12], the fault record (instance, cell,
// step), then the final fields [r][7][g*g].
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "tm/tm.h"

static int die(tm_ctx* ctx, const char* what)
{
    std::fprintf(stderr, "%s: %s\n", what, ctx ? tm_last_error(ctx) : "no context");
    if (ctx) tm_destroy(ctx);
    return 1;
}

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s fields.bin out.bin g r seed n_steps\n", argv[0]);
        return 2;
    }
    const int g = std::atoi(argv[3]), r = std::atoi(argv[4]);
    const uint32_t seed = (uint32_t)std::strtoul(argv[5], nullptr, 10);
    const int64_t n = std::atoll(argv[6]);
    constexpr int NF = 7, SpB = 3, Sp = 5;
    const size_t cells = (size_t)g * g;
    std::vector<double> host(NF * cells * r);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) return 3;
    std::fclose(f);

    // the driver's values (km, hr, kg), as rietkerk_DorFPE_2D_MultiSp receives them
    const double c = 10000, gmax = 0.05 * std::pow(10, -3.0) / 24.0, d = 0.25 / 24.0, alpha = 0.2 / 24.0;
    const double W0 = 0.2, rW = 0.2 / 24.0, dx = 0.1, dt = 0.1, R = 1.0 / 24.0;
    const double D[Sp] = {0.00025 / 24.0, 0.01028, 0.215965, 0.00025 / 24.0, 0.025 / 24.0};
    const double K[3] = {0, 5, 5000};
    const double sigma[SpB] = {std::sqrt(D[0] / (dx * dx)), 1, 10};
    const double v[SpB] = {0, 0.45427, 0.40587};
    const double dtAdv[SpB] = {0, 0.11817455, 0.765270868};
    const double aij = 3.6 * std::pow(10.0, -6.08) * std::pow(20.0, -0.37);
    const double ajm = 3.6 * std::pow(10.0, -6.08) * std::pow(100.0, -0.37);
    const double A[SpB][SpB] = {{0, aij, 0}, {aij, 0, ajm}, {0, ajm, 0}};
    const double H[SpB][SpB] = {{0, 1, 0}, {1, 0, 1}, {0, 1, 0}};
    const double E[SpB] = {1.0, 0.45, 0.85};
    const double M[SpB] = {d, 0.061609 * std::pow(20.0, -0.25) / 8760.0, 0.061609 * std::pow(100.0, -0.25) / 8760.0};
    const double pR[SpB] = {0.0, 1.04285 / 2.0, 1.25536 / 2.0};
    const double Dt_Analytic = dtAdv[1];   // the smallest positive dtAdv (FKE:197-203)

    tm_ctx* ctx = nullptr;
    if (tm_create(&ctx, 0) != TM_OK) return die(ctx, "tm_create");
    tm_fke_params p = {};
    p.a = R; p.c = c; p.gmax = gmax; p.alpha = alpha; p.rW = rW; p.W0 = W0;
    for (int k = 0; k < 3; ++k) p.K[k] = K[k];
    std::memcpy(p.A, A, sizeof p.A);
    std::memcpy(p.H, H, sizeof p.H);
    for (int s = 0; s < SpB; ++s) { p.E[s] = E[s]; p.M[s] = M[s]; p.v[s] = v[s]; p.sigma[s] = sigma[s]; p.pR[s] = pR[s]; }
    p.D_P = D[0]; p.D_W = D[Sp - 2]; p.D_O = D[Sp - 1];
    for (int s = 1; s < SpB; ++s) { p.D_mov[s] = D[s]; p.dt_adv[s] = dtAdv[s]; }
    p.dt = dt; p.dt_analytic = Dt_Analytic; p.dx = dx;
    p.flush_eps = 1e-10;                  // rietkerk_bjork_basic.h:119
    p.CI = 4; p.mv_invariance = 0; p.sampler = TM_SAMPLER_EXACT;
    std::vector<tm_instance_desc> inst(r);
    for (int j = 0; j < r; ++j) inst[j] = {g, 0, seed, (uint32_t)j, 0};
    if (tm_configure(ctx, TM_MODEL_FKE3, TM_F64, &p, 1, inst.data(), r) != TM_OK) return die(ctx, "tm_configure");
    for (int j = 0; j < r; ++j)
        if (tm_upload_fields(ctx, j, host.data() + (size_t)j * NF * cells) != TM_OK) return die(ctx, "tm_upload");

    std::vector<double> rows((size_t)r * n * 2 * NF);
    const tm_status st = tm_run(ctx, n, rows.data());
    if (st != TM_OK && st != TM_FAULT_NUMERIC) return die(ctx, "tm_run");
    std::vector<double> mov((size_t)r * 4 * SpB);
    if (tm_mov_stats(ctx, mov.data()) != TM_OK) return die(ctx, "tm_mov_stats");
    int32_t fi = -1;
    int64_t fc = -1, fs = -1;
    if (tm_fault_info(ctx, &fi, &fc, &fs) != TM_OK) return die(ctx, "tm_fault_info");
    for (int j = 0; j < r; ++j)
        if (tm_download_fields(ctx, j, host.data() + (size_t)j * NF * cells) != TM_OK) return die(ctx, "tm_download");
    tm_destroy(ctx);

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 4;
    const double fault[3] = {(double)fi, (double)fc, (double)fs};
    std::fwrite(rows.data(), sizeof(double), rows.size(), o);
    std::fwrite(mov.data(), sizeof(double), mov.size(), o);
    std::fwrite(fault, sizeof(double), 3, o);
    std::fwrite(host.data(), sizeof(double), host.size(), o);
    std::fclose(o);
    return 0;
}
