// rietkerk_loop.cpp -- a C++ caller of include/tm/tm.h, built with the host
// compiler and linked to libtm.so (no HIP headers): the drop-in body of the
// Rietkerk loop of rietkerk_DorFPE_2D_MultiSp (Rietkerk_FPE/rietkerk_bjork_FKE.cpp:
// 3887-4468) for r replicates, run between the reference's measurement times.
//
//   rietkerk_loop <fields.bin> <out.bin> <g> <r> <seed> <t_max>
//
// fields.bin: the initial frames, FP64 [r][3][g*g] (species-major, as Rho_dt).
// out.bin: FP64 n_meas, the measured steps, Rho_M [r][n_meas][6] (occupied count
// and mean of P, W, O after each measured step, as FKE:4024), the fault record
// (instance, cell, step), then the final fields [r][3][g*g].
#include <cstdint>
#include <cstdio>
#include <cstdlib>
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
        std::fprintf(stderr, "usage: %s fields.bin out.bin g r seed t_max\n", argv[0]);
        return 2;
    }
    const int g = std::atoi(argv[3]), r = std::atoi(argv[4]);
    const uint32_t seed = (uint32_t)std::strtoul(argv[5], nullptr, 10);
    const double t_max = std::atof(argv[6]);
    const size_t cells = (size_t)g * g;
    std::vector<double> host(3 * cells * r);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) return 3;
    std::fclose(f);

    // the reference's arguments (rietkerk_bjork_basic.h:450-452), values of the C2 configuration
    const double a = 1.1, c = 10.0, gmax = 0.05, alpha = 0.2, rW = 0.2, W0 = 0.2, dt = 0.005, dx = 2.0;
    const double D[3] = {0.1, 0.1, 100.0}, K[3] = {0.0, 5.0, 5.0}, sigma[1] = {0.5}, M[1] = {0.25};

    tm_ctx* ctx = nullptr;
    if (tm_create(&ctx, 0) != TM_OK) return die(ctx, "tm_create");
    tm_rietkerk_params p = {};
    p.R = a; p.c = c; p.gmax = gmax; p.alpha = alpha; p.rW = rW; p.W0 = W0; p.d = M[0];
    for (int s = 0; s < 3; ++s) { p.D[s] = D[s]; p.K[s] = K[s]; }
    p.sigma = sigma[0]; p.dt = dt; p.dx = dx;
    p.flush_eps = 1e-10;                 // rietkerk_bjork_basic.h:119
    p.sampler = TM_SAMPLER_EXACT;
    std::vector<tm_instance_desc> inst(r);
    for (int j = 0; j < r; ++j) inst[j] = {g, 0, seed, (uint32_t)j, 0};
    if (tm_configure(ctx, TM_MODEL_RIETKERK, TM_F64, &p, 1, inst.data(), r) != TM_OK) return die(ctx, "tm_configure");
    for (int j = 0; j < r; ++j)
        if (tm_upload_fields(ctx, j, host.data() + (size_t)j * 3 * cells) != TM_OK) return die(ctx, "tm_upload_fields");

    // the reference's measurement times and the steps its loop logs them at
    const int n_t = tm_log10_time_bins(t_max, dt, nullptr, 0);
    std::vector<double> t_meas(n_t);
    tm_log10_time_bins(t_max, dt, t_meas.data(), n_t);
    std::vector<int64_t> meas_step(n_t);
    const int n_meas = tm_measure_steps(t_meas.data(), n_t, dt, meas_step.data());

    std::vector<double> rho_m((size_t)r * n_meas * 6);
    std::vector<double> rows;
    int64_t done = 0;
    for (int m = 0; m < n_meas; ++m) {
        const int64_t n = meas_step[m] - done;
        if (n > 0) {
            rows.resize((size_t)r * n * 6);
            const tm_status st = tm_run(ctx, n, rows.data());
            if (st != TM_OK && st != TM_FAULT_NUMERIC) return die(ctx, "tm_run");
            done = meas_step[m];
            for (int j = 0; j < r; ++j)
                for (int k = 0; k < 6; ++k)
                    rho_m[((size_t)j * n_meas + m) * 6 + k] = rows[((size_t)j * n + (n - 1)) * 6 + k];
        } else {   // a measurement at the current state
            std::vector<double> cur((size_t)r * 6);
            if (tm_measure(ctx, cur.data()) != TM_OK) return die(ctx, "tm_measure");
            for (int j = 0; j < r; ++j)
                for (int k = 0; k < 6; ++k) rho_m[((size_t)j * n_meas + m) * 6 + k] = cur[(size_t)j * 6 + k];
        }
    }
    int32_t fi = -1;
    int64_t fc = -1, fs = -1;
    if (tm_fault_info(ctx, &fi, &fc, &fs) != TM_OK) return die(ctx, "tm_fault_info");
    for (int j = 0; j < r; ++j)
        if (tm_download_fields(ctx, j, host.data() + (size_t)j * 3 * cells) != TM_OK)
            return die(ctx, "tm_download_fields");
    tm_destroy(ctx);

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 4;
    const double hdr = (double)n_meas;
    std::fwrite(&hdr, sizeof(double), 1, o);
    for (int m = 0; m < n_meas; ++m) {
        const double s = (double)meas_step[m];
        std::fwrite(&s, sizeof(double), 1, o);
    }
    std::fwrite(rho_m.data(), sizeof(double), rho_m.size(), o);
    const double fr[3] = {(double)fi, (double)fc, (double)fs};
    std::fwrite(fr, sizeof(double), 3, o);
    std::fwrite(host.data(), sizeof(double), host.size(), o);
    std::fclose(o);
    return 0;
}
