// ref_ssp_harness.cpp -- C entry points around the reference's own legacy DP
// Dornic integrator, its running replicate statistics and its measurement
// times, built from the reference source where it lies
// (Percolation_FD/stochasticSPDP.cpp) into oracle/_ref/libref_ssp.so by
// oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used for the ensemble parity
// test of the DP step and as the "reference" CPU baseline of bench.py.
//
// percolationDornic_2D (stochasticSPDP.cpp:940-1143) integrates
//   d rho = (a rho - b rho^2 + D lap rho) dt + sigma sqrt(rho) dW
// with the Dornic Poisson-Gamma linear/noise step (stochasticSPDP.cpp:1032-1061,
// libstdc++ poisson/gamma distributions, no Gaussian shortcuts) and the exact
// nonlinear update rho/(1 + b rho dt) (stochasticSPDP.cpp:1064-1068).  Its
// stream is std::mt19937_64 seeded from std::random_device (stochasticSPDP.cpp:973),
// so agreement with it is statistical only.  Output rows (stochasticSPDP.cpp:1130-1141):
//   a, L, t, <<rho>>_surviving, Var, #surviving replicates, <#active>, surv_rep.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <thread>
#include <vector>

using std::vector;

// The definition's signature (stochasticSPDP.cpp:940-941); SPDP.h:96-97 declares a stale one.
void percolationDornic_2D(vector<vector<double>>& Rho, vector<double>& t_meas, double Rh0[],
                          double t_max, double a, double b, double D, double sigma, double dt,
                          double dx, int r, int g);

// Running statistics over surviving replicates (stochasticSPDP.cpp:173-207) and
// the log-spaced measurement times (stochasticSPDP.cpp:336-377), as defined there.
void var_mean_incremental_surv_runs(double t_avg_var_rep_N[][5], double X_curr[][2], int size);
std::vector<double> logarithm10_time_bins(double t_max, double dt);

extern "C" {

// One replicate folded into the running table: tab[size][5] = (t, mean, var,
// #surviving, mean #active), x[size][2] = (#active, mean) of the replicate.
void ref_ssp_var_mean_surv(double* tab, double* x, int size)
{
    var_mean_incremental_surv_runs(reinterpret_cast<double(*)[5]>(tab), reinterpret_cast<double(*)[2]>(x), size);
}

// The reference's measurement times; returns their number, writes at most cap.
int ref_ssp_log10_time_bins(double t_max, double dt, double* out, int cap)
{
    const vector<double> t = logarithm10_time_bins(t_max, dt);
    std::cout.flush();
    fflush(stdout);
    for (size_t i = 0; i < t.size() && (int)i < cap; ++i) out[i] = t[i];
    return (int)t.size();
}

// One call of the reference integrator: r replicates of one a-value from the
// initial frame rho0 (g*g).  Rows are written to out[max_rows][8]; returns the
// number of rows, or -1 when max_rows is too small.
int ref_ssp_percolation_dornic(int g, double a, double b, double D, double sigma, double dt,
                               double dx, int r, const double* rho0, const double* t_meas,
                               int n_meas, double t_max, double* out, int max_rows)
{
    vector<vector<double>> rows;
    vector<double> tm(t_meas, t_meas + n_meas);
    vector<double> r0(rho0, rho0 + (size_t)g * g);
    percolationDornic_2D(rows, tm, r0.data(), t_max, a, b, D, sigma, dt, dx, r, g);
    std::cout.flush();   // the integrator's progress lines go out while the caller's redirect holds
    fflush(stdout);
    if ((int)rows.size() > max_rows) return -1;
    for (size_t i = 0; i < rows.size(); ++i)
        for (size_t j = 0; j < 8; ++j) out[i * 8 + j] = j < rows[i].size() ? rows[i][j] : 0.0;
    return (int)rows.size();
}

// nthreads concurrent, independent calls (the reference's thread-per-parameter
// parallelism, rietkerk_bjork_FKE.cpp:4778-4811); rows are discarded.  Used to
// time the reference on several host cores.
int ref_ssp_parallel(int nthreads, int g, double a, double b, double D, double sigma, double dt,
                     double dx, int r, const double* rho0, const double* t_meas, int n_meas,
                     double t_max)
{
    vector<std::thread> th;
    for (int k = 0; k < nthreads; ++k)
        th.emplace_back([=]() {
            vector<vector<double>> rows;
            vector<double> tm(t_meas, t_meas + n_meas);
            vector<double> r0(rho0, rho0 + (size_t)g * g);
            percolationDornic_2D(rows, tm, r0.data(), t_max, a, b, D, sigma, dt, dx, r, g);
        });
    for (auto& t : th) t.join();
    std::cout.flush();
    fflush(stdout);
    return 0;
}

// The time steps one call executed, summed over its replicates, from its rows:
// a replicate stops at the first logged time with no occupied cell
// (stochasticSPDP.cpp:1019-1020), so the survivors at each logged time times the
// steps to the next logged time count the steps run (to within one logging
// interval per replicate).
static double steps_of_rows(const vector<vector<double>>& rows, double dt)
{
    double steps = 0.0;
    for (size_t i = 0; i + 1 < rows.size(); ++i)
        steps += rows[i][5] * (double)(long)((rows[i + 1][2] - rows[i][2]) / dt + 0.5);
    if (!rows.empty()) steps += rows.back()[5];
    return steps;
}

// As ref_ssp_parallel, returning the time steps executed by all calls together
// (replicate-steps), for a rate in cell-updates per second.
double ref_ssp_parallel_steps(int nthreads, int g, double a, double b, double D, double sigma, double dt,
                              double dx, int r, const double* rho0, const double* t_meas, int n_meas,
                              double t_max)
{
    vector<std::thread> th;
    vector<double> steps(nthreads, 0.0);
    for (int k = 0; k < nthreads; ++k)
        th.emplace_back([=, &steps]() {
            vector<vector<double>> rows;
            vector<double> tm(t_meas, t_meas + n_meas);
            vector<double> r0(rho0, rho0 + (size_t)g * g);
            percolationDornic_2D(rows, tm, r0.data(), t_max, a, b, D, sigma, dt, dx, r, g);
            steps[k] = steps_of_rows(rows, dt);
        });
    for (auto& t : th) t.join();
    std::cout.flush();
    fflush(stdout);
    double total = 0.0;
    for (double v : steps) total += v;
    return total;
}

}  // extern "C"
