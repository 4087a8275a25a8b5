// tm_launch.h -- host-side launch table for the step kernels.  Each kernel
// variant is one (tile width, tile height, threads) instantiation; the host
// context picks one by default or through the autotuner.
#pragma once
#include <hip/hip_runtime.h>
#include "tm_kernels.h"

namespace tmk {

typedef void (*StepLaunch)(const StepArgs&, int grid, hipStream_t);

struct Variant {
    int tx, ty, nt;
    int wg_per_cu;   // > 0: persistent grid of wg_per_cu workgroups per CU
    int pw;          // rho(t) partials written per tile
    const char* name;
    StepLaunch fn;
    int ax, ay;      // split Rietkerk step: tile shape of the Dornic kernel fn_a,
    StepLaunch fn_a; // launched (one tile per workgroup) before fn; 0 = fused step
    int row_major;   // tile order within an instance: 0 = down vertical strips, 1 = along rows
    int spl = 1;     // steps per launch of fn (> 1: temporally blocked DP step)
    StepLaunch fn_one = nullptr;   // spl > 1: the single-step kernel of the same tiles
    int adaptive = 0;              // spl > 1, > 0: single steps while more than this per-mille of
                                   // the cells are occupied, spl-step launches below (tm_api.cpp)
};

// Variant tables per (model, precision); defined in tm_kernels_f64.hip / _f32.hip.
const Variant* dp_variants_f64(int* n);
const Variant* rk_variants_f64(int* n);
const Variant* dp_variants_f32(int* n);
const Variant* rk_variants_f32(int* n);

// Vegetation + consumer models (tm_consumers.hip), FP64 only.
const Variant* dpc2_variants_f64(int* n);
const Variant* dpc3_variants_f64(int* n);

// Consumer models: per-instance tables of the step (device), set at configure
// (DPCInstInfo) or rewritten at the start of every step (DPCStepState).
constexpr int kDpcChunkCells = 4096;   // cells per statistics chunk
struct DPCInstInfo {
    int32_t R;          // perception radius in sites, (int)(max pR / dx)
    int32_t disk_off;   // first offset of its perception table in DPCWork::disk
    int32_t disk_len;   // offsets within radius R
    int32_t chunk0;     // first statistics chunk of the instance
    int32_t nchunks;
    int32_t pad;
    double fr[3];       // pR[s] / (dx r_max): perception radius relative to the largest
};
struct DPCStepState {
    double mean[3];     // spatial means of the densities at the step's start
    double inv[3];      // 1 / mean, or 1 / eps below eps
    int32_t eff[3];     // perception neighbours each species uses this step
    int32_t pad;
};
struct DPCWork {
    const tm_dpc_params* params;
    const DPCInstInfo* info;
    DPCStepState* st;
    double* chunk;        // [nchunks][4]: density sums, vegetation occupied count
    const int2* chunks;   // [nchunks]: (instance, chunk index within it)
    const int2* disk;     // perception offsets (row, column) of every table
};
// LDS of the consumer step kernel: NS density images with the halo, plus the
// perception offsets of the largest radius.
inline size_t dpc_lds_bytes(int ns, int tx, int ty, int halo)
{
    return sizeof(double) * (size_t)ns * (tx + 2 * halo) * (ty + 2 * halo) +
           sizeof(int) * (size_t)(2 * halo + 1) * (2 * halo + 1);
}

// Rietkerk consumer models (tm_movement.hip), FP64 only.  fke_setup builds the
// batch's device tables and scratch behind one allocation (*work, released by
// the caller with hipFree): per-point constants, statistics chunks, perception
// offsets, the 256-cell block table and the RK4 stage and movement buffers.
// Returns TM_OK, or a tm_status with *err naming the problem.
const Variant* fke2_variants_f64(int* n);
const Variant* fke3_variants_f64(int* n);
int fke_setup(int model, const tm_fke_params* params, int n_params, const tm_instance_desc* inst, int n_inst,
              const int64_t* cell_off, int64_t total_cells, void** work, int* nchunks, int* nblocks,
              const char** err);
// Movement statistics of the current fields into out [n_inst][4 SpB] (device).
void launch_fke_mov_stats(int model, const StepArgs& a, double* out, hipStream_t s);

// Default variant index per (model, precision, batch size in cells).
int default_variant(int model, int precision, int64_t total_cells);

void launch_measure_f64(int nf, const DevInst* inst, const DevTile* tiles, int ntiles, int tx, int ty,
                        const void* fields, double2* partials, hipStream_t s);
void launch_measure_f32(int nf, const DevInst* inst, const DevTile* tiles, int ntiles, int tx, int ty,
                        const void* fields, double2* partials, hipStream_t s);
void launch_combine(const CombineArgs& a, int n_inst, int nsteps, hipStream_t s);
void launch_prep(int model, const DevInst* inst, const void* params, const DevState* state, int n_inst,
                 int nsteps, void* table, hipStream_t s);
void launch_bump(DevState* state, int n, hipStream_t s);
size_t const_size(int model);

// Initial frames (tm_frames.hip), one launch for the whole batch.  Values per
// parameter point: speckle frames use perc, a = high, b = low; Gaussian frames
// use a = amp; uniform frames a = lo, b = hi.  mode 0 = speckle, 1 = Gaussian,
// 2 = uniform; blk0[n_inst + 1] = first workgroup of each instance
// (kFrameCellsPerBlock cells per workgroup).
constexpr int kFrameCellsPerBlock = 4096;
struct FrameVals {
    double perc;
    double a[kMaxFields], b[kMaxFields];
};
struct FrameSd {   // Gaussian widths L / v[f], per field (sd_div[kMaxFields] at launch)
    double v[kMaxFields];
};
void launch_init_frames(int prec, int mode, const DevInst* inst, const int64_t* blk0, int n_inst, int64_t nblocks,
                        int nf, int n_gauss, const double* sd_div, const FrameVals* vals, void* fields,
                        hipStream_t s);

// Sum of field 0's occupied counts over the instances in one row (tm_ensemble.hip).
void launch_live_count(const double* rows, int64_t rows_cap, int w, int64_t row, int n_inst, double* out,
                       hipStream_t s);

// Sweep series and replicate statistics (tm_ensemble.hip).
void launch_gather_series(const double* rows, int64_t rows_cap, int w, const int64_t* steps, int m0, int m1,
                          int64_t s0, int n_meas, int n_inst, double* series, hipStream_t s);
// nsp > 0: the multi-species form, [4 nsp] columns per row (species_stats_kernel);
// nsp = 0: field 0's statistics plus the other fields' means, [nf + 3] columns.
void launch_replicate_stats(const double* series, int n_inst, int n_meas, int nf, int nsp, const int32_t* grp_off,
                            const int32_t* grp_inst, int n_groups, int flags, int32_t* first_extinct,
                            double* out, hipStream_t s);

}  // namespace tmk
