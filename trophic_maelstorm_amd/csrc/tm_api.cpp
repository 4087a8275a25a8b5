// tm_api.cpp -- host side of the C ABI declared in include/tm/tm.h.
//
// A context owns: the instance descriptor and tile tables, the double-buffered
// fields of every instance (one allocation, instance blocks [nf][L*L] back to
// back), the tile partials of the per-step rho(t) sums, the device step state,
// the rho(t) rows and one HIP stream.  tm_run enqueues blocks of `graph_steps`
// steps (constants, step launches, rows fold, counter bump) directly on the
// stream: replaying the same block as a hipGraph measured 2-8% slower per step
// on MI355X (ROCm 7.2) for every model, from the 20 us C2 step to the 1 ms C3
// step, so graphs only serve tm_time_steps.  The host tracks the same step
// counter the kernels advance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "tm/tm.h"
#include "tm_io.h"
#include "tm_launch.h"

using namespace tmk;

struct tm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string arch;
    int n_cu = 0;

    // configuration
    int model = 0, prec = 0, nf = 0;
    size_t param_size = 0;
    std::vector<unsigned char> params;
    int n_params = 0;
    std::vector<tm_instance_desc> inst;
    std::vector<int64_t> cell_off;
    int64_t total_cells = 0;
    size_t esize = 0;

    // launch shape
    int variant = -1;
    int graph_steps = 100;
    int xcd = 1;
    std::vector<DevTile> tiles;
    std::vector<DevTile> tiles_a;   // split Rietkerk step: Dornic kernel tiles
    std::vector<DevInst> dinst;

    // device memory
    void* d_params = nullptr;
    DevInst* d_inst = nullptr;
    DevTile* d_tiles = nullptr;
    DevTile* d_tiles_a = nullptr;
    void* d_aux = nullptr;         // split Rietkerk step: P* field [total_cells]
    void* d_fields[2] = {nullptr, nullptr};
    double2* d_partials = nullptr;
    void* d_consts = nullptr;      // [nslots][n_inst] per-step constants
    int nslots = 0;
    DevState* d_state = nullptr;
    double* d_rows = nullptr;
    double* d_rows_pack = nullptr;   // rows of a run packed contiguously for one host copy
    int64_t rows_cap = 0;
    int64_t* d_draws = nullptr;
    bool record_draws = false;
    double* h_pinned = nullptr;   // small pinned staging for state writes

    int parity = 0;
    int64_t counter = 0;

    hipGraphExec_t time_graph[2] = {nullptr, nullptr};   // tm_time_steps blocks (no rows fold), [parity]
    int time_graph_steps = 0;                            // ... of this many steps
    hipEvent_t time_ev[2] = {nullptr, nullptr};

    // Multi-step variants: blocked = launch K steps at a time, else single steps.
    // Adaptive ones choose per block from the occupied fraction of field 0 at
    // the end of the previous block (live_frac; < 0: not known yet).
    bool blocked = true;
    double live_frac = -1.0;
    double* d_live = nullptr;   // [n_inst][2 nf] rows of a measure of the current state, then 2 sums
    hipEvent_t live_ev[2] = {nullptr, nullptr};   // per-block live counts copied to h_pinned[2 + j]
    bool live_pending[2] = {false, false};
    int64_t live_block = 0;
    int64_t steps_single = 0, steps_blocked = 0;   // steps tm_run advanced by each kind of launch

    // device time of the last tm_run's launches (events on the context stream)
    hipEvent_t run_ev[2] = {nullptr, nullptr};
    double last_run_ms = 0.0;
    int64_t last_run_steps = 0;

    // sweep series (tm_run_series) and replicate statistics (tm_replicate_stats)
    double* d_series = nullptr;        // [n_inst][n_series][2 nf]
    int64_t series_cap = 0;            // doubles
    int n_series = 0;                  // measurements of the current series (0: none)
    int64_t* d_series_steps = nullptr;
    int series_steps_cap = 0;
    int32_t* d_groups = nullptr;       // [n_params + 1] offsets, [n_inst] members, [n_inst] first extinct
    int64_t groups_cap = 0;
    double* d_stats = nullptr;         // [n_params][n_series][nf + 3], consumer models [4 species]
    int64_t stats_cap = 0;

    // initial-frame launch table: instance workgroup starts, per-point values
    void* d_frame_tab = nullptr;
    int64_t frame_tab_cap = 0;

    // consumer models: DPCWork, per-instance tables, chunk sums, perception offsets
    void* d_dpc = nullptr;
    int dpc_nchunks = 0;
    int dpc_halo = 1;
    int fke_nblocks = 0;           // Rietkerk consumer models: 256-cell blocks (FkeWork in d_dpc)
    double* d_mov = nullptr;       // tm_mov_stats output [n_inst][4 SpB]
};

namespace {

const char* kDefaultTuneCache = "/.cache/trophic_maelstorm_amd/tune.txt";

#define HIPCHK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return TM_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

tm_status fail(tm_ctx* c, tm_status s, const std::string& msg)
{
    if (c) c->err = msg;
    return s;
}

// The fields were replaced: the live measure of the old ones no longer applies.
void invalidate_live(tm_ctx* c)
{
    c->live_frac = -1.0;
    c->live_pending[0] = c->live_pending[1] = false;
}

const Variant* variant_table(int model, int prec, int* n)
{
    if (model == TM_MODEL_DP) return prec == TM_F64 ? dp_variants_f64(n) : dp_variants_f32(n);
    if (model == TM_MODEL_RIETKERK) return prec == TM_F64 ? rk_variants_f64(n) : rk_variants_f32(n);
    if (model == TM_MODEL_DP2 && prec == TM_F64) return dpc2_variants_f64(n);
    if (model == TM_MODEL_DP3 && prec == TM_F64) return dpc3_variants_f64(n);
    if (model == TM_MODEL_FKE2 && prec == TM_F64) return fke2_variants_f64(n);
    if (model == TM_MODEL_FKE3 && prec == TM_F64) return fke3_variants_f64(n);
    *n = 0;
    return nullptr;
}

bool is_consumer_model(int model) { return model == TM_MODEL_DP2 || model == TM_MODEL_DP3; }
bool is_fke_model(int model) { return model == TM_MODEL_FKE2 || model == TM_MODEL_FKE3; }

// Fields per instance: the densities, then the consumers' perception fields.
int nfields_of(int model)
{
    switch (model) {
    case TM_MODEL_DP: return 1;
    case TM_MODEL_DP2: return 3;
    case TM_MODEL_DP3: return 5;
    case TM_MODEL_FKE2: return 5;   // P, G, W, O, gamma_1
    case TM_MODEL_FKE3: return 7;   // P, G, Pr, W, O, gamma_1, gamma_2
    default: return 3;
    }
}

// Density fields (species) per instance: what a FRAME file holds.
int species_of(int model)
{
    switch (model) {
    case TM_MODEL_DP2: return 2;
    case TM_MODEL_DP3: return 3;
    case TM_MODEL_FKE2: return 4;   // every density the reference's frames and Rho_M hold
    case TM_MODEL_FKE3: return 5;
    default: return nfields_of(model);
    }
}

// computeNeighboringSitesCentral (multiSPDP.cpp:1442-1461): the offsets (row,
// column) of [-R, R]^2 within radius R, generated row-major and sorted by
// std::sort on squared distance alone.  The sort is not stable; the order of
// equal distances is the one that call makes of this input sequence, which
// the reference's truncated perception sums depend on.
std::vector<std::pair<int, int>> perception_offsets(int R)
{
    std::vector<std::pair<int, int>> v;
    for (int x = -R; x <= R; ++x)
        for (int y = -R; y <= R; ++y)
            if (x * x + y * y <= R * R) v.emplace_back(x, y);
    std::sort(v.begin(), v.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return (a.first * a.first + a.second * a.second) < (b.first * b.first + b.second * b.second);
    });
    return v;
}

// r_max_effective = max pR / dx (multiSPDP.cpp:2429) of one parameter point.
double perception_rmax(const tm_dpc_params& p, int ns)
{
    double m = p.pR[0];
    for (int s = 1; s < ns; ++s)
        if (p.pR[s] > m) m = p.pR[s];
    return m / p.dx;
}

void drop_graphs(tm_ctx* c)
{
    for (auto& g : c->time_graph) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
    c->time_graph_steps = 0;
}

template <typename P>
void free_dev(P*& p)
{
    if (p) hipFree((void*)p);
    p = nullptr;
}

void release_config(tm_ctx* c)
{
    drop_graphs(c);
    free_dev(c->d_params);
    free_dev(c->d_inst);
    free_dev(c->d_tiles);
    free_dev(c->d_tiles_a);
    free_dev(c->d_aux);
    c->tiles_a.clear();
    free_dev(c->d_fields[0]);
    free_dev(c->d_fields[1]);
    free_dev(c->d_partials);
    free_dev(c->d_consts);
    free_dev(c->d_rows);
    free_dev(c->d_rows_pack);
    free_dev(c->d_draws);
    free_dev(c->d_series);
    free_dev(c->d_series_steps);
    free_dev(c->d_groups);
    free_dev(c->d_stats);
    free_dev(c->d_frame_tab);
    free_dev(c->d_live);
    free_dev(c->d_dpc);
    free_dev(c->d_mov);
    c->dpc_nchunks = 0;
    c->dpc_halo = 1;
    c->fke_nblocks = 0;
    invalidate_live(c);
    c->series_cap = c->groups_cap = c->stats_cap = c->frame_tab_cap = 0;
    c->series_steps_cap = c->n_series = 0;
    c->rows_cap = 0;
    c->nslots = 0;
    c->model = 0;
}

const Variant& cur_variant(tm_ctx* c)
{
    int n;
    const Variant* v = variant_table(c->model, c->prec, &n);
    return v[c->variant];
}

// (Re)build the tile table for the current variant and upload it with the
// instance descriptors; (re)size the partials buffer.
tm_status build_tiles(tm_ctx* c)
{
    drop_graphs(c);
    const Variant& v = cur_variant(c);
    c->tiles.clear();
    c->dinst.clear();
    for (size_t i = 0; i < c->inst.size(); ++i) {
        const tm_instance_desc& d = c->inst[i];
        DevInst di{};
        di.L = d.L;
        di.pidx = d.params_index;
        di.seed = d.seed;
        di.inst_id = d.instance_id;
        di.step0 = d.step0;
        di.cell_off = c->cell_off[i];
        di.tile0 = (int32_t)c->tiles.size();
        // Vertical strips: consecutive tiles share a halo row, but the
        // workgroups in flight then read the same columns a row stride apart.
        // Along rows: the workgroups in flight read contiguous addresses.
        if (v.row_major) {
            for (int y0 = 0; y0 < d.L; y0 += v.ty)
                for (int x0 = 0; x0 < d.L; x0 += v.tx)
                    c->tiles.push_back(DevTile{(int32_t)i, x0, y0, d.L, c->cell_off[i]});
        } else {
            for (int x0 = 0; x0 < d.L; x0 += v.tx)
                for (int y0 = 0; y0 < d.L; y0 += v.ty)
                    c->tiles.push_back(DevTile{(int32_t)i, x0, y0, d.L, c->cell_off[i]});
        }
        di.ntiles = (int32_t)c->tiles.size() - di.tile0;
        c->dinst.push_back(di);
    }
    free_dev(c->d_tiles);
    free_dev(c->d_partials);
    HIPCHK(c, hipMalloc((void**)&c->d_tiles, sizeof(DevTile) * c->tiles.size()));
    HIPCHK(c, hipMemcpy(c->d_tiles, c->tiles.data(), sizeof(DevTile) * c->tiles.size(), hipMemcpyHostToDevice));
    free_dev(c->d_tiles_a);
    c->tiles_a.clear();
    if (v.fn_a) {   // split step: the Dornic kernel's own tiles, and the P* field
        for (size_t i = 0; i < c->inst.size(); ++i)
            for (int x0 = 0; x0 < c->inst[i].L; x0 += v.ax)
                for (int y0 = 0; y0 < c->inst[i].L; y0 += v.ay)
                    c->tiles_a.push_back(DevTile{(int32_t)i, x0, y0, c->inst[i].L, c->cell_off[i]});
        HIPCHK(c, hipMalloc((void**)&c->d_tiles_a, sizeof(DevTile) * c->tiles_a.size()));
        HIPCHK(c, hipMemcpy(c->d_tiles_a, c->tiles_a.data(), sizeof(DevTile) * c->tiles_a.size(),
                            hipMemcpyHostToDevice));
        if (!c->d_aux && hipMalloc(&c->d_aux, c->esize * (size_t)c->total_cells) != hipSuccess)
            return fail(c, TM_ERR_NOMEM, "P* field allocation failed");
    }
    if (!c->d_inst) HIPCHK(c, hipMalloc((void**)&c->d_inst, sizeof(DevInst) * c->dinst.size()));
    HIPCHK(c, hipMemcpy(c->d_inst, c->dinst.data(), sizeof(DevInst) * c->dinst.size(), hipMemcpyHostToDevice));
    c->nslots = 64;   // a power of two >= the graph block (kernels mask the slot index)
    while (c->nslots < c->graph_steps) c->nslots *= 2;
    HIPCHK(c, hipMalloc((void**)&c->d_partials,
                        sizeof(double2) * (size_t)c->nslots * c->tiles.size() * v.pw * c->nf));
    free_dev(c->d_consts);
    HIPCHK(c, hipMalloc(&c->d_consts, const_size(c->model) * (size_t)c->nslots * c->inst.size()));
    return TM_OK;
}

StepArgs step_args(tm_ctx* c, int parity, int k)
{
    StepArgs a{};
    a.inst = c->d_inst;
    a.tiles = c->d_tiles;
    a.consts = c->d_consts;
    a.in = c->d_fields[parity];
    a.out = c->d_fields[parity ^ 1];
    a.partials = c->d_partials;
    a.state = c->d_state;
    a.draws = c->record_draws ? c->d_draws : nullptr;
    a.tiles_a = c->d_tiles_a;
    a.aux = c->d_aux;
    a.ntiles_a = (int32_t)c->tiles_a.size();
    a.ntiles = (int32_t)c->tiles.size();
    a.nslots = c->nslots;
    a.k = k;
    a.n_inst = (int32_t)c->inst.size();
    a.xcd = c->xcd;
    a.work = c->d_dpc;
    a.nchunks = c->dpc_nchunks;
    a.halo = c->dpc_halo;
    a.nblocks = c->fke_nblocks;
    return a;
}

// Persistent variants launch wg_per_cu workgroups per CU, each walking a
// contiguous tile range; the others launch one workgroup per tile.
int grid_of(tm_ctx* c)
{
    const Variant& v = cur_variant(c);
    const int n = (int)c->tiles.size();
    if (v.wg_per_cu <= 0) return n;
    return std::max(1, std::min(n, c->n_cu * v.wg_per_cu));
}

// One launch advancing steps k .. k + m - 1: m = the variant's steps per launch,
// or 1 (a multi-step variant's single-step kernel: single-step mode, or a
// block's last steps).
void launch_step(tm_ctx* c, int parity, int k, int m = 1)
{
    const StepArgs a = step_args(c, parity, k);
    const Variant& v = cur_variant(c);
    if (v.fn_a) v.fn_a(a, (int)c->tiles_a.size(), c->stream);   // split: P* first
    const StepLaunch fn = (v.spl > 1 && m == 1) ? v.fn_one : v.fn;
    fn(a, grid_of(c), c->stream);
}

// Bookkeeping of tm_mode_steps: n steps of one block in the current mode.
void count_steps(tm_ctx* c, int64_t n)
{
    const int spl = cur_variant(c).spl > 1 && c->blocked ? cur_variant(c).spl : 1;
    const int64_t blocked = spl > 1 ? (n / spl) * spl : 0;
    c->steps_blocked += blocked;
    c->steps_single += n - blocked;
}

// Steps per launch in the current mode.
int steps_per_launch(tm_ctx* c)
{
    const Variant& v = cur_variant(c);
    return (v.spl > 1 && c->blocked) ? v.spl : 1;
}

// Number of step launches (= buffer flips) that advance n steps in one block.
int launches_of(tm_ctx* c, int64_t n)
{
    const int spl = steps_per_launch(c);
    return (int)(n / spl + n % spl);
}

void launch_fold(tm_ctx* c, int nsteps)
{
    CombineArgs a{};
    a.inst = c->d_inst;
    a.partials = c->d_partials;
    a.state = c->d_state;
    a.rows = c->d_rows;
    a.ntiles = (int32_t)c->tiles.size();
    a.nslots = c->nslots;
    a.nf = c->nf;
    a.nsteps = nsteps;
    a.rows_cap = c->rows_cap;
    a.measure = 0;
    a.pw = cur_variant(c).pw;
    launch_combine(a, (int)c->inst.size(), nsteps, c->stream);
}

// One block of n <= nslots steps starting from buffer `parity`: constants,
// n step launches, the rows fold (optional) and the counter bump.
void enqueue_block(tm_ctx* c, int parity, int n, bool fold)
{
    launch_prep(c->model, c->d_inst, c->d_params, c->d_state, (int)c->inst.size(), n, c->d_consts, c->stream);
    const int spl = steps_per_launch(c);
    for (int k = 0; k < n;) {
        const int m = k + spl <= n ? spl : 1;
        launch_step(c, parity, k, m);
        parity ^= 1;
        k += m;
    }
    if (fold) launch_fold(c, n);
    launch_bump(c->d_state, n, c->stream);
}

// Capture one block of n steps (with or without the rows fold) for a parity:
// the graphs tm_time_steps replays.
tm_status capture_block(tm_ctx* c, int parity, int n, bool fold, hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_block(c, parity, n, fold);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return fail(c, TM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(c, TM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return TM_OK;
}

tm_status ensure_rows(tm_ctx* c, int64_t n_steps)
{
    if (c->rows_cap >= n_steps) return TM_OK;
    drop_graphs(c);
    free_dev(c->d_rows);
    free_dev(c->d_rows_pack);
    c->rows_cap = 0;
    const int64_t cap = std::max<int64_t>(n_steps, 256);
    const size_t bytes = sizeof(double) * (size_t)cap * c->inst.size() * 2 * c->nf;
    HIPCHK(c, hipMalloc((void**)&c->d_rows, bytes));
    HIPCHK(c, hipMalloc((void**)&c->d_rows_pack, bytes));
    c->rows_cap = cap;
    return TM_OK;
}

tm_status check_ready(tm_ctx* c)
{
    if (!c) return TM_ERR_ARG;
    if (!c->model) return fail(c, TM_ERR_STATE, "tm_configure has not been called");
    return TM_OK;
}

tm_status set_run_base(tm_ctx* c)
{
    // run_base <- counter, ordered on the stream before the step launches.
    *(int64_t*)c->h_pinned = c->counter;
    HIPCHK(c, hipMemcpyAsync((char*)c->d_state + offsetof(DevState, run_base), c->h_pinned, sizeof(int64_t),
                             hipMemcpyHostToDevice, c->stream));
    return TM_OK;
}

// rho(t) row of the current state of every instance into out[i * stride * 2 nf]
// (device): tile partials of the current fields, then the rows fold.
void enqueue_measure(tm_ctx* c, double* out, int64_t stride)
{
    const Variant& v = cur_variant(c);
    if (c->prec == TM_F64)
        launch_measure_f64(c->nf, c->d_inst, c->d_tiles, (int)c->tiles.size(), v.tx, v.ty, c->d_fields[c->parity],
                           c->d_partials, c->stream);
    else
        launch_measure_f32(c->nf, c->d_inst, c->d_tiles, (int)c->tiles.size(), v.tx, v.ty, c->d_fields[c->parity],
                           c->d_partials, c->stream);
    CombineArgs a{};
    a.inst = c->d_inst;
    a.partials = c->d_partials;
    a.state = c->d_state;
    a.rows = out;
    a.ntiles = (int32_t)c->tiles.size();
    a.nslots = c->nslots;
    a.nf = c->nf;
    a.nsteps = 1;
    a.rows_cap = stride;
    a.measure = 1;
    a.pw = 1;
    launch_combine(a, (int)c->inst.size(), 1, c->stream);
}

tm_status ensure_live(tm_ctx* c)
{
    if (!c->d_live)
        HIPCHK(c, hipMalloc((void**)&c->d_live, sizeof(double) * (2 * c->nf * c->inst.size() + 2)));
    for (auto& e : c->live_ev)
        if (!e) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return TM_OK;
}

// The occupied fraction of field 0 of the current state (one measure and a
// wait), for an adaptive multi-step variant whose mode is not known yet.
tm_status live_fraction_now(tm_ctx* c, double* frac)
{
    tm_status s = ensure_live(c);
    if (s != TM_OK) return s;
    const int n = (int)c->inst.size();
    double* sums = c->d_live + 2 * c->nf * n;
    enqueue_measure(c, c->d_live, 1);
    launch_live_count(c->d_live, 1, 2 * c->nf, 0, n, sums, c->stream);
    double occ = 0.0;
    HIPCHK(c, hipMemcpyAsync(&occ, sums, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *frac = occ / (double)c->total_cells;
    return TM_OK;
}

// An adaptive variant takes single steps while more than its threshold
// (Variant::adaptive, per mille) of the field-0 cells is occupied: there the
// K-step kernel's recomputed rings cost more sampling than its saved field
// traffic is worth (measured on the DP quench, where the 4-step launches win
// once about 2% of the cells or fewer are occupied).
tm_status update_mode(tm_ctx* c)
{
    const Variant& v = cur_variant(c);
    if (v.spl == 1) return TM_OK;
    if (!v.adaptive) {
        c->blocked = true;
        return TM_OK;
    }
    if (c->live_frac < 0.0) {
        tm_status s = live_fraction_now(c, &c->live_frac);
        if (s != TM_OK) return s;
    }
    c->blocked = c->live_frac <= v.adaptive * 1e-3;
    return TM_OK;
}

// Advance n_steps with the rows fold; rows stay on the device.  Adaptive
// multi-step variants re-choose their mode between blocks: after each block a
// one-workgroup kernel sums field 0's occupied counts of the block's last row
// and the sum is copied to pinned host memory behind an event; the choice
// uses the newest sum whose event has completed (one or two blocks old), so
// the host never waits for the device.
tm_status advance(tm_ctx* c, int64_t n_steps)
{
    tm_status s = ensure_rows(c, n_steps);
    if (s != TM_OK) return s;
    if ((s = update_mode(c)) != TM_OK) return s;
    if ((s = set_run_base(c)) != TM_OK) return s;
    const bool adaptive = cur_variant(c).spl > 1 && cur_variant(c).adaptive;
    if (adaptive && (s = ensure_live(c)) != TM_OK) return s;
    int64_t left = n_steps;
    const int B = c->graph_steps;
    volatile double* h_sum = (volatile double*)c->h_pinned + 2;   // [2], after the run-base staging word
    auto after_block = [&]() -> tm_status {
        if (!adaptive) return TM_OK;
        // The sum of two blocks back, waited for: the host stays at most two
        // blocks ahead of the device, which keeps the choice at most two blocks
        // old while the device always has the newest block queued.  The sum of
        // the block before, if already there, is newer.
        const int j = (int)(c->live_block & 1);          // slot of two blocks back, reused now
        if (c->live_pending[j]) {
            HIPCHK(c, hipEventSynchronize(c->live_ev[j]));
            c->live_frac = h_sum[j] / (double)c->total_cells;
            c->live_pending[j] = false;
        }
        if (c->live_pending[j ^ 1] && hipEventQuery(c->live_ev[j ^ 1]) == hipSuccess) {
            c->live_frac = h_sum[j ^ 1] / (double)c->total_cells;
            c->live_pending[j ^ 1] = false;
        }
        double* sums = c->d_live + 2 * c->nf * c->inst.size();
        launch_live_count(c->d_rows, c->rows_cap, 2 * c->nf, n_steps - left - 1, (int)c->inst.size(), sums + j,
                          c->stream);
        HIPCHK(c, hipMemcpyAsync((void*)(h_sum + j), sums + j, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(c->live_ev[j], c->stream));
        c->live_pending[j] = true;
        c->live_block += 1;
        return update_mode(c);
    };
    while (B > 0 && left >= B) {
        enqueue_block(c, c->parity, B, true);
        HIPCHK(c, hipGetLastError());
        count_steps(c, B);
        c->counter += B;
        c->parity ^= (launches_of(c, B) & 1);
        left -= B;
        if ((s = after_block()) != TM_OK) return s;
    }
    while (left > 0) {
        const int chunk = (int)std::min<int64_t>(left, c->nslots);
        enqueue_block(c, c->parity, chunk, true);
        HIPCHK(c, hipGetLastError());
        count_steps(c, chunk);
        c->parity ^= (launches_of(c, chunk) & 1);
        c->counter += chunk;
        left -= chunk;
        if ((s = after_block()) != TM_OK) return s;
    }
    return TM_OK;
}

// First fault recorded by the kernels: (instance, cell, instance step), or all -1.
tm_status read_fault(tm_ctx* c, int32_t* inst, int64_t* cell, int64_t* step)
{
    DevState st;
    HIPCHK(c, hipMemcpy(&st, c->d_state, sizeof(st), hipMemcpyDeviceToHost));
    *inst = -1;
    *cell = -1;
    *step = -1;
    if (st.fault_counter != ~0ull) {
        *inst = (int32_t)((st.fault_key >> 40) & 0xFFFFF);   // bits 60..63: the step's place in its launch
        *cell = (int64_t)(st.fault_key & ((1ull << 40) - 1));
        *step = (*inst >= 0 && *inst < (int)c->inst.size() ? c->inst[*inst].step0 : 0) + (int64_t)st.fault_counter;
    }
    return TM_OK;
}

tm_status fault_status(tm_ctx* c)
{
    int32_t inst;
    int64_t cell, step;
    tm_status s = read_fault(c, &inst, &cell, &step);
    if (s != TM_OK) return s;
    if (inst >= 0) {
        char buf[160];
        snprintf(buf, sizeof buf, "numeric fault: instance %d cell %lld step %lld", inst, (long long)cell,
                 (long long)step);
        c->err = buf;
        return TM_FAULT_NUMERIC;
    }
    return TM_OK;
}

// Zero state with no fault recorded.
// Consumer models: perception tables per parameter point, statistics chunks
// and per-instance tables, in one device allocation behind a DPCWork header.
tm_status setup_dpc(tm_ctx* c)
{
    const int ns = species_of(c->model);
    const tm_dpc_params* P = (const tm_dpc_params*)c->params.data();
    std::vector<int2> disk;
    std::vector<int> p_off((size_t)c->n_params), p_len((size_t)c->n_params), p_R((size_t)c->n_params);
    int halo = 1;
    for (int g = 0; g < c->n_params; ++g) {
        const double rmax = perception_rmax(P[g], ns);
        const int R = (int)rmax;
        const std::vector<std::pair<int, int>> o = perception_offsets(R);
        p_off[(size_t)g] = (int)disk.size();
        p_len[(size_t)g] = (int)o.size();
        p_R[(size_t)g] = R;
        for (const auto& q : o) disk.push_back(make_int2(q.first, q.second));
        halo = std::max(halo, R);
    }
    std::vector<DPCInstInfo> info(c->inst.size());
    std::vector<int2> chunks;
    for (size_t i = 0; i < c->inst.size(); ++i) {
        const tm_instance_desc& d = c->inst[i];
        const tm_dpc_params& p = P[d.params_index];
        const double rmax = perception_rmax(p, ns);
        DPCInstInfo& in = info[i];
        in = DPCInstInfo{};
        in.R = p_R[(size_t)d.params_index];
        in.disk_off = p_off[(size_t)d.params_index];
        in.disk_len = p_len[(size_t)d.params_index];
        for (int s = 0; s < ns; ++s) in.fr[s] = p.pR[s] / (p.dx * rmax);   // multiSPDP.cpp:2434-2435
        const int64_t cells = (int64_t)d.L * d.L;
        const int nch = (int)((cells + kDpcChunkCells - 1) / kDpcChunkCells);
        in.chunk0 = (int)chunks.size();
        in.nchunks = nch;
        for (int k = 0; k < nch; ++k) chunks.push_back(make_int2((int)i, k));
    }
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_info = up(sizeof(DPCWork));
    const size_t o_st = o_info + up(sizeof(DPCInstInfo) * info.size());
    const size_t o_chunk = o_st + up(sizeof(DPCStepState) * info.size());
    const size_t o_chunks = o_chunk + up(sizeof(double) * 4 * chunks.size());
    const size_t o_disk = o_chunks + up(sizeof(int2) * chunks.size());
    const size_t total = o_disk + up(sizeof(int2) * disk.size());
    free_dev(c->d_dpc);
    if (hipMalloc(&c->d_dpc, total) != hipSuccess) return fail(c, TM_ERR_NOMEM, "consumer tables allocation failed");
    char* base = (char*)c->d_dpc;
    DPCWork w{};
    w.params = (const tm_dpc_params*)c->d_params;
    w.info = (const DPCInstInfo*)(base + o_info);
    w.st = (DPCStepState*)(base + o_st);
    w.chunk = (double*)(base + o_chunk);
    w.chunks = (const int2*)(base + o_chunks);
    w.disk = (const int2*)(base + o_disk);
    std::vector<char> host(total, 0);
    std::memcpy(host.data(), &w, sizeof(w));
    std::memcpy(host.data() + o_info, info.data(), sizeof(DPCInstInfo) * info.size());
    std::memcpy(host.data() + o_chunks, chunks.data(), sizeof(int2) * chunks.size());
    std::memcpy(host.data() + o_disk, disk.data(), sizeof(int2) * disk.size());
    HIPCHK(c, hipMemcpy(c->d_dpc, host.data(), total, hipMemcpyHostToDevice));
    c->dpc_nchunks = (int)chunks.size();
    c->dpc_halo = halo;
    return TM_OK;
}

tm_status reset_state(tm_ctx* c)
{
    DevState st{};
    st.fault_counter = ~0ull;
    st.fault_key = ~0ull;
    HIPCHK(c, hipMemcpy(c->d_state, &st, sizeof(st), hipMemcpyHostToDevice));
    return TM_OK;
}

std::string tune_key(tm_ctx* c)
{
    std::map<int, int> sizes;
    for (auto& d : c->inst) sizes[d.L] += 1;
    std::ostringstream k;
    k << c->arch << "|m" << c->model << "|p" << c->prec << "|";
    for (auto& kv : sizes) k << kv.first << "x" << kv.second << ",";
    return k.str();
}

std::string tune_path(const char* p)
{
    if (p && *p) return p;
    if (const char* e = getenv("TM_TUNE_CACHE")) return e;
    const char* h = getenv("HOME");
    return std::string(h ? h : "/tmp") + kDefaultTuneCache;
}

void mkdirs(const std::string& path)
{
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

}  // namespace

extern "C" {

tm_status tm_create(tm_ctx** out, int device)
{
    if (!out) return TM_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return TM_ERR_NODEV;
    tm_ctx* c = new tm_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return TM_ERR_NODEV;
    }
    c->arch = prop.gcnArchName;
    c->n_cu = prop.multiProcessorCount;
    if (c->arch.rfind("gfx950", 0) != 0) {
        delete c;
        return TM_ERR_NODEV;   // code objects are built for gfx950 only
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->d_state, sizeof(DevState)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_pinned, 64, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&c->run_ev[0]) != hipSuccess || hipEventCreate(&c->run_ev[1]) != hipSuccess) {
        tm_destroy(c);
        return TM_ERR_HIP;
    }
    *out = c;
    return TM_OK;
}

void tm_destroy(tm_ctx* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    release_config(c);
    free_dev(c->d_state);
    if (c->h_pinned) hipHostFree(c->h_pinned);
    for (auto e : c->live_ev)
        if (e) hipEventDestroy(e);
    for (auto e : c->run_ev)
        if (e) hipEventDestroy(e);
    for (auto e : c->time_ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* tm_last_error(const tm_ctx* c) { return c ? c->err.c_str() : "null context"; }

tm_status tm_configure(tm_ctx* c, int model, int precision, const void* params, int n_params,
                       const tm_instance_desc* instances, int n_inst)
{
    if (!c) return TM_ERR_ARG;
    if (model != TM_MODEL_DP && model != TM_MODEL_RIETKERK && !is_consumer_model(model) && !is_fke_model(model))
        return fail(c, TM_ERR_ARG, "unknown model");
    if (precision != TM_F64 && precision != TM_F32) return fail(c, TM_ERR_ARG, "unknown precision");
    if ((is_consumer_model(model) || is_fke_model(model)) && precision != TM_F64)
        return fail(c, TM_ERR_ARG, "the consumer models run in TM_F64 only");
    if (!params || n_params <= 0 || !instances || n_inst <= 0) return fail(c, TM_ERR_ARG, "empty batch");
    if (is_consumer_model(model)) {
        const tm_dpc_params* P = (const tm_dpc_params*)params;
        const int ns = species_of(model);
        for (int g = 0; g < n_params; ++g) {
            if (!(P[g].dx > 0.0) || !(P[g].dt > 0.0)) return fail(c, TM_ERR_ARG, "dt and dx must be positive");
            for (int s = 0; s < ns; ++s)
                if (!(P[g].pR[s] >= 0.0)) return fail(c, TM_ERR_ARG, "perception radii must be >= 0");
            const double rmax = perception_rmax(P[g], ns);
            if (!(rmax > 0.0)) return fail(c, TM_ERR_ARG, "the largest perception radius must be positive");
            if (!(rmax < TM_DPC_MAX_RADIUS + 1.0))
                return fail(c, TM_ERR_ARG, "perception radius above TM_DPC_MAX_RADIUS sites");
        }
        for (int i = 0; i < n_inst; ++i) {
            const tm_instance_desc& d = instances[i];
            if (d.params_index < 0 || d.params_index >= n_params) continue;   // reported below
            if ((int)perception_rmax(P[d.params_index], ns) >= d.L)
                return fail(c, TM_ERR_ARG, "perception radius must be smaller than L");
        }
    }
    for (int i = 0; i < n_inst; ++i) {
        const tm_instance_desc& d = instances[i];
        if (d.L < 1 || d.L > 65536) return fail(c, TM_ERR_ARG, "instance L out of range [1, 65536]");
        if ((int64_t)d.L * d.L > 0xFFFFFFFFll) return fail(c, TM_ERR_ARG, "instance too large for 32-bit cell ids");
        if (d.params_index < 0 || d.params_index >= n_params) return fail(c, TM_ERR_ARG, "params_index out of range");
        if (d.step0 < 0) return fail(c, TM_ERR_ARG, "negative step0");
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    release_config(c);
    c->model = model;
    c->prec = precision;
    c->nf = nfields_of(model);
    c->esize = precision == TM_F64 ? sizeof(double) : sizeof(float);
    c->param_size = model == TM_MODEL_DP         ? sizeof(tm_dp_params)
                    : is_consumer_model(model) ? sizeof(tm_dpc_params)
                    : is_fke_model(model)      ? sizeof(tm_fke_params)
                                               : sizeof(tm_rietkerk_params);
    c->params.assign((const unsigned char*)params, (const unsigned char*)params + c->param_size * n_params);
    c->n_params = n_params;
    c->inst.assign(instances, instances + n_inst);
    c->cell_off.resize(n_inst);
    c->total_cells = 0;
    for (int i = 0; i < n_inst; ++i) {
        c->cell_off[i] = c->total_cells;
        c->total_cells += (int64_t)instances[i].L * instances[i].L;
    }
    c->variant = default_variant(model, precision, c->total_cells);
    {   // whole multi-step launches per graph block, as tm_set_launch makes them
        const int spl = cur_variant(c).spl;
        if (spl > 1 && c->graph_steps > 0) c->graph_steps = (c->graph_steps + spl - 1) / spl * spl;
    }
    HIPCHK(c, hipMalloc(&c->d_params, c->param_size * n_params));
    HIPCHK(c, hipMemcpy(c->d_params, c->params.data(), c->param_size * n_params, hipMemcpyHostToDevice));
    const size_t fbytes = c->esize * (size_t)c->nf * c->total_cells;
    for (int b = 0; b < 2; ++b) {
        if (hipMalloc(&c->d_fields[b], fbytes) != hipSuccess) {
            release_config(c);
            return fail(c, TM_ERR_NOMEM, "field allocation failed");
        }
        HIPCHK(c, hipMemsetAsync(c->d_fields[b], 0, fbytes, c->stream));
    }
    HIPCHK(c, hipMalloc((void**)&c->d_draws, sizeof(int64_t) * (size_t)c->total_cells));
    HIPCHK(c, hipMemsetAsync(c->d_draws, 0xFF, sizeof(int64_t) * (size_t)c->total_cells, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (reset_state(c) != TM_OK) return TM_ERR_HIP;
    c->parity = 0;
    c->counter = 0;
    tm_status s = build_tiles(c);
    if (s != TM_OK) return s;
    if (is_consumer_model(model) && (s = setup_dpc(c)) != TM_OK) return s;
    if (is_fke_model(model)) {
        const char* err = "";
        const int rc = fke_setup(model, (const tm_fke_params*)c->params.data(), n_params, c->inst.data(), n_inst,
                                 c->cell_off.data(), c->total_cells, &c->d_dpc, &c->dpc_nchunks, &c->fke_nblocks,
                                 &err);
        if (rc != TM_OK) {
            const std::string msg = err;
            release_config(c);
            return fail(c, (tm_status)rc, msg);
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TM_OK;
}

tm_status tm_upload_fields(tm_ctx* c, int i, const double* host)
{
    if (c) invalidate_live(c);   // new fields
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !host) return fail(c, TM_ERR_ARG, "bad instance or pointer");
    const int64_t n = (int64_t)c->nf * c->inst[i].L * c->inst[i].L;
    char* dst = (char*)c->d_fields[c->parity] + c->esize * c->nf * c->cell_off[i];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->prec == TM_F64) {
        HIPCHK(c, hipMemcpy(dst, host, sizeof(double) * n, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(host, host + n);
        HIPCHK(c, hipMemcpy(dst, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    return TM_OK;
}

tm_status tm_download_fields(tm_ctx* c, int i, double* host)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !host) return fail(c, TM_ERR_ARG, "bad instance or pointer");
    const int64_t n = (int64_t)c->nf * c->inst[i].L * c->inst[i].L;
    const char* src = (const char*)c->d_fields[c->parity] + c->esize * c->nf * c->cell_off[i];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->prec == TM_F64) {
        HIPCHK(c, hipMemcpy(host, src, sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(c, hipMemcpy(tmp.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < n; ++k) host[k] = tmp[k];
    }
    return TM_OK;
}

// One launch of the frame kernel over every instance: the workgroup start of
// each instance and the values of each parameter point go to a small device table.
static tm_status enqueue_frames(tm_ctx* c, int mode, int n_gauss, const double* sd_div,
                                const std::vector<FrameVals>& vals)
{
    invalidate_live(c);   // new fields
    const int n = (int)c->inst.size();
    std::vector<int64_t> blk0((size_t)n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const int64_t cells = (int64_t)c->inst[(size_t)k].L * c->inst[(size_t)k].L;
        blk0[(size_t)k + 1] = blk0[(size_t)k] + (cells + kFrameCellsPerBlock - 1) / kFrameCellsPerBlock;
    }
    const size_t bytes = sizeof(int64_t) * blk0.size() + sizeof(FrameVals) * vals.size();
    if ((int64_t)bytes > c->frame_tab_cap) {
        free_dev(c->d_frame_tab);
        c->frame_tab_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_frame_tab, bytes));
        c->frame_tab_cap = (int64_t)bytes;
    }
    std::vector<char> host(bytes);
    std::memcpy(host.data(), blk0.data(), sizeof(int64_t) * blk0.size());
    std::memcpy(host.data() + sizeof(int64_t) * blk0.size(), vals.data(), sizeof(FrameVals) * vals.size());
    HIPCHK(c, hipMemcpyAsync(c->d_frame_tab, host.data(), bytes, hipMemcpyHostToDevice, c->stream));
    const int64_t* d_blk0 = (const int64_t*)c->d_frame_tab;
    const FrameVals* d_vals = (const FrameVals*)((const char*)c->d_frame_tab + sizeof(int64_t) * blk0.size());
    launch_init_frames(c->prec, mode, c->d_inst, d_blk0, n, blk0[(size_t)n], c->nf, n_gauss, sd_div, d_vals,
                       c->d_fields[c->parity], c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));   // the staging vector lives until the copy is done
    return TM_OK;
}

tm_status tm_init_uniform(tm_ctx* c, const double* lo, const double* hi)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!lo || !hi) return fail(c, TM_ERR_ARG, "null bounds");
    std::vector<FrameVals> vals((size_t)c->n_params);
    for (auto& v : vals) {
        v = FrameVals{};
        for (int f = 0; f < c->nf; ++f) {
            v.a[f] = lo[f];
            v.b[f] = hi[f];
        }
    }
    double sd_none[kMaxFields];
    std::fill(sd_none, sd_none + kMaxFields, 1.0);
    return enqueue_frames(c, 2, 0, sd_none, vals);
}

// ---- output layer: frames, snapshots ----------------------------------------

tm_status tm_write_frame_csv(tm_ctx* c, int i, const char* path, double a, const char* header, int digits)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !path || digits < 1 || digits > 17)
        return fail(c, TM_ERR_ARG, "tm_write_frame_csv: bad instance, path or digits");
    const int L = c->inst[(size_t)i].L;
    std::vector<double> host((size_t)c->nf * L * L);
    if ((s = tm_download_fields(c, i, host.data())) != TM_OK) return s;
    // a FRAME file holds the densities (multiSPDP.cpp:1596-1645 writes Rho_dt)
    if (tm_format_frame_csv(path, host.data(), species_of(c->model), L, a, header, digits) != 0)
        return fail(c, TM_ERR_ARG, std::string("tm_write_frame_csv: cannot write ") + path);
    return TM_OK;
}

tm_status tm_save_snapshot(tm_ctx* c, int i, const char* path)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !path) return fail(c, TM_ERR_ARG, "tm_save_snapshot: bad instance or path");
    const tm_instance_desc& d = c->inst[(size_t)i];
    const size_t bytes = c->esize * (size_t)c->nf * d.L * d.L;
    std::vector<char> host(bytes);
    const char* src = (const char*)c->d_fields[c->parity] + c->esize * c->nf * c->cell_off[(size_t)i];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(host.data(), src, bytes, hipMemcpyDeviceToHost));
    const tmio::SnapInfo info{c->model, c->prec, c->nf, d.L, d.seed, d.instance_id, d.step0 + c->counter,
                              d.params_index};
    if (tmio::write_snapshot(path, info, host.data(), bytes) != 0)
        return fail(c, TM_ERR_ARG, std::string("tm_save_snapshot: cannot write ") + path);
    return TM_OK;
}

tm_status tm_load_snapshot(tm_ctx* c, int i, const char* path, int64_t* step)
{
    if (c) invalidate_live(c);   // new fields
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !path) return fail(c, TM_ERR_ARG, "tm_load_snapshot: bad instance or path");
    tmio::SnapInfo info;
    std::vector<char> data;
    if (tmio::read_snapshot(path, &info, &data) != 0)
        return fail(c, TM_ERR_ARG, std::string("tm_load_snapshot: not a readable snapshot: ") + path);
    const tm_instance_desc& d = c->inst[(size_t)i];
    if (info.model != c->model || info.nf != c->nf || info.L != d.L)
        return fail(c, TM_ERR_ARG, "tm_load_snapshot: model or lattice size differs from the instance");
    char* dst = (char*)c->d_fields[c->parity] + c->esize * c->nf * c->cell_off[(size_t)i];
    const size_t n = (size_t)c->nf * d.L * d.L;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (info.precision == c->prec) {
        HIPCHK(c, hipMemcpy(dst, data.data(), data.size(), hipMemcpyHostToDevice));
    } else if (c->prec == TM_F64) {   // FP32 snapshot into FP64 fields: exact
        std::vector<double> v(n);
        for (size_t k = 0; k < n; ++k) v[k] = ((const float*)data.data())[k];
        HIPCHK(c, hipMemcpy(dst, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
    } else {                          // FP64 snapshot into FP32 fields: rounded as tm_upload_fields does
        std::vector<float> v(n);
        for (size_t k = 0; k < n; ++k) v[k] = (float)((const double*)data.data())[k];
        HIPCHK(c, hipMemcpy(dst, v.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (step) *step = info.step;
    return TM_OK;
}

tm_status tm_init_speckle(tm_ctx* c, const double* perc, const double* high, const double* low)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!perc || !high || !low) return fail(c, TM_ERR_ARG, "tm_init_speckle: null values");
    std::vector<FrameVals> vals((size_t)c->n_params);
    for (int g = 0; g < c->n_params; ++g) {
        FrameVals& v = vals[(size_t)g];
        v = FrameVals{};
        v.perc = perc[g];
        for (int f = 0; f < c->nf; ++f) {
            v.a[f] = high[(size_t)g * c->nf + f];
            v.b[f] = low[(size_t)g * c->nf + f];
        }
    }
    double sd_none[kMaxFields];
    std::fill(sd_none, sd_none + kMaxFields, 1.0);
    return enqueue_frames(c, 0, 0, sd_none, vals);
}

tm_status tm_init_gauss(tm_ctx* c, int n_gauss, const double* sd_div, int n_sd, const double* amp)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!amp || n_gauss < 0 || n_gauss > c->nf) return fail(c, TM_ERR_ARG, "tm_init_gauss: bad n_gauss or amp");
    if (n_gauss > 0 && (!sd_div || n_sd < 1)) return fail(c, TM_ERR_ARG, "tm_init_gauss: need sd_div[n_sd >= 1]");
    // Missing widths repeat the last one given (rietkerk_bjork_FKE.cpp:1118-1127).
    double sd[kMaxFields];
    std::fill(sd, sd + kMaxFields, 1.0);
    for (int f = 0; f < n_gauss; ++f) {
        sd[f] = sd_div[std::min(f, n_sd - 1)];
        if (!(sd[f] > 0.0)) return fail(c, TM_ERR_ARG, "tm_init_gauss: sd_div must be > 0");
    }
    std::vector<FrameVals> vals((size_t)c->n_params);
    for (int g = 0; g < c->n_params; ++g) {
        vals[(size_t)g] = FrameVals{};
        for (int f = 0; f < c->nf; ++f) vals[(size_t)g].a[f] = amp[(size_t)g * c->nf + f];
    }
    return enqueue_frames(c, 1, n_gauss, sd, vals);
}

int tm_mft_frame_values(int homogeneous, double R, double R_c, double dP, int nf, const double* c_spread,
                        const double* mft, double* high, double* low)
{
    if (nf < 1 || nf > 3 || !c_spread || !mft || !high || !low) return -1;
    const int o = R < R_c ? 0 : nf;   // below R_c the first nf entries, else the last nf
    for (int f = 0; f < nf; ++f) {
        const double e = mft[o + f], cs = c_spread[o + f];
        high[f] = low[f] = e == 0 ? cs : cs * e;   // a zero MFT value leaves the bare scale factor
    }
    if (!homogeneous) {
        // field 0: the MFT value, raised to max(dP, c_spread[.] E_0) in the speckles; at or
        // above R_c the background is the bare E_0 (rietkerk_bjork_FKE.cpp:596-630)
        const double e0 = mft[o], cs0 = c_spread[o];
        high[0] = std::max(dP, cs0 * e0);
        low[0] = o == 0 ? cs0 * e0 : e0;
    }
    return 0;
}

tm_status tm_run(tm_ctx* c, int64_t n_steps, double* rows)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (n_steps < 0) return fail(c, TM_ERR_ARG, "negative n_steps");
    if (n_steps == 0) return TM_OK;
    HIPCHK(c, hipEventRecord(c->run_ev[0], c->stream));
    if ((s = advance(c, n_steps)) != TM_OK) return s;
    HIPCHK(c, hipEventRecord(c->run_ev[1], c->stream));
    if (rows) {
        // The run's rows are strided on the device (rows_cap steps per instance):
        // packed there first, then one contiguous copy.  A strided copy into
        // pageable host memory goes one instance row at a time (about 7 us each).
        const size_t w = sizeof(double) * 2 * c->nf;
        const double* src = c->d_rows;
        if (n_steps < c->rows_cap && c->inst.size() > 1) {
            HIPCHK(c, hipMemcpy2DAsync(c->d_rows_pack, w * n_steps, c->d_rows, w * c->rows_cap, w * n_steps,
                                       c->inst.size(), hipMemcpyDeviceToDevice, c->stream));
            src = c->d_rows_pack;
        }
        HIPCHK(c, hipMemcpyAsync(rows, src, w * n_steps * c->inst.size(), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->run_ev[0], c->run_ev[1]));
    c->last_run_ms = ms;
    c->last_run_steps = n_steps;
    return fault_status(c);
}

tm_status tm_mode_steps(const tm_ctx* c, int64_t* single_steps, int64_t* blocked_steps)
{
    if (!c || !single_steps || !blocked_steps) return TM_ERR_ARG;
    *single_steps = c->steps_single;
    *blocked_steps = c->steps_blocked;
    return TM_OK;
}

tm_status tm_last_run_time(const tm_ctx* c, double* ms, int64_t* n_steps)
{
    if (!c || !ms) return TM_ERR_ARG;
    *ms = c->last_run_ms;
    if (n_steps) *n_steps = c->last_run_steps;
    return TM_OK;
}

tm_status tm_measure(tm_ctx* c, double* rows)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!rows) return fail(c, TM_ERR_ARG, "null rows");
    double* d_out = nullptr;
    const size_t bytes = sizeof(double) * 2 * c->nf * c->inst.size();
    HIPCHK(c, hipMalloc((void**)&d_out, bytes));
    enqueue_measure(c, d_out, 1);
    hipError_t e = hipMemcpyAsync(rows, d_out, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d_out);
    if (e != hipSuccess) return fail(c, TM_ERR_HIP, std::string("tm_measure: ") + hipGetErrorString(e));
    return TM_OK;
}

tm_status tm_mov_stats(tm_ctx* c, double* out)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!out) return fail(c, TM_ERR_ARG, "null out");
    if (!is_fke_model(c->model)) return fail(c, TM_ERR_STATE, "tm_mov_stats: not a Rietkerk consumer model");
    const int sp = c->model == TM_MODEL_FKE3 ? 3 : 2;
    const size_t bytes = sizeof(double) * 4 * sp * c->inst.size();
    if (!c->d_mov) HIPCHK(c, hipMalloc((void**)&c->d_mov, bytes));
    launch_fke_mov_stats(c->model, step_args(c, c->parity, 0), c->d_mov, c->stream);
    HIPCHK(c, hipGetLastError());
    hipError_t e = hipMemcpyAsync(out, c->d_mov, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, TM_ERR_HIP, std::string("tm_mov_stats: ") + hipGetErrorString(e));
    return TM_OK;
}

tm_status tm_set_record_draws(tm_ctx* c, int enable)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (c->record_draws != (enable != 0)) drop_graphs(c);
    c->record_draws = enable != 0;
    return TM_OK;
}

tm_status tm_download_draws(tm_ctx* c, int i, int64_t* host)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !host) return fail(c, TM_ERR_ARG, "bad instance or pointer");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(host, c->d_draws + c->cell_off[i], sizeof(int64_t) * c->inst[i].L * c->inst[i].L,
                        hipMemcpyDeviceToHost));
    return TM_OK;
}

tm_status tm_fault_info(tm_ctx* c, int32_t* inst, int64_t* cell, int64_t* step)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int32_t i;
    int64_t ce, st;
    if ((s = read_fault(c, &i, &ce, &st)) != TM_OK) return s;
    if (inst) *inst = i;
    if (cell) *cell = ce;
    if (step) *step = st;
    return TM_OK;
}

tm_status tm_get_step(tm_ctx* c, int i, int64_t* step)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (i < 0 || i >= (int)c->inst.size() || !step) return fail(c, TM_ERR_ARG, "bad instance or pointer");
    *step = c->inst[i].step0 + c->counter;
    return TM_OK;
}

int tm_variant_count(int model, int precision)
{
    int n = 0;
    variant_table(model, precision, &n);
    return n;
}

const char* tm_variant_name(int model, int precision, int variant)
{
    int n = 0;
    const Variant* v = variant_table(model, precision, &n);
    if (!v || variant < 0 || variant >= n) return nullptr;
    return v[variant].name;
}

tm_status tm_set_launch(tm_ctx* c, int variant, int graph_steps)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (graph_steps < 0 || graph_steps > 4096) return fail(c, TM_ERR_ARG, "graph_steps out of range [0, 4096]");
    const int n = tm_variant_count(c->model, c->prec);
    if (variant < -1 || variant >= n) return fail(c, TM_ERR_ARG, "variant out of range");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->variant = variant < 0 ? default_variant(c->model, c->prec, c->total_cells) : variant;
    // A graph block of a multi-step variant is a whole number of its launches:
    // a remainder would run as single steps in every block.
    const int spl = cur_variant(c).spl;
    c->graph_steps = spl > 1 ? (graph_steps + spl - 1) / spl * spl : graph_steps;
    return build_tiles(c);
}

tm_status tm_get_launch(const tm_ctx* c, int* variant, int* graph_steps)
{
    if (!c) return TM_ERR_ARG;
    if (!c->model) return TM_ERR_STATE;
    if (variant) *variant = c->variant;
    if (graph_steps) *graph_steps = c->graph_steps;
    return TM_OK;
}

void* tm_stream(tm_ctx* c) { return c ? (void*)c->stream : nullptr; }

tm_status tm_time_steps(tm_ctx* c, int64_t n_steps, double* ms_total, int64_t* launches)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (n_steps <= 0 || !ms_total) return fail(c, TM_ERR_ARG, "n_steps must be positive");
    const Variant& v = cur_variant(c);
    if (v.spl > 1 && v.adaptive)
        return fail(c, TM_ERR_STATE,
                    "tm_time_steps runs no rows fold, so an adaptive variant could not choose its launch mode "
                    "between blocks: time it with tm_run and tm_last_run_time");
    if ((s = update_mode(c)) != TM_OK) return s;
    // Step kernels only (no rows fold): graphs of up to 100 steps, one per
    // starting parity, captured once per context and launch shape and kept with
    // the run graphs, replayed between two events on the context stream.
    const int B = (int)std::min<int64_t>(std::min<int64_t>(n_steps, 100), c->nslots);
    const int flip = launches_of(c, B) & 1;
    if (c->time_graph_steps != B) {
        for (auto& g : c->time_graph) {
            if (g) hipGraphExecDestroy(g);
            g = nullptr;
        }
        c->time_graph_steps = 0;
        for (int par = 0; par < 2; ++par)   // a block from either parity
            if ((s = capture_block(c, par, B, false, &c->time_graph[par])) != TM_OK) return s;
        c->time_graph_steps = B;
    }
    for (auto& e : c->time_ev)
        if (!e) HIPCHK(c, hipEventCreate(&e));
    HIPCHK(c, hipEventRecord(c->time_ev[0], c->stream));
    int64_t left = n_steps;
    while (left >= B) {
        HIPCHK(c, hipGraphLaunch(c->time_graph[c->parity], c->stream));
        c->parity ^= flip;
        left -= B;
    }
    int64_t nlaunch = (n_steps / B) * launches_of(c, B);
    if (left > 0) {
        enqueue_block(c, c->parity, (int)left, false);
        c->parity ^= (launches_of(c, left) & 1);
        nlaunch += launches_of(c, left);
    }
    HIPCHK(c, hipEventRecord(c->time_ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->time_ev[1]));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->time_ev[0], c->time_ev[1]));
    count_steps(c, n_steps);
    c->counter += n_steps;
    *ms_total = ms;
    if (launches) *launches = nlaunch;
    return fault_status(c);
}

tm_status tm_reset(tm_ctx* c)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((s = reset_state(c)) != TM_OK) return s;
    c->counter = 0;
    c->parity = 0;
    return TM_OK;
}

tm_status tm_autotune(tm_ctx* c, const char* cache_path, int* out_variant, int* out_graph_steps)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    const std::string key = tune_key(c);
    const std::string path = tune_path(cache_path);
    {   // cached choice
        std::ifstream f(path);
        std::string k;
        int v, g;
        while (f >> k >> v >> g) {
            if (k == key && v >= 0 && v < tm_variant_count(c->model, c->prec)) {
                if ((s = tm_set_launch(c, v, g)) != TM_OK) return s;
                if (out_variant) *out_variant = v;
                if (out_graph_steps) *out_graph_steps = c->graph_steps;
                return TM_OK;
            }
        }
    }
    // Dry timing: launches read the current buffer and write the other one
    // without advancing the step counter, so the batch state is unchanged.
    const int nv = tm_variant_count(c->model, c->prec);
    const int graph_steps = c->graph_steps;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    // The variant in place when tuning starts (the default unless the caller
    // chose one) is kept unless another is at least 5% faster on this state.
    const int start_v = c->variant;
    int best_v = start_v;
    double best = 1e30, start_time = 1e30;
    int reps = 6;
    for (int v = 0; v < nv; ++v) {
        c->variant = v;
        if ((s = build_tiles(c)) != TM_OK) return s;
        if ((s = update_mode(c)) != TM_OK) return s;   // the mode this state would run in
        const int spl = steps_per_launch(c);           // steps per timed launch
        launch_prep(c->model, c->d_inst, c->d_params, c->d_state, (int)c->inst.size(), spl, c->d_consts,
                    c->stream);
        launch_step(c, c->parity, 0, spl);   // warm-up
        // Best of three trials of `reps` launches; reps grows until a trial
        // takes about a millisecond, so that close variants are told apart.
        double vbest = 1e30;
        for (int trial = 0; trial < 3; ++trial) {
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r) launch_step(c, c->parity, 0, spl);
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            vbest = std::min(vbest, (double)ms / (reps * spl));   // per step
            if (v == 0 && trial == 0 && ms < 1.0f) reps = std::min(200, (int)(reps * 1.0 / std::max(ms, 1e-3f)) + 1);
        }
        if (v == start_v) start_time = vbest;
        if (vbest < best) {
            best = vbest;
            best_v = v;
        }
    }
    if (best > 0.95 * start_time) {
        best_v = start_v;
        best = start_time;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // Block length: short steps want long blocks (fewer host round trips).
    const double per_step_ms = best;
    int g = per_step_ms < 0.05 ? 200 : (graph_steps > 0 ? graph_steps : 100);
    c->variant = best_v;
    if ((s = tm_set_launch(c, best_v, g)) != TM_OK) return s;
    g = c->graph_steps;
    mkdirs(path);
    std::ofstream f(path, std::ios::app);
    if (f) f << key << " " << best_v << " " << g << "\n";
    if (out_variant) *out_variant = best_v;
    if (out_graph_steps) *out_graph_steps = g;
    return TM_OK;
}

// ---- Sweep measurement and replicate statistics ----------------------------

int tm_perception_offsets(int R, int32_t* out, int cap)
{
    if (R < 0 || (R > 0 && !out)) return -1;
    const std::vector<std::pair<int, int>> v = perception_offsets(R);
    if ((int)v.size() > cap) return -1;
    for (size_t k = 0; k < v.size(); ++k) {
        out[2 * k] = v[k].first;
        out[2 * k + 1] = v[k].second;
    }
    return (int)v.size();
}

int tm_log10_time_bins(double t_max, double dt, double* out, int cap)
{
    // 10^p for p = 0, 0.04, 0.08, ... (p accumulated) up to floor(10 log10 t_max)/10,
    // each rounded down to the dt grid (an int number of dt) and kept above 1,
    // skipping a repeat of the previous point; then, from the end, every point
    // less than dt/2 above its predecessor is dropped
    // (rietkerk_bjork_FKE.cpp:1497-1536).
    std::vector<double> t;
    if (!(t_max > 1.0) || !(dt > 0.0)) return 0;
    const double top = std::floor(std::log10(t_max) * 10.0) / 10.0;
    for (double p = 0.0; p <= top; p += 0.04) {
        double v = (double)(int)(std::pow(10.0, p) / dt) * dt;
        if (v <= 1.0) v = v + dt;
        if (t.empty() || v != t.back()) t.push_back(v);
    }
    for (int i = (int)t.size() - 1; i > 0; --i)
        if (t[i] < t[i - 1] + dt / 2.0) t.erase(t.begin() + i);
    if (out)
        for (int i = 0; i < (int)t.size() && i < cap; ++i) out[i] = t[i];
    return (int)t.size();
}

int tm_measure_steps(const double* t_meas, int n_meas, double dt, int64_t* steps)
{
    if (!t_meas || !steps || n_meas <= 0 || !(dt > 0.0)) return 0;
    const double t_end = t_meas[n_meas - 1] + dt;
    double t = 0.0;
    int idx = 0;
    for (int64_t n = 0; t < t_end && idx < n_meas; ++n) {
        if (t == 0 || (t >= t_meas[idx] - dt / 2.0 && t < t_meas[idx] + dt / 2.0)) steps[idx++] = n;
        t += dt;
    }
    return idx;
}

tm_status tm_run_series(tm_ctx* c, const int64_t* steps, int n_meas, double* series)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!steps || n_meas <= 0) return fail(c, TM_ERR_ARG, "tm_run_series: need steps[n_meas] with n_meas > 0");
    for (int m = 0; m < n_meas; ++m)
        if (steps[m] < 0 || (m > 0 && steps[m] < steps[m - 1]))
            return fail(c, TM_ERR_ARG, "tm_run_series: steps must be non-decreasing and >= 0");
    const int n_inst = (int)c->inst.size();
    const int w = 2 * c->nf;
    const int64_t need = (int64_t)n_inst * n_meas * w;
    c->n_series = 0;
    if (need > c->series_cap) {
        free_dev(c->d_series);
        c->series_cap = 0;
        if (hipMalloc((void**)&c->d_series, sizeof(double) * (size_t)need) != hipSuccess)
            return fail(c, TM_ERR_NOMEM, "series allocation failed");
        c->series_cap = need;
    }
    if (n_meas > c->series_steps_cap) {
        free_dev(c->d_series_steps);
        c->series_steps_cap = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_series_steps, sizeof(int64_t) * (size_t)n_meas));
        c->series_steps_cap = n_meas;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_series_steps, steps, sizeof(int64_t) * (size_t)n_meas, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipEventRecord(c->run_ev[0], c->stream));
    int m = 0;
    for (; m < n_meas && steps[m] == 0; ++m) enqueue_measure(c, c->d_series + (int64_t)m * w, n_meas);
    // Advance in chunks of whole graph blocks whose rows fit a ~32 MiB rows
    // buffer (or the one already there); after each chunk, keep the rows of the
    // measurements it reached.
    int64_t chunk = std::max<int64_t>(c->rows_cap, (int64_t{32} << 20) / ((int64_t)sizeof(double) * n_inst * w));
    if (c->graph_steps > 0 && chunk >= c->graph_steps) chunk -= chunk % c->graph_steps;
    chunk = std::max<int64_t>(chunk, 1);
    const int64_t total = steps[n_meas - 1];
    for (int64_t s0 = 0; s0 < total;) {
        const int64_t len = std::min(chunk, total - s0);
        if ((s = advance(c, len)) != TM_OK) return s;
        int m1 = m;
        while (m1 < n_meas && steps[m1] <= s0 + len) ++m1;
        launch_gather_series(c->d_rows, c->rows_cap, w, c->d_series_steps, m, m1, s0, n_meas, n_inst, c->d_series,
                             c->stream);
        m = m1;
        s0 += len;
        // the next advance() rewrites the pinned staging its first copy reads
        if (s0 < total) HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->run_ev[1], c->stream));
    if (series)
        HIPCHK(c, hipMemcpyAsync(series, c->d_series, sizeof(double) * (size_t)need, hipMemcpyDeviceToHost,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->run_ev[0], c->run_ev[1]));
    c->last_run_ms = ms;
    c->last_run_steps = total;
    c->n_series = n_meas;
    return fault_status(c);
}

tm_status tm_replicate_stats(tm_ctx* c, int flags, double* stats)
{
    tm_status s = check_ready(c);
    if (s != TM_OK) return s;
    if (!stats) return fail(c, TM_ERR_ARG, "null stats");
    if (flags & ~TM_STATS_REFERENCE) return fail(c, TM_ERR_ARG, "tm_replicate_stats: unknown flags");
    if (c->n_series <= 0) return fail(c, TM_ERR_STATE, "tm_run_series has not been called");
    const int n_inst = (int)c->inst.size(), ng = c->n_params, n_meas = c->n_series;
    const int nsp = is_consumer_model(c->model) || is_fke_model(c->model) ? species_of(c->model)
                                                                           : 0;   // multi-species rows
    const int ncol = nsp > 0 ? 4 * nsp : c->nf + 3;
    // Instances grouped by parameter point, in instance (= replicate) order; then
    // the first extinct measurement per (instance, species).
    std::vector<int32_t> grp((size_t)ng + 1 + (size_t)n_inst * (1 + std::max(nsp, 1)), 0);
    for (const tm_instance_desc& d : c->inst) ++grp[(size_t)d.params_index + 1];
    for (int g = 0; g < ng; ++g) grp[(size_t)g + 1] += grp[(size_t)g];
    std::vector<int32_t> cur(grp.begin(), grp.begin() + ng);
    for (int i = 0; i < n_inst; ++i) grp[(size_t)ng + 1 + cur[(size_t)c->inst[(size_t)i].params_index]++] = i;
    const int64_t gsize = (int64_t)grp.size();
    if (gsize > c->groups_cap) {
        free_dev(c->d_groups);
        c->groups_cap = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_groups, sizeof(int32_t) * (size_t)gsize));
        c->groups_cap = gsize;
    }
    const int64_t nout = (int64_t)ng * n_meas * ncol;
    if (nout > c->stats_cap) {
        free_dev(c->d_stats);
        c->stats_cap = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_stats, sizeof(double) * (size_t)nout));
        c->stats_cap = nout;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_groups, grp.data(), sizeof(int32_t) * (size_t)(ng + 1 + n_inst),
                             hipMemcpyHostToDevice, c->stream));
    launch_replicate_stats(c->d_series, n_inst, n_meas, c->nf, nsp, c->d_groups, c->d_groups + ng + 1, ng, flags,
                           c->d_groups + ng + 1 + n_inst, c->d_stats, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(stats, c->d_stats, sizeof(double) * (size_t)nout, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TM_OK;
}

}  // extern "C"

namespace tmk {
int default_variant(int model, int precision, int64_t total_cells)
{
    const char* want;
    if (model == TM_MODEL_FKE2 || model == TM_MODEL_FKE3) {
        return 0;                          // one launch shape
    } else if (model == TM_MODEL_DP2) {
        want = "dpc2_f64_32x16_t256_w5";   // measured fastest on the consumer sweep batch
    } else if (model == TM_MODEL_DP3) {
        want = "dpc3_f64_32x16_t256_w4";
    } else if (model == TM_MODEL_DP) {
        // adaptive eight-step launches: single steps while more than 2% of the
        // cells are occupied (the two-pass step that defers rejected Gamma
        // trials), eight steps per launch once the field is mostly absorbing
        want = precision == TM_F64 ? "dp_f64_64x32_t256_x8aq_w6" : "dp_f32_128x24_t256_x8aq";
    } else {
        // Rietkerk: small lattices are latency-bound and want the single fused
        // 16 x 16 launch; batches want the split step (each P* sampled once),
        // in FP32 with the two-pass Dornic kernel on 128 x 24 tiles (C3 batch:
        // 4.6e10 against 4.3e10 cell-updates/s for the one-pass 128 x 32 kernel)
        if (total_cells < (1 << 20)) return 0;
        want = precision == TM_F64 ? "rk_f64_split_32x16_t256" : "rk_f32_split_32x16_t256_q24";
    }
    const int n = tm_variant_count(model, precision);
    for (int v = 0; v < n; ++v)
        if (std::string(tm_variant_name(model, precision, v)) == want) return v;
    return 0;
}
}  // namespace tmk
