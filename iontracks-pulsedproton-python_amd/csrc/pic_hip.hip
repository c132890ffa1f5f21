// C ABI of the gfx950 time-stepping backend (declared in include/pic_hip.h).
//
// One handle = one HIP stream + the four carrier arrays (current and next of
// both species) + scratch. The time-step loop runs here, on the host side of
// the library: for each batch of up to `batch` steps one gather launch
// computes the deposition density of every in-pulse step of the batch, then
// one sweep launch per step does deposition-add, Lax-Wendroff update, lateral
// ring and the per-step reduction into a device (n_steps, 4) record. Nothing
// is copied back per step. With graph replay on, a segment's launches are
// collected as data and replayed through graphs of kChain kernel nodes (see
// submit()).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pic_hip.h"
#include "pic_kernels.hpp"
#include "pic_sweep2.hpp"
#include "pic_adjoint.hpp"

// Staging of pic_run_pulse_static, the entry point that allocates nothing:
// the per-step record, the step offsets and the track centres of one call.
// Every GPU the library runs on holds its own copy for the life of the
// process. In host code `scratch` names a host-side stand-in, not device
// memory: kernels and copies get the per-device address from
// static_scratch() below, never the array's name.
constexpr size_t kStaticScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

thread_local char g_err[1024] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(PIC_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define HIP_LAUNCH_CHECK()                                                                                  \
    do {                                                                                                    \
        hipError_t e_ = hipGetLastError();                                                                  \
        if (e_ != hipSuccess) return fail(PIC_E_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct SweepShape {
    int tj = 0, chunk = 0, n_windows = 0, wpb = 0, n_kg = 0, n_jt = 0, n_ic = 0, v = 1, win_step = 0, tpb = 1;
    int64_t blocks() const { return (int64_t)n_kg * ((n_jt + tpb - 1) / tpb) * n_ic; }
    int threads() const { return 64 * wpb * tpb; }
};

// One kernel launch held as data: the kernel, its shape and a copy of its
// single by-value argument. `kind` classifies it for the timing events
// (0 sweep, 1 deposition gather, 2 other).
constexpr size_t kMaxArg = 256;
struct KLaunch {
    const void *func = nullptr;
    dim3 grid, block;
    int kind = 0;
    alignas(16) unsigned char arg[kMaxArg];

    template <typename T>
    static KLaunch make(const void *f, dim3 g, dim3 b, int kind, const T &a) {
        static_assert(sizeof(T) <= kMaxArg, "kernel argument larger than a launch record");
        KLaunch l;
        l.func = f;
        l.grid = g;
        l.block = b;
        l.kind = kind;
        memcpy(l.arg, &a, sizeof(T));
        return l;
    }
};

}  // namespace

struct pic_sim {
    pic_params p{};
    int prec = PIC_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t gstream = nullptr;  // gathers of the next segment, beside this one's sweeps
    hipEvent_t seg_go = nullptr;    // a segment's first sweep is done: the next gather may start
    hipEvent_t seg_dens = nullptr;  // a segment's gather is done: its sweeps may start
    size_t elem = 8;
    int64_t pitch = 0;  // row pitch in elements (nzb; FP32 rounded up to even, or to a multiple of four)
    int64_t n_vox = 0;  // allocated elements per array, no_xy^2 * pitch
    void *cur_p = nullptr, *cur_n = nullptr, *nxt_p = nullptr, *nxt_n = nullptr;
    double *dens = nullptr;        // [dens_slots][no_xy^2]
    double *sweep_part = nullptr;  // [2 slots][sweep_part_cap][3]
    int64_t sweep_part_cap = 0;
    double *dep_part = nullptr;    // [dens_slots][gather tiles]
    int64_t batch = 1;             // steps per gather launch
    int64_t *offsets_d = nullptr;
    // a step whose partials are written but not yet folded into its record
    struct Pending {
        const double *part = nullptr;
        int n = 0;
        const double *dep = nullptr;
        int n_dep = 0;
        double *out4 = nullptr;
    } pending;
    double *scratch4 = nullptr;  // 8 doubles for the kernel-level calls
    double *zero_dens = nullptr; // [no_xy^2] zeros: the density a step without tracks adds
    double *per_step = nullptr;
    int64_t n_steps = 0;
    double *xs = nullptr, *ys = nullptr;
    int64_t n_tracks = 0;
    std::vector<int64_t> offsets;
    int force_tj = 0, force_chunk = 0, force_v = 0, force_tpb = 0;
    int steps_per_pass = 0;        // 1, 2 (pairs of steps through sweep2_kernel), 0: the library's choice
    int pass2_chunk = 0;           // planes per block of sweep2_kernel, 0: the library's choice
    double *pair_part = nullptr;   // [2 launches][2 steps][pair_part_cap][3]: the two-step pass's partials
    int64_t pair_part_cap = 0;
    int pair_slot = 0;             // which half of pair_part the next pass writes
    pic::FoldArgs pair_pending[2]{};  // the last pass's two steps, not yet folded into their records
    bool pair_open = false;
    int cus = 256;                 // compute units of the device (queried once)
    bool shape_valid = false;      // `shape` matches the tiling knobs
    SweepShape shape;
    bool timing = false;
    int timing_group = 1;  // sweep launches per event pair
    int group_open = 0;    // sweeps launched since the open pair's first event
    std::vector<hipEvent_t> ev_pool;  // reused across calls
    std::vector<int> ev_kind;         // per recorded event: 0 sweep, 1 deposit, 2 ring + reduce
    std::vector<int> ev_count;        // per pair (at its first event): launches it brackets
    size_t ev_used = 0;
    bool in_run = false;  // events are recorded inside pic_run_steps only
    bool graphs = false;  // replay a run's launches through graph chains (never while timing)
    std::vector<KLaunch> *sink = nullptr;  // open segment: launches are collected here
    std::vector<KLaunch> seg;              // a run's segment buffer, reused
    // Tracks binned by gather tile for the uploaded run (pic_set_gather_mode):
    // counters, scans and cursors per (step of a segment, tile), entries per
    // step, and the entry list, sized for the busiest segment of the upload.
    int gather_mode = PIC_GATHER_AUTO;
    bool bins_ready = false;
    int *bin_cnt = nullptr, *bin_off = nullptr, *bin_cur = nullptr, *bin_tot = nullptr, *bin_list = nullptr;
    int64_t bin_cap = 0;
};

// Many configurations advanced together: per global step one sweep launch
// covers the blocks of every configuration still running (their block ranges
// concatenated), each block reading its configuration's arguments from a
// device table. The handles keep their own arrays, tracks and records.
struct pic_batch {
    std::vector<pic_sim *> sims;  // caller's order
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int prec = PIC_FP64, v = 1, wpb = 1;
    bool reflect = false;
    int force_tj = 0, force_chunk = 0;
    int64_t waves_min = 0;
    std::vector<std::pair<int, int>> cands;       // candidate (rows per tile, planes per block), preferred first
    std::vector<std::vector<SweepShape>> shapes;  // [candidate][sims[c]]
    // Per-segment argument tables (a segment is one gather batch of steps),
    // double-buffered so the host builds segment k+1 while the device runs
    // segment k: pinned host staging, device copy, and the event that marks
    // the copy out of a host slot as done.
    int64_t slot_steps = 0, slot_cfgs = 0;
    pic::SweepArgs *descs_h[2] = {nullptr, nullptr}, *descs_d[2] = {nullptr, nullptr};
    int *offs_h[2] = {nullptr, nullptr}, *offs_d[2] = {nullptr, nullptr};
    // the segment's gathers: one GatherArgs per configuration with tracks, and block offsets
    pic::GatherArgs *gdesc_h[2] = {nullptr, nullptr}, *gdesc_d[2] = {nullptr, nullptr};
    int *goffs_h[2] = {nullptr, nullptr}, *goffs_d[2] = {nullptr, nullptr};
    hipEvent_t slot_ev[2] = {nullptr, nullptr};
    bool slot_busy[2] = {false, false};
    double *zero_dens = nullptr;  // the density of a step without tracks, no_xy^2 of the largest grid
    std::vector<pic::SweepArgs> descs_build;  // host-side build of one segment's tables
    std::vector<int> offs_build, goffs_build;
    std::vector<pic::GatherArgs> gdesc_build;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;  // pairs around each launch: kind 0 sweep, 1 gather
    std::vector<int> ev_kind;
    size_t ev_used = 0;
    bool graphs = false;       // replay each segment's launches through graph chains (never while timing)
    std::vector<KLaunch> seg;  // one segment's launches
    // Lanes: the configurations that run much longer than the others (over
    // 0.8 of the longest run, at most a quarter of them) are advanced by a
    // sub-batch of their own, on its own stream and host thread, beside
    // sub-batches of the rest (lane_split); their latency-bound tail of small
    // launches then overlaps the rest's sweeps instead of following them. The
    // sub-batches are built on the first run that splits and kept while the
    // split holds.
    static constexpr int kMaxLanes = 3;
    pic_batch *lanes[kMaxLanes] = {nullptr, nullptr, nullptr};
    int n_lanes = 0;
    std::vector<int> lane_of;  // the split the lanes were built for: lane of sims[c]
    bool is_lane = false;      // a sub-batch: never splits again
};

// Reverse-mode derivative of one handle's run: checkpoints of the state
// every K steps during the forward pass; per segment (last first) the states
// are recomputed from the checkpoint into a store of K+1 array pairs, then
// the adjoint steps run backwards through them.
struct pic_adjoint {
    pic_sim *s = nullptr;
    int64_t K = 0, nseg = 0, T = 0;
    int64_t M = 1;            // segments per outer checkpoint (1: every segment's state is kept)
    int64_t n_outer = 0;      // ceil(nseg / M)
    int chunk = 2;            // i planes per adjoint block
    double *ckpt = nullptr;   // [n_outer][2][n_vox]: the state at the start of every M-th segment
    double *inner = nullptr;  // [M][2][n_vox] (M > 1): one outer span's segment states, recomputed
    double *store = nullptr;  // [K + 1][2][n_vox]
    double *mu = nullptr;     // [2 buffers][2 species][n_vox]
    double *dgf = nullptr, *dh = nullptr;  // [K][no_xy^2]
    double *dpart = nullptr;               // [K][gather tiles][2]
    double *acc = nullptr;                 // [adjoint blocks][2 * kAdjNV] running compensated sums
    double *grad = nullptr;                // [kAdjNV] folded totals
    double *dinj_rec = nullptr;            // [T][2]
    double *scratch = nullptr;             // sweep records of the recompute (discarded)
    double *zeros = nullptr;               // [no_xy^2] the density (and derivatives) of a step without tracks
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float forward_ms = 0, backward_ms = 0;
    size_t bytes = 0;
};

namespace {

int sm_count(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || v <= 0) v = 256;
    return v;
}

// Windows (waves) per tile in a block of `tpb` tiles. A block holds every
// window of its rows when they fit in 8 waves (the sweep kernels' 512-thread
// bound): a row is then read and written whole by one CU. Longer rows are
// split into equal groups where the window count allows: 4 windows when it
// divides by 4 (one- and two-value layouts only), else 2 when it is even, 3
// when it divides by 3, else 2 (the last group short). A short group is what
// costs: its blocks hold a full block's place on a CU with some waves idle.
// Measured on MI355X, sweep times on one device. Two FP64 values, 2-row
// lockstep pairs: 208^2 x 506 (5 windows), groups of 2 179 us (1: 181,
// 3+2: 215, 4+1: 243; all 5 in one-tile blocks: 190); 408^2 x 631 (6), 2
// 0.777 ms (3: 0.807, 1: 0.888, all 6 in one-tile blocks: 0.808); 248^2 x 806
// (7), 2 388 us (1: 386, 3: 399, 4: 404, all 7 in one-tile blocks: 399);
// 506^2 x 1010 (9), 3 1.82 ms (2: 1.89, 1 and 4: 1.94, 8+1 in one-tile
// blocks: 3.04). Two FP32 values, 4-row tiles: 506^2 x 1010 (8 windows), two
// groups of 4 in lockstep pairs 0.97 ms (2: 0.97; all 8 in one-tile blocks:
// 1.04); 2656^2 x 1010 (8), 25.6 ms (2: 26.1; all 8: 27.1); 648^2 x 806 (7),
// all 7 in one-tile blocks 1.31 ms (pairs of tiles, groups of 1: 1.36, 2:
// 1.39, 4: 1.49); 408^2 x 631 (5), all 5 0.448 ms (groups of 2 and 1 in pairs
// of tiles: 0.441 and 0.442; 4: 0.641).
int window_group(int n_windows, int tpb, bool wide) {
    const int cap = std::max(1, 8 / tpb);
    if (n_windows <= cap) return n_windows;
    if (!wide && cap >= 4 && n_windows % 4 == 0) return 4;
    for (int d = 2; d <= std::min(cap, 3); ++d)
        if (n_windows % d == 0) return d;
    return std::min(cap, 2);
}

// Waves of a sweep tiling: blocks of `tpb` tiles of `tj` rows times `wpb`
// windows, over every (window group, tile group, chunk of `chunk` planes).
int64_t sweep_waves(int64_t lateral, int64_t n_windows, int wpb, int tj, int tpb, int chunk) {
    return ceil_div(n_windows, wpb) * ceil_div(ceil_div(lateral, tj), tpb) * ceil_div(lateral, chunk) * wpb * tpb;
}

// k values per lane: FP64 fields one or two (8- or 16-byte loads); FP32
// fields one, two or four (4-, 8- or 16-byte loads). FP32 grids take four when
// four-value windows take no more lanes than two-value ones and the four-value
// tiling (3-row tiles, two per block) gives at least 6 waves per CU at
// 16-plane chunks; two otherwise. Measured on MI355X, each pair on one device:
// 206^2 x 410 and 208^2 x 406, four values (3-row pairs) 67.8 and 69.7 us
// against 72.9 and 74.3 for the two-value tiling (4-row tiles of 8 planes);
// 166^2 and 246^2 x 210 (one window), two values 27.1 and 50.8 us against
// 28.6 and 52.0; 2656^2 x 1010 (nine two-value windows, five four-value
// ones), 26.9 against 30.2 ms; the 16^2..47^2 x 106..206 grids of a
// saturation curve run one after another, 84 against 108 ms. Grids that took
// four values with 2-row tiles are 1-9 % faster with 3-row ones
// (compute_sweep_shape).
bool four_values_suit(const pic_params *p, int cus) {
    const int64_t n = p->nzb - 2, w4 = pic::window_outputs(4), w2 = pic::window_step(2, n);
    const int64_t lateral = p->no_xy - 2;
    const bool lanes_ok = ceil_div(n, w4) * w4 <= ceil_div(n, w2) * w2;
    const int nw = (int)ceil_div(n, w4);
    const bool waves_ok = sweep_waves(lateral, nw, window_group(nw, 2, true), 3, 2, 16) >= (int64_t)cus * 6;
    return lanes_ok && waves_ok;
}

// FP64 grids take two values per lane (2-row tiles in lockstep pairs) when
// that tiling gives at least 4 waves per CU at 4-plane chunks, whatever the
// lanes the wider windows leave idle; one value per lane otherwise. Measured
// on MI355X, sweep time against the one-value tiling, same device: 1066^2 x
// 410, 3.07 against 3.67 ms; 208^2 x 406, 125 against 155 us; 536^2 x 210,
// 411 against 487 us; 506^2 x 1010, 1.82 against 2.11 ms; 246^2, 166^2 and
// 106^2 x 210, 92, 37 and 20 against 109, 39 and 22 us; 128^2, 88^2 and 48^2
// x 406, 38, 23 and 13 against 44, 35 and 15 us; 46^2 x 210 (484 waves), 9.5
// against 8.5 us.
bool fp64_pairs_suit(const pic_params *p, int cus) {
    const int64_t lateral = p->no_xy - 2;
    const int nw = (int)ceil_div(p->nzb - 2, pic::window_outputs(2));
    return sweep_waves(lateral, nw, window_group(nw, 2, true), 2, 2, 4) >= (int64_t)cus * 4;
}

int values_per_lane(const pic_sim *s, int force_v) {
    if (s->prec == PIC_FP64) {
        if (force_v > 0) return force_v;
        if (s->force_tj > 2) return 1;  // two FP64 values per lane take 1- and 2-row tiles
        return fp64_pairs_suit(&s->p, s->cus) ? 2 : 1;
    }
    if (force_v > 0) return force_v;
    if (s->force_tj == 3) return 4;  // tiles of three rows exist for four values
    if (s->force_tj > 3) return 2;   // tiles of four rows exist for one and two values (eight for one)
    return four_values_suit(&s->p, s->cus) ? 4 : 2;
}

// Planes per block of the four-value tiling, whose blocks hold two waves per
// SIMD (8 per CU): the longest chunk of 40, 32, 16, 8 or 4 planes that still
// gives four rounds of resident waves, so no CU marches much more than
// another; failing that, the shortest chunk whose waves fit in one round (the
// most waves that run at once), searched over every length from 4 to 32
// planes: the busiest CU's blocks each march chunk + 2 planes, so the
// shortest chunk that still fits one round is the fastest. Measured on MI355X
// (3-row tiles, two per block, each grid against its other chunks on one
// device): 1066^2 x 410, 40 planes 1.561 ms (32: 1.569, 36: 1.562, 44:
// 1.586, 48: 1.578, 64: 1.579, 24: 1.579); 726^2 x 410, 40 planes 731 us
// (32: 745, 36: 734, 44: 776, 48: 747); 486^2 x 410, 16 planes 332 us (8:
// 337, 32: 358); 326^2 x 410, 8 planes 156 us (16: 166, 32: 188); 208^2 x
// 406, 15 planes (490 blocks for 512 places) 68.3 us (16, 455 blocks: 69.8;
// 14, 525 blocks, a second round: 88.9; 12: 82.7); 206^2 x 410, 14 planes
// 66.3 us (15: 66.1, 16: 66.8, 12: 74.4); 536^2 x 210, 8 planes 227 us (16:
// 232, 32: 259); 366^2 x 210, 4 planes 116 us (8: 117, 16: 128).
int four_value_chunk(int64_t lateral, int64_t n_windows, int wpb, int tj, int tpb, int cus) {
    const int64_t round = (int64_t)cus * 8;
    for (int c : {40, 32, 16, 8, 4})
        if (sweep_waves(lateral, n_windows, wpb, tj, tpb, c) >= 4 * round) return c;
    for (int c = 4; c < 32; ++c)
        if (sweep_waves(lateral, n_windows, wpb, tj, tpb, c) <= round) return c;
    return 32;
}

// Planes per block of two-FP64-value tiles: the chunk whose busiest CU
// marches the fewest planes, counted as rounds of resident blocks (8 waves
// per CU) times chunk + 2 planes (a march loads two planes before its first
// output); a shorter chunk is taken only when that count is 2 % lower. The
// count ranks the measured chunks of every grid tried within 1-3 %: on
// MI355X, 208^2 x 406, 24 planes 124 us (16: 125, 12: 129, 8: 135, 32: 147);
// 128^2 x 406, 16 planes 38 us (8: 41, 4: 49, 32: 56); 88^2 x 406, 8 planes
// 23 us (4: 28, 16: 32); 246^2 x 210, 32 planes 92 us (4-16: 92-95); 1066^2
// x 410, 32 planes 3.07 ms (16: 3.11; 48 and 64, outside the list: 3.03).
int pair_chunk(const SweepShape &sh, int64_t planes, int cus) {
    const int64_t slots = (int64_t)cus * std::max(1, 8 / (sh.wpb * sh.tpb));
    const int64_t groups = (int64_t)sh.n_kg * ceil_div(sh.n_jt, sh.tpb);
    int best = 32;
    double best_cost = 0;
    for (int c : {32, 24, 16, 12, 8, 4}) {
        const double cost = (double)ceil_div(groups * ceil_div(planes, c), slots) * (c + 2);
        if (c == 32 || cost < 0.98 * best_cost) {
            best = c;
            best_cost = cost;
        }
    }
    return best;
}

// Fixed tiling of one configuration: v values per lane, tj rows per tile,
// `chunk` planes per block, wpb windows (waves) per block.
SweepShape fixed_shape(const pic_sim *s, int v, int tj, int chunk, int wpb) {
    SweepShape sh;
    sh.v = v;
    sh.win_step = pic::window_step(v, s->p.nzb - 2);
    sh.n_windows = (int)ceil_div(s->p.nzb - 2, sh.win_step);
    sh.wpb = wpb;
    sh.n_kg = (int)ceil_div(sh.n_windows, wpb);
    const int64_t rows = s->p.no_xy - 2, planes = s->p.no_xy - 2;
    sh.tj = tj;
    sh.n_jt = (int)ceil_div(rows, tj);
    sh.chunk = (int)std::min<int64_t>(chunk, planes);
    sh.n_ic = (int)ceil_div(planes, sh.chunk);
    return sh;
}

// Rows per tile (registers) and planes per block: enough blocks to keep every
// CU busy, as few halo rows and chunk prologues as that allows.
SweepShape compute_sweep_shape(const pic_sim *s);

// the handle's tiling, recomputed only when a knob changes
SweepShape sweep_shape(pic_sim *s) {
    if (!s->shape_valid) {
        s->shape = compute_sweep_shape(s);
        s->shape_valid = true;
    }
    return s->shape;
}

SweepShape compute_sweep_shape(const pic_sim *s) {
    SweepShape sh;
    const int64_t no_xy = s->p.no_xy, nzb = s->p.nzb;
    sh.v = values_per_lane(s, s->force_v);
    // FP64 windows of two values step by 124 (the 126 step of FP32 pairs
    // needs single-value edge stores)
    sh.win_step = s->prec == PIC_FP64 ? pic::window_outputs(sh.v) : pic::window_step(sh.v, nzb - 2);
    sh.n_windows = (int)ceil_div(nzb - 2, sh.win_step);
    // 16-byte lanes: four FP32 or two FP64 values
    const bool wide = sh.v == 4 || (s->prec == PIC_FP64 && sh.v == 2);
    // Tiles per block: two in lockstep for 16-byte lanes and for rows of 8 or
    // more windows (window_group); one otherwise. A forced count is cut to
    // what a block of every window of a row holds (to 4 on longer rows).
    const int fit = std::max(1, 8 / std::min(sh.n_windows, 8));
    sh.tpb = s->force_tpb > 0 ? std::min(s->force_tpb, sh.n_windows > 8 ? 4 : fit) : (wide || sh.n_windows >= 8 ? 2 : 1);
    sh.wpb = window_group(sh.n_windows, sh.tpb, wide);
    sh.n_kg = (int)ceil_div(sh.n_windows, sh.wpb);
    const int64_t rows = no_xy - 2, planes = no_xy - 2;
    const int cus = s->cus;
    // Enough waves for ~8 per CU, with the tallest tile (fewest halo rows) and
    // the longest chunk (fewest prologue planes) that still get there.
    // Measured on MI355X over whole runs (depositing steps included), with the
    // buffer-addressed sweep: 4-row tiles of 16 planes are best for one and two
    // values per lane on 208^2 x 406 (FP32 80.5 us, against 84.4 for 2-row
    // tiles; FP64 154 us, against 162 for 2-row tiles of 8 planes) and on
    // 1066^2 x 410 (FP32 two values 1.88 ms; FP64 3.65 ms, against 4.02 for
    // 2-row tiles of 8 planes), before whole-lane stores took 4 to 8 % off the
    // FP32 figures; 12- and 24-plane chunks tie with 16 on 208^2 x 406, 14,
    // 18 and 20 lose 10-13 % (the busiest CU marches more block-planes). Small
    // grids fall through to 1-row tiles of 4 planes.
    //
    // Four values per lane: the 2-row tile (175-189 registers, two waves per
    // SIMD), with chunks shortened first until there are ~16 waves per CU --
    // two rounds of the 8 a CU holds (208^2 x 406: 8 planes, 78.5 us, against
    // 84.6 for 16). Two values per lane aim for 16 waves per CU too: on
    // 208^2 x 406, 4-row tiles of 8 planes (5408 waves) against 16 (2704),
    // measured on three boxes: 72.1 against 74.2, 71.2 against 73.2 and 75.2
    // against 78.4 us.
    const int64_t waves_wanted = (int64_t)cus * (sh.v >= 2 ? 16 : 8);
    int tj = s->force_tj;
    int chunk = s->force_chunk;
    // tallest tiles: three rows of four FP32 values, two of two FP64 values
    // (a third row spills registers in the absorbing-wall kernel)
    const int max_tj = (sh.v == 1 && s->prec == PIC_FP32) ? 8 : sh.v == 4 ? 3 : wide ? 2 : 4;
    const int chunk0 = 16;
    if (tj <= 0 && wide) tj = max_tj;
    if (tj <= 0) {
        // the tallest tile that reaches the wave count, with two values per
        // lane at chunks of 16 or, failing that, 8 planes before a shorter
        // tile is tried (208^2 x 406: 4-row tiles of 8 planes, not 2-row
        // tiles of 16, which took 78.9 against 72-75 us)
        const int c0 = chunk > 0 ? chunk : chunk0;
        const int c_min = chunk > 0 || sh.v != 2 ? c0 : 8;
        tj = 1;
        bool found = false;
        for (int c = max_tj; c >= 1 && !found; c /= 2)
            for (int ch = c0; ch >= c_min && !found; ch /= 2)
                if (ceil_div(rows, c) * sh.n_windows * ceil_div(planes, ch) >= waves_wanted) {
                    tj = c;
                    found = true;
                }
    }
    // Four values per lane: two tiles per block marching in lockstep (with
    // 2-row tiles on 1066^2 x 410: 1.58 ms per step, against 1.59 for 16 or
    // 64 planes, 1.63 and 1.60 for four tiles of 32 or 64, 1.69 for one tile
    // of 16 free-running), and tiles of three rows (1.570 against 1.588 ms
    // with 2-row pairs; 208^2 x 406, 69.7 against 74.1 us; three tiles per
    // block 1.73 ms and 91 us). The register cost of a third row keeps the two
    // waves per SIMD of 2-row tiles (223-239 against 175-189 VGPRs), and a pair
    // then reads 8 rows for 6 it writes, against 6 for 4. Tiles of four rows
    // and two values were measured slower in lockstep (208^2 x 406: 75 us
    // alone, 82-97 us in pairs or fours) and keep one tile per block on rows
    // of up to 7 windows (window_group covers longer rows).
    sh.tj = tj;
    sh.n_jt = (int)ceil_div(rows, tj);
    if (chunk <= 0 && sh.v == 4) chunk = four_value_chunk(rows, sh.n_windows, sh.wpb, tj, sh.tpb, cus);
    if (chunk <= 0 && wide) chunk = pair_chunk(sh, planes, cus);
    if (chunk <= 0) {
        chunk = chunk0;
        while (chunk > 4 && ceil_div(rows, tj) * sh.n_windows * ceil_div(planes, chunk) < waves_wanted) chunk /= 2;
        // Grids too small for 8 waves per CU with 1-row tiles of 4 planes are
        // latency-bound: every plane a block marches is one more dependent
        // memory round, so they take 2-plane chunks (the short march, which
        // runs no step past its chunk). Measured on MI355X, FP32: 39^2 x 206
        // 6.8 against 8.15 us per sweep, 21^2 x 166 6.0 against 7.4, 34^2 x
        // 210 6.3 against 6.7; FP64 39^2 x 206 8.15 against 8.8. 106^2 x 210
        // (5408 waves at 4 planes) gains 3 %, 166^2 x 210 loses 12 %.
        if (tj == 1 && chunk == 4 && sh.v <= 2 && ceil_div(rows, 1) * sh.n_windows * ceil_div(planes, 4) < (int64_t)cus * 8)
            chunk = 2;
    }
    sh.chunk = (int)std::min<int64_t>(chunk, planes);
    sh.n_ic = (int)ceil_div(planes, sh.chunk);
    return sh;
}

// The sweep kernel instantiation for a tiling, deposition flag and wall.
template <typename S, typename A, int TJ, int V>
const void *sweep_fn_tj(bool dep, bool reflect) {
    if (reflect)
        return dep ? (const void *)pic::sweep_kernel<S, A, TJ, true, V, true>
                   : (const void *)pic::sweep_kernel<S, A, TJ, false, V, true>;
    return dep ? (const void *)pic::sweep_kernel<S, A, TJ, true, V, false>
               : (const void *)pic::sweep_kernel<S, A, TJ, false, V, false>;
}

template <typename S, typename A, int V>
const void *sweep_fn_typed(int tj, bool dep, bool reflect) {
    switch (tj) {
        case 1: return sweep_fn_tj<S, A, 1, V>(dep, reflect);
        case 2: return sweep_fn_tj<S, A, 2, V>(dep, reflect);
        case 3:
            if constexpr (V == 4) {
                return sweep_fn_tj<S, A, 3, V>(dep, reflect);
            } else {
                return nullptr;
            }
        case 4:
            if constexpr (V == 4 || (V == 2 && sizeof(S) == 8)) {
                return nullptr;
            } else {
                return sweep_fn_tj<S, A, 4, V>(dep, reflect);
            }
        case 8:
            if constexpr (V >= 2 || sizeof(S) == 8) {
                return nullptr;
            } else {
                return sweep_fn_tj<S, A, 8, V>(dep, reflect);
            }
        default: return nullptr;
    }
}

// The depositing kernel of 1-row tiles marching exactly `ch` (1 or 2) planes.
template <typename S, typename A, int V>
const void *sweep_fn_short(int ch, bool reflect) {
    if (ch == 1)
        return reflect ? (const void *)pic::sweep_kernel<S, A, 1, true, V, true, 1>
                       : (const void *)pic::sweep_kernel<S, A, 1, true, V, false, 1>;
    return reflect ? (const void *)pic::sweep_kernel<S, A, 1, true, V, true, 2>
                   : (const void *)pic::sweep_kernel<S, A, 1, true, V, false, 2>;
}

// null for a combination that is not compiled (the tiling checks reject it
// first). Depositing 1-row tiles of one or two planes take the short march,
// which runs no step past the chunk (the general march runs its plane
// buffers' count); the arithmetic is the same.
const void *sweep_fn(int prec, int v, int tj, bool dep, bool reflect, int chunk = 0) {
    if (tj == 1 && dep && chunk >= 1 && chunk <= 2 && (prec == PIC_FP64 ? v == 1 : v <= 2)) {
        if (prec == PIC_FP64) return sweep_fn_short<double, double, 1>(chunk, reflect);
        return v == 2 ? sweep_fn_short<float, float, 2>(chunk, reflect) : sweep_fn_short<float, float, 1>(chunk, reflect);
    }
    if (prec == PIC_FP64)
        return v == 2 ? sweep_fn_typed<double, double, 2>(tj, dep, reflect) : sweep_fn_typed<double, double, 1>(tj, dep, reflect);
    if (v == 4) return sweep_fn_typed<float, float, 4>(tj, dep, reflect);
    return v == 2 ? sweep_fn_typed<float, float, 2>(tj, dep, reflect) : sweep_fn_typed<float, float, 1>(tj, dep, reflect);
}

template <typename S, typename A, int V>
const void *batch_fn_typed(int tj, bool reflect) {
    switch (tj) {
        case 1: return reflect ? (const void *)pic::batch_sweep_kernel<S, A, 1, V, true>
                               : (const void *)pic::batch_sweep_kernel<S, A, 1, V, false>;
        case 2: return reflect ? (const void *)pic::batch_sweep_kernel<S, A, 2, V, true>
                               : (const void *)pic::batch_sweep_kernel<S, A, 2, V, false>;
        case 4:
            if constexpr (sizeof(S) == 8) {
                return nullptr;
            } else {
                return reflect ? (const void *)pic::batch_sweep_kernel<S, A, 4, V, true>
                               : (const void *)pic::batch_sweep_kernel<S, A, 4, V, false>;
            }
        default: return nullptr;
    }
}

const void *batch_fn(int prec, int v, int tj, bool reflect, int chunk = 0) {
    if (tj == 1 && chunk >= 1 && chunk <= 2 && (prec == PIC_FP64 || v == 2)) {
        if (prec == PIC_FP64)
            return chunk == 1 ? (reflect ? (const void *)pic::batch_sweep_kernel<double, double, 1, 1, true, 1>
                                         : (const void *)pic::batch_sweep_kernel<double, double, 1, 1, false, 1>)
                              : (reflect ? (const void *)pic::batch_sweep_kernel<double, double, 1, 1, true, 2>
                                         : (const void *)pic::batch_sweep_kernel<double, double, 1, 1, false, 2>);
        return chunk == 1 ? (reflect ? (const void *)pic::batch_sweep_kernel<float, float, 1, 2, true, 1>
                                     : (const void *)pic::batch_sweep_kernel<float, float, 1, 2, false, 1>)
                          : (reflect ? (const void *)pic::batch_sweep_kernel<float, float, 1, 2, true, 2>
                                     : (const void *)pic::batch_sweep_kernel<float, float, 1, 2, false, 2>);
    }
    if (prec == PIC_FP64) return batch_fn_typed<double, double, 1>(tj, reflect);
    return v == 2 ? batch_fn_typed<float, float, 2>(tj, reflect) : batch_fn_typed<float, float, 1>(tj, reflect);
}

int launch_now(const KLaunch &l, hipStream_t st) {
    void *argv[1] = {(void *)l.arg};
    HIP_TRY(hipLaunchKernel(l.func, l.grid, l.block, argv, 0, st));
    return PIC_OK;
}

// Graph chains. A chain is a graph of kChain kernel nodes in sequence,
// instantiated once and then re-parameterised (kernel, grid and argument of
// every node) for each use, so a run's launches cost one graph launch per
// kChain kernels instead of one launch each, and the device starts each
// kernel as the previous one ends. Instances are pooled per device; one is
// re-parameterised only after its previous launch has finished, so an update
// never touches a launch in flight.
//
// Completion is read from memory, not asked of the runtime: a last node
// (chain_mark_kernel) stores the launch's number to a pinned host word. The
// HIP runtime refuses calls that are unsafe during stream capture -- event
// queries and synchronisations among them -- from every thread while any
// stream of the process is capturing, and its graph machinery may capture
// while another thread instantiates or launches a graph; an event query here
// was refused that way ("stream is capturing") when two threads drove chains
// at once. The pool therefore makes no such call: under g_chain_mu it only
// reads the marks, sets node parameters and launches graphs.
constexpr int kChain = 32;
constexpr int kChainPool = 4;

struct Chain {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t nodes[kChain + 1] = {};  // kChain kernels, then the mark
    unsigned long long *mark = nullptr;     // pinned host word the last node writes
    unsigned long long seq = 0;             // number of the last launch
    uint64_t last_use = 0;
};

std::mutex g_chain_mu;
std::vector<std::vector<Chain *>> g_chains;  // [device]
unsigned long long g_chain_uses = 0;

bool chain_done(const Chain *c) {
    return __atomic_load_n(c->mark, __ATOMIC_ACQUIRE) >= c->seq;
}

hipKernelNodeParams node_params(const KLaunch &l, void **argv) {
    hipKernelNodeParams p{};
    p.func = (void *)l.func;
    p.gridDim = l.grid;
    p.blockDim = l.block;
    p.sharedMemBytes = 0;
    argv[0] = (void *)l.arg;
    p.kernelParams = argv;
    p.extra = nullptr;
    return p;
}

KLaunch mark_launch(Chain *c, unsigned long long seq) {
    return KLaunch::make((const void *)pic::chain_mark_kernel, dim3(1), dim3(64), 2, pic::ChainMark{c->mark, seq});
}

int chain_create(const KLaunch *ls, unsigned long long seq, Chain **out) {
    Chain *c = new Chain();
    auto cleanup = [&](int code) {
        if (c->exec) hipGraphExecDestroy(c->exec);
        if (c->graph) hipGraphDestroy(c->graph);
        if (c->mark) hipHostFree(c->mark);
        delete c;
        return code;
    };
    if (hipHostMalloc((void **)&c->mark, sizeof(unsigned long long), hipHostMallocCoherent) != hipSuccess)
        return cleanup(fail(PIC_E_NOMEM, "hipHostMalloc of a chain mark failed"));
    *c->mark = 0;
    if (hipGraphCreate(&c->graph, 0) != hipSuccess) return cleanup(fail(PIC_E_HIP, "hipGraphCreate failed"));
    const KLaunch mark = mark_launch(c, seq);
    for (int i = 0; i <= kChain; ++i) {
        void *argv[1];
        const hipKernelNodeParams p = node_params(i < kChain ? ls[i] : mark, argv);
        if (hipGraphAddKernelNode(&c->nodes[i], c->graph, i ? &c->nodes[i - 1] : nullptr, i ? 1 : 0, &p) != hipSuccess)
            return cleanup(fail(PIC_E_HIP, "hipGraphAddKernelNode failed"));
    }
    if (hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0) != hipSuccess)
        return cleanup(fail(PIC_E_HIP, "hipGraphInstantiate failed"));
    c->seq = seq;
    *out = c;
    return PIC_OK;
}

// run ls[0 .. kChain) on st as one graph launch
int chain_launch(const KLaunch *ls, int device, hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_chain_mu);
    if ((int)g_chains.size() <= device) g_chains.resize(device + 1);
    std::vector<Chain *> &pool = g_chains[device];
    const unsigned long long seq = ++g_chain_uses;
    Chain *c = nullptr;
    bool fresh = false;
    for (Chain *q : pool)
        if (chain_done(q)) {
            c = q;
            break;
        }
    if (!c && (int)pool.size() < kChainPool) {
        int rc = chain_create(ls, seq, &c);
        if (rc) return rc;
        pool.push_back(c);
        fresh = true;
    }
    if (!c) {  // all in flight: wait for the least recently launched to finish
        c = *std::min_element(pool.begin(), pool.end(), [](Chain *x, Chain *y) { return x->last_use < y->last_use; });
        const auto t0 = std::chrono::steady_clock::now();
        while (!chain_done(c)) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120))
                return fail(PIC_E_HIP, "a graph chain launched earlier did not finish within 120 s");
            std::this_thread::yield();
        }
    }
    if (!fresh) {
        for (int i = 0; i <= kChain; ++i) {
            void *argv[1];
            const KLaunch mark = mark_launch(c, seq);
            hipKernelNodeParams p = node_params(i < kChain ? ls[i] : mark, argv);
            HIP_TRY(hipGraphExecKernelNodeSetParams(c->exec, c->nodes[i], &p));
        }
        c->seq = seq;
    }
    HIP_TRY(hipGraphLaunch(c->exec, st));
    c->last_use = seq;
    return PIC_OK;
}

// Run a segment's launches in order on st: whole chains through graphs, the
// remainder (fewer than kChain) as ordinary launches.
int submit(const std::vector<KLaunch> &ls, int device, hipStream_t st) {
    size_t i = 0;
    for (; ls.size() - i >= (size_t)kChain; i += kChain) {
        const int rc = chain_launch(ls.data() + i, device, st);
        if (rc) return rc;
    }
    for (; i < ls.size(); ++i) {
        const int rc = launch_now(ls[i], st);
        if (rc) return rc;
    }
    return PIC_OK;
}

// a batched launch's argument
template <typename D>
pic::BatchTable<D> batch_table(const D *descs, const int *offs, int n_cfg) {
    pic::BatchTable<D> t{};
    t.descs = descs;
    t.offs = offs;
    t.n_cfg = n_cfg;
    return t;
}

// With timing on, launches are bracketed by pairs of events on the handle's
// stream: each launch, or runs of up to timing_group back-to-back sweep
// launches; pic_get_timing sums the pairs per kind.
void record(pic_sim *s, int kind, hipStream_t st = nullptr) {
    if (!s->timing || !s->in_run) return;
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e;
        // device-scope release: a timing event must not write back the L2 to
        // system memory, which would perturb the kernels it brackets
        if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) return;
        s->ev_pool.push_back(e);
        s->ev_kind.push_back(kind);
        s->ev_count.push_back(1);
    }
    s->ev_kind[s->ev_used] = kind;
    s->ev_count[s->ev_used] = 1;
    if (hipEventRecord(s->ev_pool[s->ev_used], st ? st : s->stream) == hipSuccess) ++s->ev_used;
}

// end the open run of timed sweeps, if any
void close_group(pic_sim *s) {
    if (!s->group_open) return;
    record(s, 0);
    if (s->ev_used >= 2) s->ev_count[s->ev_used - 2] = s->group_open;
    s->group_open = 0;
}

void record_batch(pic_batch *b, int kind) {
    if (!b->timing) return;
    if (b->ev_used == b->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) return;
        b->ev_pool.push_back(e);
        b->ev_kind.push_back(kind);
    }
    b->ev_kind[b->ev_used] = kind;
    if (hipEventRecord(b->ev_pool[b->ev_used], b->stream) == hipSuccess) ++b->ev_used;
}

// A launch of handle s: appended to the open segment, or launched now on st
// (between a pair of timing events when timing is on).
int emit(pic_sim *s, const KLaunch &l, hipStream_t st) {
    if (s->sink) {
        s->sink->push_back(l);
        return PIC_OK;
    }
    if (s->timing && s->in_run && l.kind == 0 && s->timing_group > 1) {
        if (!s->group_open) record(s, 0);
        const int rc = launch_now(l, st);
        if (rc) return rc;
        if (++s->group_open == s->timing_group) close_group(s);
        return PIC_OK;
    }
    if (s->timing && s->in_run) close_group(s);  // another kind of launch ends a run of sweeps
    record(s, l.kind, st);
    const int rc = launch_now(l, st);
    if (rc) return rc;
    record(s, l.kind, st);
    return PIC_OK;
}

const void *checked_sweep_fn(const pic_sim *s, const SweepShape &sh, bool dep) {
    const void *f = sweep_fn(s->prec, sh.v, sh.tj, dep, s->p.wall == PIC_WALL_REFLECTING, sh.chunk);
    if (!f) fail(PIC_E_INVALID, "no sweep kernel for %d rows per tile with %d values per lane (%s fields)", sh.tj, sh.v,
                 s->prec == PIC_FP64 ? "FP64" : "FP32");
    return f;
}

// Grow the sweep partial buffer to hold nblk blocks (a tiling with more
// blocks than the default capacity); only between runs, no step pending.
int ensure_partials(pic_sim *s, int64_t nblk) {
    if (nblk <= s->sweep_part_cap) return PIC_OK;
    if (s->pending.out4) return fail(PIC_E_INVALID, "cannot grow the partial buffer while a step is pending");
    HIP_TRY(hipStreamSynchronize(s->stream));
    double *grown = nullptr;
    if (hipMalloc(&grown, (size_t)(2 * nblk * 3) * sizeof(double)) != hipSuccess)
        return fail(PIC_E_NOMEM, "hipMalloc of %lld sweep partials failed", (long long)nblk);
    if (s->sweep_part) hipFree(s->sweep_part);
    s->sweep_part = grown;
    s->sweep_part_cap = nblk;
    return PIC_OK;
}

// Arguments of sweep step `step` of one configuration: adds the deposition
// density dens (if dep), writes its block partials to slot step % 2 and
// becomes the pending step whose record out4 is folded later; folds the
// previously pending step at its start. Updates the pending state.
int make_sweep_args(pic_sim *s, const SweepShape &sh, int64_t step, bool dep, const double *dens,
                    const double *dep_part, int n_dep, double *out4, pic::SweepArgs &a) {
    const int64_t nblk = sh.blocks();
    if (nblk > s->sweep_part_cap) return fail(PIC_E_INVALID, "sweep grid %lld exceeds partial buffer", (long long)nblk);
    double *part = s->sweep_part + (step % 2) * s->sweep_part_cap * 3;
    a = pic::SweepArgs{};
    a.pos = s->cur_p;
    a.neg = s->cur_n;
    a.pos_out = s->nxt_p;
    a.neg_out = s->nxt_n;
    a.dens = dep ? dens : nullptr;
    a.partials = part;
    a.prev_part = s->pending.part;
    a.prev_n = s->pending.n;
    a.prev_dep_part = s->pending.dep;
    a.prev_n_dep = s->pending.n_dep;
    a.prev_out4 = s->pending.out4;
    a.no_xy = (int)s->p.no_xy;
    a.nzb = (int)s->p.nzb;
    a.pitch = (int)s->pitch;
    a.nze = (int)s->p.nze;
    a.no_z = (int)s->p.no_z;
    a.n_windows = sh.n_windows;
    a.win_step = sh.win_step;
    a.wpb = sh.wpb;
    a.tpb = sh.tpb;
    a.n_kg = sh.n_kg;
    a.n_jt = sh.n_jt;
    a.n_ic = sh.n_ic;
    a.chunk = sh.chunk;
    a.mid = s->p.mid_xy;
    a.rsq = s->p.scoring_radius_sq;
    a.p_lat = s->p.p_lat;
    a.p_zm = s->p.p_zm;
    a.p_zp = s->p.p_zp;
    a.p_cen = s->p.p_cen;
    a.n_lat = s->p.n_lat;
    a.n_zm = s->p.n_zm;
    a.n_zp = s->p.n_zp;
    a.n_cen = s->p.n_cen;
    a.alpha_dt = s->p.alpha_dt;
    s->pending = pic_sim::Pending{part, (int)nblk, dep ? dep_part : nullptr, dep ? n_dep : 0, out4};
    return PIC_OK;
}

int flush_pair(pic_sim *s, hipStream_t st);

int enqueue_sweep(pic_sim *s, int64_t step, bool dep, const double *dens, const double *dep_part, int n_dep,
                  double *out4) {
    const SweepShape sh = sweep_shape(s);
    pic::SweepArgs a;
    int rc = flush_pair(s, nullptr);  // a pass's two steps, folded before this step takes the fold slot
    if (rc) return rc;
    rc = make_sweep_args(s, sh, step, dep, dens, dep_part, n_dep, out4, a);
    if (rc) return rc;
    // one extra block folds the previous step when one is pending
    const dim3 grid((unsigned)(sh.blocks() + (a.prev_out4 ? 1 : 0))), block((unsigned)sh.threads());
    // A step without tracks runs the depositing kernel too, adding a density
    // of zeros (x + 0 == x, so the fields are the same bit for bit). Measured
    // on the 5 um full electrode: the non-depositing variant, whose march
    // issues its loads with less work between them, read 1.43x the
    // algorithmic bytes against 1.17x (the neighbouring tiles' halo rows hit
    // the L2 less often) and took 1.79 against 1.68 ms per step.
    a.dens = s->zero_dens;
    if (dep) a.dens = dens;
    const void *f = checked_sweep_fn(s, sh, true);
    if (!f) return PIC_E_INVALID;
    return emit(s, KLaunch::make(f, grid, block, 0, a), s->stream);
}

// Two steps per pass (sweep2_kernel): blocks of 16 one-row waves, 12 output
// rows, one k window of 240 values (248 in the first), `chunk` planes.
struct Pass2Shape {
    int n_windows = 0, n_jg = 0, n_ic = 0, chunk = 0;
    int64_t blocks() const { return (int64_t)n_windows * n_jg * n_ic; }
};

// Planes per block of the two-step pass: one 16-wave block per CU, so the
// chunk whose busiest CU marches the fewest planes, counted as rounds of
// blocks x (chunk + 3) (a pass loads three planes before its first output).
// Measured on MI355X (sweep time per step, each grid against single steps on
// one device): 1066^2 x 410, 40 planes 1.41 ms against 1.56 (64: 1.41);
// 726^2 x 410, 24 planes 689 us against 727 (40: 718, 64: 691); 506^2 x 1010,
// 40 planes 803 against 965 (24: 811, 16: 838); 486^2 x 410, 16-24 planes
// 323 against 331 (40: 357); 536^2 x 210, 16 planes 209 against 226 (24: 228);
// 246^2 x 210, 24 planes 40.6 against 52.7 (8: 50.1, 16: 54.0); 206^2 x 410,
// 16 planes 58.6 against 66.0 (8: 67.2, 24: 78.5).
Pass2Shape pass2_shape(const pic_sim *s) {
    Pass2Shape sh;
    const int64_t lateral = s->p.no_xy - 2;
    sh.n_windows = pic::sweep2_windows((int)s->p.nzb);
    sh.n_jg = (int)ceil_div(lateral, pic::kS2Rows);
    auto with_chunk = [&](int c) {
        Pass2Shape t = sh;
        t.chunk = (int)std::min<int64_t>(c, lateral);
        t.n_ic = (int)ceil_div(lateral, t.chunk);
        return t;
    };
    if (s->pass2_chunk > 0) return with_chunk(s->pass2_chunk);
    Pass2Shape best = with_chunk(8);
    double best_cost = -1;
    for (int c : {8, 12, 16, 24, 32, 40, 48, 64}) {
        const Pass2Shape t = with_chunk(c);
        const double cost = (double)ceil_div(t.blocks(), s->cus) * (t.chunk + 3);
        if (best_cost < 0 || cost < best_cost) {
            best = t;
            best_cost = cost;
        }
    }
    return best;
}

// why a handle cannot take two steps per pass, or null
const char *pass2_unsupported(const pic_sim *s) {
    if (s->prec != PIC_FP32) return "two steps per pass need FP32 fields";
    if (s->p.no_xy < 4 || s->p.nzb < 4) return "two steps per pass need at least 2 interior voxels per axis";
    return nullptr;
}

// Whether run_steps pairs steps: when asked to, or by default where the pass
// is possible, no single-step tiling is forced, and its grid has at least 32
// (row group, window) columns of blocks -- every grid measured from 206^2 x
// 410 up ran faster per step.
bool pass2_on(const pic_sim *s) {
    if (s->steps_per_pass == 1 || pass2_unsupported(s)) return false;
    if (s->steps_per_pass == 2) return true;
    if (s->force_tj || s->force_chunk || s->force_v || s->force_tpb) return false;
    const Pass2Shape sh = pass2_shape(s);
    return (int64_t)sh.n_jg * sh.n_windows >= 32;
}

// The pass's partials buffer, grown to hold nblk blocks (between runs only).
int ensure_pair_partials(pic_sim *s, int64_t nblk) {
    if (nblk <= s->pair_part_cap) return PIC_OK;
    if (s->pair_open) return fail(PIC_E_INVALID, "cannot grow the pass partials while a pass is pending");
    HIP_TRY(hipStreamSynchronize(s->stream));
    double *grown = nullptr;
    if (hipMalloc(&grown, (size_t)(4 * nblk * 3) * sizeof(double)) != hipSuccess)
        return fail(PIC_E_NOMEM, "hipMalloc of %lld two-step partials failed", (long long)nblk);
    if (s->pair_part) hipFree(s->pair_part);
    s->pair_part = grown;
    s->pair_part_cap = nblk;
    return PIC_OK;
}

// fold the last pass's two steps into their records (one launch, a block each)
int flush_pair(pic_sim *s, hipStream_t st) {
    if (!s->pair_open) return PIC_OK;
    pic::Fold2Args f{};
    f.step[0] = s->pair_pending[0];
    f.step[1] = s->pair_pending[1];
    s->pair_open = false;
    return emit(s, KLaunch::make((const void *)pic::fold2_kernel, dim3(2), dim3(256), 2, f), st ? st : s->stream);
}

// Steps `step` and `step + 1` in one sweep2_kernel launch. Its first blocks
// fold what the previous launch left (a single step, or a pass's two steps);
// its own two steps are folded by the next launch or the run's end.
int enqueue_sweep2(pic_sim *s, int64_t step, const bool dep[2], const double *dens[2], const double *dep_part[2],
                   int n_dep, double *out4[2]) {
    const Pass2Shape sh = pass2_shape(s);
    const int64_t nblk = sh.blocks();
    if (nblk > s->pair_part_cap) return fail(PIC_E_INVALID, "two-step grid %lld exceeds its partial buffer", (long long)nblk);
    pic::Sweep2Args a{};
    a.n_prev = 0;
    if (s->pending.out4) {
        a.prev[a.n_prev++] = pic::FoldArgs{s->pending.part, s->pending.dep, s->pending.out4, s->pending.n, s->pending.n_dep};
        s->pending = pic_sim::Pending{};
    }
    if (s->pair_open) {
        a.prev[a.n_prev++] = s->pair_pending[0];
        a.prev[a.n_prev++] = s->pair_pending[1];
        s->pair_open = false;
    }
    double *base = s->pair_part + (int64_t)s->pair_slot * 2 * s->pair_part_cap * 3;
    s->pair_slot ^= 1;
    a.pos = static_cast<const float *>(s->cur_p);
    a.neg = static_cast<const float *>(s->cur_n);
    a.pos_out = static_cast<float *>(s->nxt_p);
    a.neg_out = static_cast<float *>(s->nxt_n);
    a.dens1 = dep[0] ? dens[0] : s->zero_dens;
    a.dens2 = dep[1] ? dens[1] : s->zero_dens;
    a.part1 = base;
    a.part2 = base + s->pair_part_cap * 3;
    a.no_xy = (int)s->p.no_xy;
    a.nzb = (int)s->p.nzb;
    a.nze = (int)s->p.nze;
    a.no_z = (int)s->p.no_z;
    a.pitch = (int)s->pitch;
    a.n_windows = sh.n_windows;
    a.n_jg = sh.n_jg;
    a.n_ic = sh.n_ic;
    a.chunk = sh.chunk;
    a.mid = s->p.mid_xy;
    a.rsq = s->p.scoring_radius_sq;
    a.p_lat = s->p.p_lat;
    a.p_zm = s->p.p_zm;
    a.p_zp = s->p.p_zp;
    a.n_lat = s->p.n_lat;
    a.n_zm = s->p.n_zm;
    a.n_zp = s->p.n_zp;
    a.alpha_dt = s->p.alpha_dt;
    a.reflect = s->p.wall == PIC_WALL_REFLECTING ? 1 : 0;
    const int rc = emit(s,
                        KLaunch::make((const void *)pic::sweep2_kernel, dim3((unsigned)(nblk + a.n_prev)),
                                      dim3(64 * pic::kS2Waves), 0, a),
                        s->stream);
    if (rc) return rc;
    for (int q = 0; q < 2; ++q)
        s->pair_pending[q] = pic::FoldArgs{q == 0 ? a.part1 : a.part2, dep[q] ? dep_part[q] : nullptr, out4[q],
                                           (int)nblk, dep[q] ? n_dep : 0};
    s->pair_open = true;
    return PIC_OK;
}

// fold the pending step's partials into its record (end of a run of steps)
int flush_pending(pic_sim *s, hipStream_t st = nullptr) {
    if (s->pair_open) return flush_pair(s, st);  // a pass and a single step are never pending together
    if (!s->pending.out4) return PIC_OK;
    close_group(s);
    record(s, 2);
    hipLaunchKernelGGL(pic::reduce_kernel, dim3(1), dim3(256), 0, st ? st : s->stream, s->pending.part, s->pending.n,
                       s->pending.dep, s->pending.n_dep, s->pending.out4);
    HIP_LAUNCH_CHECK();
    record(s, 2);
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    return PIC_OK;
}

int64_t gather_blocks(const pic_sim *s) {
    const int64_t nt = ceil_div(s->p.no_xy, pic::kGT);
    return nt * nt;
}

// Density and deposition-partial slots: step t uses slot t % dens_slots. Two
// gather batches and two steps: the gather of the next segment runs while this
// segment's sweeps read theirs, and this segment's first launch still folds
// the previous segment's last step, or its last two after a two-step pass.
int64_t dens_slots(int64_t batch) { return 2 * batch + 2; }

// arguments of the gather of steps from `first` (slot = step mod dens_slots);
// offsets are device offsets into xs/ys
pic::GatherArgs make_gather_args(const pic_sim *s, const double *xs, const double *ys, const int64_t *offsets,
                                 int64_t first) {
    pic::GatherArgs g{};
    g.xs = xs;
    g.ys = ys;
    g.offsets = offsets;
    g.first_step = first;
    g.n_slots = dens_slots(s->batch);
    g.dens = s->dens;
    g.partials = s->dep_part;
    g.no_xy = (int)s->p.no_xy;
    g.no_z = (int)s->p.no_z;
    g.ntj = (int)ceil_div(s->p.no_xy, pic::kGT);
    g.cutoff = s->p.cutoff_voxels;
    g.h2b2 = s->p.h2_over_b2;
    g.gf = s->p.gaussian_factor;
    g.mid = s->p.mid_xy;
    g.rsq = s->p.scoring_radius_sq;
    return g;
}

// deposition densities of steps [first, first+count) (count <= batch) into
// their dens and dep_part slots. With `offsets_h` (the host copy of the
// step offsets) and the handle's bins ready, the tracks are first binned by
// tile (clear, count, scan, fill) on the same stream, and each gather block
// reads only its bin; otherwise every block scans its step's tracks. The
// results are the same bit for bit.
int enqueue_gather(pic_sim *s, const double *xs, const double *ys, const int64_t *offsets, int64_t first,
                   int64_t count, hipStream_t st = nullptr, const int64_t *offsets_h = nullptr) {
    st = st ? st : s->stream;
    pic::GatherArgs g = make_gather_args(s, xs, ys, offsets, first);
    if (offsets_h && s->bins_ready) {
        const int64_t tiles = gather_blocks(s);
        int64_t most = 0;
        for (int64_t t = first; t < first + count; ++t) most = std::max(most, offsets_h[t + 1] - offsets_h[t]);
        pic::BinArgs b{};
        b.xs = xs;
        b.ys = ys;
        b.offsets = offsets;
        b.first_step = first;
        b.cnt = s->bin_cnt;
        b.off = s->bin_off;
        b.cur = s->bin_cur;
        b.tot = s->bin_tot;
        b.list = s->bin_list;
        b.n_cnt = count * tiles;
        b.no_xy = (int)s->p.no_xy;
        b.ntj = (int)ceil_div(s->p.no_xy, pic::kGT);
        b.cutoff = s->p.cutoff_voxels;
        const unsigned clear_blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(b.n_cnt, 256), 1024));
        const dim3 per_track((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(most, 256), 64)), (unsigned)count);
        const KLaunch passes[] = {
            KLaunch::make((const void *)pic::bin_clear_kernel, dim3(clear_blocks), dim3(256), 1, b),
            KLaunch::make((const void *)pic::bin_count_kernel, per_track, dim3(256), 1, b),
            KLaunch::make((const void *)pic::bin_scan_kernel, dim3((unsigned)count), dim3(256), 1, b),
            KLaunch::make((const void *)pic::bin_fill_kernel, per_track, dim3(256), 1, b),
        };
        for (const KLaunch &l : passes) {
            const int rc = emit(s, l, st);
            if (rc) return rc;
        }
        g.bins.cnt = s->bin_cnt;
        g.bins.off = s->bin_off;
        g.bins.tot = s->bin_tot;
        g.bins.list = s->bin_list;
    }
    return emit(s, KLaunch::make((const void *)pic::gather_kernel, dim3((unsigned)gather_blocks(s), (unsigned)count),
                                 dim3(256), 1, g),
                st);
}

// Tiles a track's stencil box can meet, at most: a box spans at most
// floor(2 cutoff) + 1 cells per axis, hence (w - 1) / 16 + 2 tiles.
int64_t bin_tiles_per_track(const pic_sim *s) {
    const int64_t ntj = ceil_div(s->p.no_xy, pic::kGT);
    const double w = std::floor(2.0 * s->p.cutoff_voxels) + 1.0;
    const int64_t axis = w >= (double)(ntj * pic::kGT) ? ntj : std::min<int64_t>(ntj, ((int64_t)w - 1) / pic::kGT + 2);
    return axis * axis;
}

void free_bins(pic_sim *s) {
    int **bufs[] = {&s->bin_cnt, &s->bin_off, &s->bin_cur, &s->bin_tot, &s->bin_list};
    for (int **b : bufs) {
        if (*b) hipFree(*b);
        *b = nullptr;
    }
    s->bins_ready = false;
    s->bin_cap = 0;
}

// Size the bins for the uploaded tracks when the gather mode asks for them
// (automatic: grids of at least 256 gather tiles, where scanning every track
// per tile costs most -- 4489 tiles x 6646 tracks per step on the 5 um full
// electrode). Capacity: the most entries any window of `batch` consecutive
// steps can hold. A list that would pass 1 GB (a huge cutoff) keeps the scan.
int ensure_bins(pic_sim *s) {
    free_bins(s);
    if (s->gather_mode == PIC_GATHER_SCAN || s->n_steps <= 0) return PIC_OK;
    const int64_t tiles = gather_blocks(s);
    if (s->gather_mode == PIC_GATHER_AUTO && tiles < 256) return PIC_OK;
    int64_t most = 0;
    for (int64_t t = 0; t < s->n_steps; ++t)
        most = std::max(most, s->offsets[std::min(t + s->batch, s->n_steps)] - s->offsets[t]);
    const int64_t per = bin_tiles_per_track(s);
    const int64_t cap = std::max<int64_t>(1, most) * per;
    if (most > INT_MAX / std::max<int64_t>(per, 1) || cap * 4 > (1LL << 30)) {
        if (s->gather_mode == PIC_GATHER_BINNED)
            return fail(PIC_E_INVALID, "binned deposition would need %lld list entries (over 1 GB); use the scan",
                        (long long)cap);
        return PIC_OK;
    }
    const size_t counters = (size_t)s->batch * tiles * sizeof(int);
    const size_t need = 3 * counters + (size_t)s->batch * sizeof(int) + (size_t)cap * sizeof(int);
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if ((double)need > 0.9 * (double)fr) {
        if (s->gather_mode == PIC_GATHER_BINNED)
            return fail(PIC_E_NOMEM, "binned deposition needs %.3f GiB, %.3f GiB free", need / 1073741824.0,
                        fr / 1073741824.0);
        return PIC_OK;
    }
    if (hipMalloc(&s->bin_cnt, counters) != hipSuccess || hipMalloc(&s->bin_off, counters) != hipSuccess ||
        hipMalloc(&s->bin_cur, counters) != hipSuccess || hipMalloc(&s->bin_tot, (size_t)s->batch * sizeof(int)) != hipSuccess ||
        hipMalloc(&s->bin_list, (size_t)cap * sizeof(int)) != hipSuccess) {
        free_bins(s);
        return fail(PIC_E_NOMEM, "hipMalloc of the deposition bins failed");
    }
    s->bin_cap = cap;
    s->bins_ready = true;
    return PIC_OK;
}

int enqueue_deposit_reduce(pic_sim *s, double *out4) {
    hipLaunchKernelGGL(pic::reduce_kernel, dim3(1), dim3(256), 0, s->stream, s->sweep_part, 0, s->dep_part,
                       (int)gather_blocks(s), out4);  // slot 0: step 0 of a one-step gather
    HIP_LAUNCH_CHECK();
    return PIC_OK;
}

void swap_buffers(pic_sim *s) {
    std::swap(s->cur_p, s->nxt_p);
    std::swap(s->cur_n, s->nxt_n);
}

// Elements from one (i, j) row to the next: FP32 rows are padded to an even
// length, so every lane's first value is 8-byte aligned (four-value lanes
// then load 16 bytes at 8-byte alignment, which buffer loads allow). The
// pitch also sets how rows and planes fall on HBM channels, and every padding
// beyond the least measured slower: 1066^2 x 410 with four values per lane
// 1.43-1.48 ms at pitch 410 against 1.46-1.52 at 412 (same box), 1.72 at
// 420, 1.80 at 428, 1.92-1.96 at 436..456; with two values 1.74 ms at 410,
// 1.83 at 412 and 416; 2656^2 x 1010, 26.9 ms at 1010 against 40.8 at 1012.
int64_t row_pitch(const pic_params *p, int precision) {
    if (precision != PIC_FP32) return p->nzb;
    return (p->nzb + 1) / 2 * 2;
}

int check_params(const pic_params *p, int precision) {
    if (!p) return fail(PIC_E_INVALID, "params is NULL");
    if (precision != PIC_FP64 && precision != PIC_FP32) return fail(PIC_E_INVALID, "precision must be PIC_FP64 or PIC_FP32");
    if (p->no_xy < 3 || p->nzb < 3) return fail(PIC_E_INVALID, "grid too small: no_xy=%lld nzb=%lld", (long long)p->no_xy, (long long)p->nzb);
    if (p->no_xy > 65535 || p->nzb > 65535) return fail(PIC_E_INVALID, "grid axis exceeds 65535");
    if (p->nze < 0 || p->no_z < 0 || p->nze + p->no_z > p->nzb) return fail(PIC_E_INVALID, "gap layers outside the grid");
    if (p->wall != PIC_WALL_ABSORBING && p->wall != PIC_WALL_REFLECTING) return fail(PIC_E_INVALID, "wall must be PIC_WALL_ABSORBING or PIC_WALL_REFLECTING");
    if (!(p->cutoff_voxels > 0.0)) return fail(PIC_E_INVALID, "cutoff_voxels must be positive");
    // the sweep addresses one i plane of an array through a buffer resource
    // with 32-bit offsets
    const int64_t plane_bytes = p->no_xy * row_pitch(p, precision) * (precision == PIC_FP64 ? 8 : 4);
    if (plane_bytes + pic::kFieldSlack >= INT32_MAX)
        return fail(PIC_E_INVALID, "one i plane (%lld x %lld) exceeds 2 GB", (long long)p->no_xy, (long long)p->nzb);
    return PIC_OK;
}

// steps per gather launch: as many as fit in ~256 MB of density, at most
// 128, and an even count where more than one fits, so that two steps per
// pass pair every step of a segment
int64_t batch_steps(const pic_params *p) {
    const int64_t per = p->no_xy * p->no_xy * 8;
    const int64_t n = std::max<int64_t>(1, std::min<int64_t>(128, (int64_t)(256LL << 20) / per));
    return n > 1 ? n & ~(int64_t)1 : n;
}

size_t partial_cap_bytes(const pic_params *p) {
    // generous upper bound on sweep blocks: one per (row, window, 8 planes)
    const int64_t rows = p->no_xy - 2;
    const int64_t windows = ceil_div(p->nzb - 2, pic::kOut);
    const int64_t cap = rows * windows * std::max<int64_t>(1, ceil_div(rows, 8)) + 1024;
    return (size_t)cap * 3 * sizeof(double);
}

// Tiling of a batch: one block size (wpb waves) and one value layout for
// every configuration; the tile rows and planes per block are chosen per step
// from a short candidate list -- the first candidate whose blocks give the
// configurations running at that step at least waves_min waves. Measured on
// MI355X over a 32-point saturation curve (16^2..47^2 x 106..206 grids):
// 2-row tiles of 4 planes are fastest while eight or more grids run, 1-row
// tiles of 4 planes once a few are left, and 1-row tiles of 2 planes (the
// short march) for the last few; taller tiles lose to the register cost of
// the multi-configuration kernel, longer chunks to the march length.
int batch_tiling(pic_batch *b) {
    int v = 2, wpb = 8;
    for (pic_sim *s : b->sims) v = std::min(v, values_per_lane(s, 0));
    if (b->prec == PIC_FP64) v = 1;  // the batched FP64 sweep takes one value per lane
    for (pic_sim *s : b->sims)
        wpb = std::min(wpb, (int)ceil_div(s->p.nzb - 2, pic::window_step(v, s->p.nzb - 2)));
    const int cus = b->sims[0]->cus;
    std::vector<std::pair<int, int>> cands;
    if (b->force_tj > 0 || b->force_chunk > 0)
        cands = {{b->force_tj > 0 ? b->force_tj : 1, b->force_chunk > 0 ? b->force_chunk : 4}};
    else
        cands = {{2, 4}, {1, 4}, {1, 2}};
    std::vector<std::vector<SweepShape>> shapes(cands.size());
    for (size_t ci = 0; ci < cands.size(); ++ci)
        for (pic_sim *s : b->sims) shapes[ci].push_back(fixed_shape(s, v, cands[ci].first, cands[ci].second, wpb));
    for (size_t c = 0; c < b->sims.size(); ++c) {
        int64_t most = 0;
        for (size_t ci = 0; ci < cands.size(); ++ci) most = std::max(most, shapes[ci][c].blocks());
        const int rc = ensure_partials(b->sims[c], most);
        if (rc) return rc;
    }
    b->v = v;
    b->wpb = wpb;
    b->waves_min = (int64_t)cus * 6;
    b->cands = cands;
    b->shapes = shapes;
    return PIC_OK;
}

void free_batch_slots(pic_batch *b) {
    for (int k = 0; k < 2; ++k) {
        if (b->descs_h[k]) hipHostFree(b->descs_h[k]);
        if (b->offs_h[k]) hipHostFree(b->offs_h[k]);
        if (b->gdesc_h[k]) hipHostFree(b->gdesc_h[k]);
        if (b->goffs_h[k]) hipHostFree(b->goffs_h[k]);
        if (b->descs_d[k]) hipFree(b->descs_d[k]);
        if (b->offs_d[k]) hipFree(b->offs_d[k]);
        if (b->gdesc_d[k]) hipFree(b->gdesc_d[k]);
        if (b->goffs_d[k]) hipFree(b->goffs_d[k]);
        b->descs_h[k] = b->descs_d[k] = nullptr;
        b->offs_h[k] = b->offs_d[k] = nullptr;
        b->gdesc_h[k] = b->gdesc_d[k] = nullptr;
        b->goffs_h[k] = b->goffs_d[k] = nullptr;
        b->slot_busy[k] = false;
    }
    b->slot_steps = b->slot_cfgs = 0;
}

// argument-table slots for segments of `steps` steps (stream idle)
int ensure_batch_slots(pic_batch *b, int64_t steps) {
    const int64_t n = (int64_t)b->sims.size();
    if (steps <= b->slot_steps && n <= b->slot_cfgs) return PIC_OK;
    free_batch_slots(b);
    for (int k = 0; k < 2; ++k) {
        if (hipHostMalloc((void **)&b->descs_h[k], (size_t)(steps * n) * sizeof(pic::SweepArgs)) != hipSuccess ||
            hipHostMalloc((void **)&b->offs_h[k], (size_t)(steps * (n + 1)) * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->descs_d[k], (size_t)(steps * n) * sizeof(pic::SweepArgs)) != hipSuccess ||
            hipMalloc(&b->offs_d[k], (size_t)(steps * (n + 1)) * sizeof(int)) != hipSuccess ||
            hipHostMalloc((void **)&b->gdesc_h[k], (size_t)n * sizeof(pic::GatherArgs)) != hipSuccess ||
            hipHostMalloc((void **)&b->goffs_h[k], (size_t)(n + 1) * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->gdesc_d[k], (size_t)n * sizeof(pic::GatherArgs)) != hipSuccess ||
            hipMalloc(&b->goffs_d[k], (size_t)(n + 1) * sizeof(int)) != hipSuccess) {
            free_batch_slots(b);
            return fail(PIC_E_NOMEM, "allocation of the batch argument tables failed");
        }
    }
    b->slot_steps = steps;
    b->slot_cfgs = n;
    return PIC_OK;
}

// --- batch lanes -----------------------------------------------------

// The lanes of a batch, by run length: part[c] = 0 for the configurations
// running over 0.8 of the longest run, 1 for those over 0.36 of it, 2 for the
// rest. The long lane needs at least four configurations in all and takes at
// most a quarter of them; the middle lane is kept only when it and the last
// lane both hold two configurations or more, else it joins the last. Returns
// the number of lanes, 0 when there is no split. Measured on the 32-point
// saturation curve, lanes driven from Python threads with the split switched
// off, median of five passes: one lane 36.9 ms; the longest run apart 33.1;
// the longest (2505 steps) | the eight of 927-1993 steps | the 23 of at most
// 883, 31.1 (16.5, 20.5 and 15.0 ms alone); four lanes balanced by a cost
// model 32.3; six or eight lanes 44.7-45.4 (more streams than the four
// hardware queues of a process).
int lane_split(const pic_batch *b, std::vector<int> &part) {
    const int n = (int)b->sims.size();
    if (n < 4) return 0;
    int64_t longest = 0;
    for (const pic_sim *s : b->sims) longest = std::max(longest, s->n_steps);
    part.assign(n, 2);
    int count[3] = {0, 0, 0};
    for (int c = 0; c < n; ++c) {
        const double r = (double)b->sims[c]->n_steps / (double)longest;
        part[c] = r > 0.8 ? 0 : r > 0.36 ? 1 : 2;
        ++count[part[c]];
    }
    if (count[0] < 1 || count[0] > n / 4) return 0;
    if (count[1] < 2 || count[2] < 2) {  // two lanes: the long one and the rest
        for (int c = 0; c < n; ++c)
            if (part[c] == 2) part[c] = 1;
        return 2;
    }
    return 3;
}

// Run steps [first, first+count) as lanes (sub-batches over the split `part`
// of n_lanes lanes), each driven by its own host thread on its own stream; the
// batch's stream then waits for all of them, and each handle's stream for its
// lane.
int run_lanes(pic_batch *b, const std::vector<int> &part, int n_lanes, int64_t first, int64_t count) {
    bool built = part == b->lane_of && n_lanes == b->n_lanes;
    for (int L = 0; L < n_lanes && built; ++L) built = b->lanes[L] != nullptr;
    if (!built) {
        for (pic_batch *&l : b->lanes) {
            pic_batch_destroy(l);
            l = nullptr;
        }
        b->lane_of.clear();
        b->n_lanes = 0;
        for (int L = 0; L < n_lanes; ++L) {
            std::vector<pic_sim *> members;
            for (size_t c = 0; c < b->sims.size(); ++c)
                if (part[c] == L) members.push_back(b->sims[c]);
            int rc = pic_batch_create(members.data(), (int)members.size(), &b->lanes[L]);
            if (!rc && (b->force_tj > 0 || b->force_chunk > 0))
                rc = pic_batch_set_tiling(b->lanes[L], b->force_tj, b->force_chunk);
            if (rc) return rc;
            b->lanes[L]->is_lane = true;
            b->lanes[L]->graphs = b->graphs;
        }
        b->lane_of = part;
        b->n_lanes = n_lanes;
    }
    int rc[pic_batch::kMaxLanes] = {PIC_OK, PIC_OK, PIC_OK};
    std::string err[pic_batch::kMaxLanes];
    auto go = [&](int L) {
        pic_batch *l = b->lanes[L];
        int64_t longest = 0;
        for (const pic_sim *s : l->sims) longest = std::max(longest, s->n_steps);
        const int64_t c = std::min(first + count, longest) - first;
        if (c > 0) rc[L] = pic_batch_run(l, first, c);
        if (rc[L]) err[L] = g_err;
    };
    // lanes 0 .. n_lanes-2 on worker threads, the last on this one
    const int device = b->sims[0]->device;
    std::thread workers[pic_batch::kMaxLanes];
    bool threaded[pic_batch::kMaxLanes] = {false, false, false};
    for (int L = 0; L + 1 < n_lanes; ++L) {
        try {
            workers[L] = std::thread([&, L] {
                if (hipSetDevice(device) != hipSuccess) {
                    rc[L] = fail(PIC_E_HIP, "hipSetDevice failed in a lane thread");
                    err[L] = g_err;
                    return;
                }
                go(L);
            });
            threaded[L] = true;
        } catch (...) {  // no thread to be had: that lane from this one, after the last
        }
    }
    go(n_lanes - 1);
    for (int L = 0; L + 1 < n_lanes; ++L) {
        if (threaded[L]) workers[L].join();
        else go(L);
    }
    for (int L = 0; L < n_lanes; ++L)
        if (rc[L]) {
            snprintf(g_err, sizeof g_err, "%s", err[L].c_str());
            return rc[L];
        }
    for (int L = 0; L < n_lanes; ++L) HIP_TRY(hipStreamWaitEvent(b->stream, b->lanes[L]->done, 0));
    return PIC_OK;
}

// --- static scratch --------------------------------------------------

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on `device` (which must be current):
// looked up once per device with hipGetSymbolAddress, then reused.
int static_scratch(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return fail(PIC_E_INVALID, "device %d outside 0..%d", device, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PIC_E_HIP, "hipGetSymbolAddress of the static scratch failed: %s", hipGetErrorString(e));
        g_scratch_dev[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_dev[device];
    return PIC_OK;
}

// --- adjoint helpers -------------------------------------------------

// i planes per adjoint block: as few as keep the grid within one wave of
// resident blocks (each block marches its chunk; a second wave of blocks would
// serialise behind the first), at least 2 and at most 16
int adjoint_chunk(const pic_sim *s) {
    int occ = 0;
    const void *f = s->p.wall == PIC_WALL_REFLECTING ? (const void *)pic::adjoint_kernel<true>
                                                      : (const void *)pic::adjoint_kernel<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, 256, 0) != hipSuccess || occ < 1) occ = 1;
    const int64_t per_plane = ceil_div(s->p.nzb, pic::kWave) * ceil_div(s->p.no_xy, 4);
    const int64_t chunk = ceil_div(per_plane * s->p.no_xy, (int64_t)occ * s->cus);
    return (int)std::min<int64_t>(std::max<int64_t>(chunk, 2), 16);
}

int64_t adjoint_blocks(const pic_adjoint *ad) {
    const pic_sim *s = ad->s;
    return ceil_div(s->p.nzb, pic::kWave) * ceil_div(s->p.no_xy, 4) * ceil_div(s->p.no_xy, ad->chunk);
}

// one forward sweep of step t from store slot src to slot src+1, with the
// step's deposition (dens slot of the sim) when it has tracks; the step's
// record is not kept
int adjoint_recompute_sweep(pic_adjoint *ad, int64_t t, int64_t src) {
    pic_sim *s = ad->s;
    const SweepShape sh = sweep_shape(s);
    const int64_t xy = s->p.no_xy * s->p.no_xy;
    const bool dep = s->offsets[t + 1] > s->offsets[t];
    const int n_tiles = (int)gather_blocks(s);
    const int64_t slot = t % dens_slots(s->batch);
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    pic::SweepArgs a;
    int rc = make_sweep_args(s, sh, 0, dep, s->dens + slot * xy, s->dep_part + slot * n_tiles, n_tiles, ad->scratch, a);
    if (rc) return rc;
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    double *in = ad->store + src * 2 * s->n_vox;
    a.pos = in;
    a.neg = in + s->n_vox;
    a.pos_out = in + 2 * s->n_vox;
    a.neg_out = in + 3 * s->n_vox;
    a.prev_out4 = nullptr;
    const dim3 grid((unsigned)sh.blocks()), block((unsigned)sh.threads());
    const void *f = checked_sweep_fn(s, sh, dep);
    if (!f) return PIC_E_INVALID;
    return launch_now(KLaunch::make(f, grid, block, 0, a), s->stream);
}

int adjoint_deriv_gather(pic_adjoint *ad, int64_t first, int64_t count) {
    pic_sim *s = ad->s;
    pic::GatherArgs g{};
    g.xs = s->xs;
    g.ys = s->ys;
    g.offsets = s->offsets_d;
    g.first_step = first;
    g.n_slots = dens_slots(s->batch);
    g.no_xy = (int)s->p.no_xy;
    g.no_z = (int)s->p.no_z;
    g.ntj = (int)ceil_div(s->p.no_xy, pic::kGT);
    g.cutoff = s->p.cutoff_voxels;
    g.h2b2 = s->p.h2_over_b2;
    g.gf = s->p.gaussian_factor;
    g.mid = s->p.mid_xy;
    g.rsq = s->p.scoring_radius_sq;
    g.dgf = ad->dgf;
    g.dh = ad->dh;
    g.dpart = ad->dpart;
    const int tiles = (int)gather_blocks(s);
    hipLaunchKernelGGL(pic::gather_deriv_kernel, dim3((unsigned)tiles, (unsigned)count), dim3(256), 0, s->stream, g);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(pic::fold_rows_kernel<2>, dim3((unsigned)count), dim3(256), 0, s->stream, ad->dpart, tiles, 2,
                       ad->dinj_rec + 2 * first);
    HIP_LAUNCH_CHECK();
    return PIC_OK;
}

// adjoint step t: mu_t from mu_{t+1} (buffer `cur`) into buffer cur ^ 1
int adjoint_step(pic_adjoint *ad, int64_t t, int64_t src, int cur) {
    pic_sim *s = ad->s;
    const int64_t xy = s->p.no_xy * s->p.no_xy;
    const bool dep = s->offsets[t + 1] > s->offsets[t];
    const int64_t nblk = adjoint_blocks(ad);
    pic::AdjointArgs a{};
    const double *in = ad->store + src * 2 * s->n_vox;
    a.sp = in;
    a.sn = in + s->n_vox;
    a.dens = dep ? s->dens + (t % dens_slots(s->batch)) * xy : ad->zeros;
    a.dgf = dep ? ad->dgf + (t % ad->K) * xy : ad->zeros;
    a.dh = dep ? ad->dh + (t % ad->K) * xy : ad->zeros;
    a.mp = ad->mu + (2 * cur + 0) * s->n_vox;
    a.mn = ad->mu + (2 * cur + 1) * s->n_vox;
    a.op = ad->mu + (2 * (cur ^ 1) + 0) * s->n_vox;
    a.on = ad->mu + (2 * (cur ^ 1) + 1) * s->n_vox;
    a.acc = ad->acc;
    a.no_xy = (int)s->p.no_xy;
    a.nzb = (int)s->p.nzb;
    a.nze = (int)s->p.nze;
    a.no_z = (int)s->p.no_z;
    a.pitch = (int)s->pitch;
    a.n_kb = (int)ceil_div(s->p.nzb, pic::kWave);
    a.n_jb = (int)ceil_div(s->p.no_xy, 4);
    a.chunk = ad->chunk;
    a.mid = s->p.mid_xy;
    a.rsq = s->p.scoring_radius_sq;
    a.p_lat = s->p.p_lat;
    a.p_zm = s->p.p_zm;
    a.p_zp = s->p.p_zp;
    a.p_cen = s->p.p_cen;
    a.n_lat = s->p.n_lat;
    a.n_zm = s->p.n_zm;
    a.n_zp = s->p.n_zp;
    a.n_cen = s->p.n_cen;
    a.adt = s->p.alpha_dt;
    const dim3 grid((unsigned)nblk), block(256);
    if (s->p.wall == PIC_WALL_REFLECTING)
        hipLaunchKernelGGL(pic::adjoint_kernel<true>, grid, block, 0, s->stream, a);
    else
        hipLaunchKernelGGL(pic::adjoint_kernel<false>, grid, block, 0, s->stream, a);
    HIP_LAUNCH_CHECK();
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return PIC_E_INVALID;
    snprintf(buf, len, "%s", g_err);
    return PIC_OK;
}

int pic_device_info(char *name, size_t name_len, size_t *free_bytes, size_t *total_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PIC_E_NODEVICE, "no HIP device is visible");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = tot;
    return PIC_OK;
}

int pic_mem_required(const pic_params *p, int precision, size_t *bytes) {
    int rc = check_params(p, precision);
    if (rc) return rc;
    if (!bytes) return fail(PIC_E_INVALID, "bytes is NULL");
    const size_t elem = precision == PIC_FP64 ? 8 : 4;
    const size_t vox = (size_t)p->no_xy * p->no_xy * row_pitch(p, precision);
    const size_t xy = (size_t)p->no_xy * p->no_xy;
    const size_t gt = (size_t)ceil_div(p->no_xy, pic::kGT);
    const size_t b = (size_t)dens_slots(batch_steps(p));
    *bytes = 4 * (vox * elem + pic::kFieldSlack) + b * (xy * 8 + gt * gt * 8) + xy * 8 + 2 * partial_cap_bytes(p) + 256;
    return PIC_OK;
}

int pic_create(const pic_params *p, int precision, pic_sim **out) {
    if (!out) return fail(PIC_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = check_params(p, precision);
    if (rc) return rc;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PIC_E_NODEVICE, "no HIP device is visible");
    size_t need = 0;
    pic_mem_required(p, precision, &need);
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if ((double)need > 0.9 * (double)fr)
        return fail(PIC_E_NOMEM, "grid %lldx%lldx%lld needs %.3f GiB of device memory, %.3f GiB free",
                    (long long)p->no_xy, (long long)p->no_xy, (long long)p->nzb, need / 1073741824.0,
                    fr / 1073741824.0);
    pic_sim *s = new pic_sim();
    s->p = *p;
    s->prec = precision;
    s->elem = precision == PIC_FP64 ? 8 : 4;
    s->pitch = row_pitch(p, precision);
    s->n_vox = p->no_xy * p->no_xy * s->pitch;
    HIP_TRY(hipGetDevice(&s->device));
    s->cus = sm_count(s->device);
    auto cleanup = [&](int code) {
        pic_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(PIC_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&s->seg_go, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->seg_dens, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(PIC_E_HIP, "hipEventCreate failed"));
    void **bufs[4] = {&s->cur_p, &s->cur_n, &s->nxt_p, &s->nxt_n};
    for (void **b : bufs) {
        // kFieldSlack zero bytes past the last row: a four-value lane at the
        // end of a row whose length is 2 mod 4 loads 8 bytes past it
        const size_t bytes = s->n_vox * s->elem + pic::kFieldSlack;
        if (hipMalloc(b, bytes) != hipSuccess) return cleanup(fail(PIC_E_NOMEM, "hipMalloc of a carrier array failed"));
        if (hipMemsetAsync(*b, 0, bytes, s->stream) != hipSuccess) return cleanup(fail(PIC_E_HIP, "hipMemsetAsync failed"));
    }
    const size_t xy = (size_t)p->no_xy * p->no_xy;
    s->sweep_part_cap = (int64_t)(partial_cap_bytes(p) / (3 * sizeof(double)));
    s->batch = batch_steps(p);
    const int64_t slots = dens_slots(s->batch);
    if (hipMalloc(&s->dens, slots * xy * 8) != hipSuccess ||
        hipMalloc(&s->sweep_part, 2 * s->sweep_part_cap * 3 * 8) != hipSuccess ||
        hipMalloc(&s->dep_part, slots * gather_blocks(s) * 8) != hipSuccess ||
        hipMalloc(&s->scratch4, 8 * 8) != hipSuccess || hipMalloc(&s->zero_dens, xy * 8) != hipSuccess)
        return cleanup(fail(PIC_E_NOMEM, "hipMalloc of scratch failed"));
    // the two-step pass's partials, for its default chunk (a forced chunk
    // grows them in pic_set_steps_per_pass)
    if (!pass2_unsupported(s)) {
        const int r2 = ensure_pair_partials(s, pass2_shape(s).blocks());
        if (r2) return cleanup(r2);
    }
    if (hipMemsetAsync(s->dens, 0, slots * xy * 8, s->stream) != hipSuccess ||
        hipMemsetAsync(s->zero_dens, 0, xy * 8, s->stream) != hipSuccess ||
        hipMemsetAsync(s->scratch4, 0, 64, s->stream) != hipSuccess)
        return cleanup(fail(PIC_E_HIP, "hipMemsetAsync failed"));
    if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup(fail(PIC_E_HIP, "initialisation failed"));
    *out = s;
    return PIC_OK;
}

int pic_destroy(pic_sim *s) {
    if (!s) return PIC_OK;
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->gstream) hipStreamSynchronize(s->gstream);
    void *bufs[] = {s->cur_p, s->cur_n, s->nxt_p, s->nxt_n, s->dens, s->sweep_part, s->dep_part,
                    s->scratch4, s->zero_dens, s->per_step, s->xs, s->ys, s->offsets_d, s->pair_part};
    for (void *b : bufs)
        if (b) hipFree(b);
    free_bins(s);
    for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
    if (s->seg_go) hipEventDestroy(s->seg_go);
    if (s->seg_dens) hipEventDestroy(s->seg_dens);
    if (s->gstream) hipStreamDestroy(s->gstream);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return PIC_OK;
}

int pic_upload_tracks(pic_sim *s, const double *xs, const double *ys, int64_t n_tracks, const int64_t *offsets,
                      int64_t n_steps) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (n_tracks < 0 || n_steps < 0 || !offsets) return fail(PIC_E_INVALID, "bad track arrays");
    if (n_tracks > 0 && (!xs || !ys)) return fail(PIC_E_INVALID, "xs/ys NULL with n_tracks > 0");
    if (offsets[0] != 0 || offsets[n_steps] != n_tracks) return fail(PIC_E_INVALID, "offsets must run from 0 to n_tracks");
    for (int64_t i = 0; i < n_steps; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(PIC_E_INVALID, "offsets must be non-decreasing");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipStreamSynchronize(s->gstream));  // a side gather reads the track arrays
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    free_bins(s);
    if (s->xs) hipFree(s->xs);
    if (s->ys) hipFree(s->ys);
    if (s->per_step) hipFree(s->per_step);
    if (s->offsets_d) hipFree(s->offsets_d);
    s->xs = s->ys = s->per_step = nullptr;
    s->offsets_d = nullptr;
    s->n_steps = 0;
    const size_t tb = (size_t)std::max<int64_t>(n_tracks, 1) * 8;
    // checked before allocating, at 0.9 of free memory like the grid itself
    // (the full electrode's 24.6 M tracks take 394 MB)
    const size_t need = 2 * tb + (size_t)std::max<int64_t>(n_steps, 1) * 4 * 8 + (size_t)(n_steps + 1) * 8;
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if ((double)need > 0.9 * (double)fr)
        return fail(PIC_E_NOMEM, "%lld tracks over %lld steps need %.3f GiB of device memory, %.3f GiB free",
                    (long long)n_tracks, (long long)n_steps, need / 1073741824.0, fr / 1073741824.0);
    if (hipMalloc(&s->xs, tb) != hipSuccess || hipMalloc(&s->ys, tb) != hipSuccess ||
        hipMalloc(&s->per_step, (size_t)std::max<int64_t>(n_steps, 1) * 4 * 8) != hipSuccess ||
        hipMalloc(&s->offsets_d, (size_t)(n_steps + 1) * 8) != hipSuccess)
        return fail(PIC_E_NOMEM, "hipMalloc for %lld tracks failed", (long long)n_tracks);
    HIP_TRY(hipMemcpyAsync(s->offsets_d, offsets, (size_t)(n_steps + 1) * 8, hipMemcpyHostToDevice, s->stream));
    if (n_tracks > 0) {
        HIP_TRY(hipMemcpyAsync(s->xs, xs, n_tracks * 8, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->ys, ys, n_tracks * 8, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemsetAsync(s->per_step, 0, (size_t)std::max<int64_t>(n_steps, 1) * 4 * 8, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->offsets.assign(offsets, offsets + n_steps + 1);
    s->n_tracks = n_tracks;
    s->n_steps = n_steps;
    return ensure_bins(s);
}

int pic_set_gather_mode(pic_sim *s, int mode) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (mode != PIC_GATHER_AUTO && mode != PIC_GATHER_SCAN && mode != PIC_GATHER_BINNED)
        return fail(PIC_E_INVALID, "gather mode must be PIC_GATHER_AUTO, PIC_GATHER_SCAN or PIC_GATHER_BINNED");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipStreamSynchronize(s->gstream));
    const int old = s->gather_mode;
    s->gather_mode = mode;
    const int rc = ensure_bins(s);
    if (rc) {
        s->gather_mode = old;
        ensure_bins(s);
    }
    return rc;
}

int pic_get_gather_mode(pic_sim *s, int *binned) {
    if (!s || !binned) return fail(PIC_E_INVALID, "NULL argument");
    *binned = s->bins_ready ? 1 : 0;
    return PIC_OK;
}

}  // extern "C"

namespace {

// The tracks and per-step record a run works on: the handle's uploaded
// arrays (pic_run_steps) or the static scratch (pic_run_pulse_static).
struct TrackView {
    const double *xs = nullptr, *ys = nullptr;
    const int64_t *offsets_d = nullptr;  // device step offsets, n_steps + 1
    const int64_t *offsets_h = nullptr;  // the same on the host
    double *per_step = nullptr;          // device (n_steps, 4) record
    int64_t n_steps = 0;
    bool uploaded = false;               // the handle's upload, whose bins the handle holds
};

TrackView uploaded_tracks(const pic_sim *s) {
    TrackView t;
    t.xs = s->xs;
    t.ys = s->ys;
    t.offsets_d = s->offsets_d;
    t.offsets_h = s->offsets.data();
    t.per_step = s->per_step;
    t.n_steps = s->n_steps;
    t.uploaded = true;
    return t;
}

// Steps [first, first+count) over the tracks of `tv`. `direct`: launch one
// kernel at a time with no timing events and no graph replay (no host or
// device memory is allocated on this path).
int run_steps_on(pic_sim *s, const TrackView &tv, int64_t first, int64_t count, bool direct) {
    HIP_TRY(hipStreamSynchronize(s->stream));  // events of an earlier call may still be pending
    HIP_TRY(hipStreamSynchronize(s->gstream));
    s->ev_used = 0;
    s->group_open = 0;
    s->in_run = !direct;  // timing events are recorded only inside a run that asks for them
    int rc = PIC_OK;
    const int n_tiles = (int)gather_blocks(s);
    const int64_t xy = s->p.no_xy * s->p.no_xy;
    const int64_t slots = dens_slots(s->batch);
    const int64_t end = first + count;
    const int64_t *off = tv.offsets_h;
    // One segment per gather batch. Launched one by one, a segment's gather
    // runs on the side stream beside the previous segment's sweeps: it starts
    // once that segment's first sweep has folded the step before it, and the
    // segment's sweeps wait for it. The first segment of a call has nothing
    // to run beside, so it is kept to kLeadSteps steps: its short gather is
    // all the call waits for before its first sweep, and the next segment's
    // gather runs beside its sweeps. (5 um full electrode, 20 steps from the
    // start: one 20-step segment waited 0.93 ms for its gather.) With graph
    // replay a segment's gather and sweeps are collected, then replayed
    // through graph chains in order.
    constexpr int64_t kLeadSteps = 2;
    const bool chained = s->graphs && !s->timing && !direct;
    auto seg_end = [&](int64_t b0) {
        return std::min(end, b0 + (b0 == first ? std::min<int64_t>(kLeadSteps, s->batch) : s->batch));
    };
    // the gather of segment [g0, seg_end(g0)) on the side stream: 1 queued, 0 no tracks, else an error
    auto side_gather = [&](int64_t g0) -> int {
        const int64_t g1 = seg_end(g0);
        if (g0 >= end || off[g1] <= off[g0]) return 0;
        int r = enqueue_gather(s, tv.xs, tv.ys, tv.offsets_d, g0, g1 - g0, s->gstream, tv.uploaded ? off : nullptr);
        if (!r && hipEventRecord(s->seg_dens, s->gstream) != hipSuccess) r = fail(PIC_E_HIP, "hipEventRecord failed");
        return r ? r : 1;
    };
    const bool use_pairs = pass2_on(s);
    bool gather_queued = false;
    if (!chained) {
        const int g = side_gather(first);
        if (g == 1) gather_queued = true;
        else if (g) rc = g;
    }
    for (int64_t b0 = first, b1 = 0; b0 < end && rc == PIC_OK; b0 = b1) {
        b1 = seg_end(b0);
        if (chained) {
            s->seg.clear();
            s->sink = &s->seg;
            if (off[b1] > off[b0])
                rc = enqueue_gather(s, tv.xs, tv.ys, tv.offsets_d, b0, b1 - b0, s->stream, tv.uploaded ? off : nullptr);
        } else if (gather_queued) {
            if (hipStreamWaitEvent(s->stream, s->seg_dens, 0) != hipSuccess) rc = fail(PIC_E_HIP, "hipStreamWaitEvent failed");
            gather_queued = false;
        }
        for (int64_t step = b0, stride = 1; step < b1 && rc == PIC_OK; step += stride) {
            stride = use_pairs && step + 1 < b1 ? 2 : 1;
            if (stride == 2) {
                bool dep[2];
                const double *dens[2], *dpart[2];
                double *out4[2];
                for (int q = 0; q < 2; ++q) {
                    const int64_t t = step + q, slot = t % slots;
                    dep[q] = off[t + 1] > off[t];
                    dens[q] = s->dens + slot * xy;
                    dpart[q] = s->dep_part + slot * n_tiles;
                    out4[q] = tv.per_step + 4 * t;
                }
                rc = enqueue_sweep2(s, step, dep, dens, dpart, n_tiles, out4);
            } else {
                const bool dep = off[step + 1] > off[step];
                const int64_t slot = step % slots;
                rc = enqueue_sweep(s, step, dep, s->dens + slot * xy, s->dep_part + slot * n_tiles, n_tiles,
                                   tv.per_step + 4 * step);
            }
            if (!rc) swap_buffers(s);
            if (!rc && !chained && step == b0 && b1 < end) {
                if (hipEventRecord(s->seg_go, s->stream) != hipSuccess ||
                    hipStreamWaitEvent(s->gstream, s->seg_go, 0) != hipSuccess) {
                    rc = fail(PIC_E_HIP, "ordering the side stream failed");
                } else {
                    const int g = side_gather(b1);
                    if (g == 1) gather_queued = true;
                    else if (g) rc = g;
                }
            }
        }
        s->sink = nullptr;
        if (chained && !rc) rc = submit(s->seg, s->device, s->stream);
    }
    if (!rc) rc = flush_pending(s);
    close_group(s);
    s->in_run = false;
    // Whatever happened above, the handle's own stream ends behind every
    // gather queued on the side stream, so any later call that synchronises
    // only that stream also waits for them (a gather still writing density
    // slots the next call reuses would otherwise race with it).
    if (hipEventRecord(s->seg_dens, s->gstream) != hipSuccess || hipStreamWaitEvent(s->stream, s->seg_dens, 0) != hipSuccess) {
        if (!rc) rc = fail(PIC_E_HIP, "ordering the handle's stream after its side stream failed");
    }
    if (rc) {  // a failed run leaves no step to fold
        s->pending = pic_sim::Pending{};
        s->pair_open = false;
    }
    return rc;
}

}  // namespace

extern "C" {

int pic_run_steps(pic_sim *s, int64_t first, int64_t count) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (first < 0 || count < 0 || first + count > s->n_steps)
        return fail(PIC_E_INVALID, "steps [%lld, %lld) outside the uploaded %lld", (long long)first,
                    (long long)(first + count), (long long)s->n_steps);
    return run_steps_on(s, uploaded_tracks(s), first, count, false);
}

int pic_read_per_step(pic_sim *s, int64_t first, int64_t count, double *out) {
    if (!s || !out) return fail(PIC_E_INVALID, "NULL argument");
    if (first < 0 || count < 0 || first + count > s->n_steps) return fail(PIC_E_INVALID, "rows outside the record");
    if (count == 0) return PIC_OK;
    HIP_TRY(hipMemcpyAsync(out, s->per_step + 4 * first, count * 4 * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PIC_OK;
}

int pic_run_pulse(pic_sim *s, const double *xs, const double *ys, const int64_t *step_offsets, int64_t n_steps,
                  double *per_step_out) {
    if (!s || !step_offsets || !per_step_out) return fail(PIC_E_INVALID, "NULL argument");
    if (n_steps < 1) return fail(PIC_E_INVALID, "n_steps must be positive");
    int rc = pic_upload_tracks(s, xs, ys, step_offsets[n_steps], step_offsets, n_steps);
    if (!rc) rc = pic_run_steps(s, 0, n_steps);
    if (!rc) rc = pic_read_per_step(s, 0, n_steps, per_step_out);
    return rc;
}

int pic_run_pulse_static(pic_sim *s, const double *xs, const double *ys, const int64_t *step_offsets,
                         int64_t n_steps, double *per_step_out) {
    if (!s || !step_offsets || !per_step_out) return fail(PIC_E_INVALID, "NULL argument");
    if (n_steps < 1) return fail(PIC_E_INVALID, "n_steps must be positive");
    const int64_t n_tracks = step_offsets[n_steps];
    if (step_offsets[0] != 0 || n_tracks < 0) return fail(PIC_E_INVALID, "offsets must run from 0 to n_tracks");
    for (int64_t i = 0; i < n_steps; ++i)
        if (step_offsets[i + 1] < step_offsets[i]) return fail(PIC_E_INVALID, "offsets must be non-decreasing");
    if (n_tracks > 0 && (!xs || !ys)) return fail(PIC_E_INVALID, "xs/ys NULL with tracks in the schedule");
    // regions of the scratch, each 256-byte aligned: record, offsets, xs, ys
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t rec = up((size_t)n_steps * 32), offb = up((size_t)(n_steps + 1) * 8), trk = up((size_t)n_tracks * 8);
    const size_t need = rec + offb + 2 * trk;
    if (need > kStaticScratchBytes)
        return fail(PIC_E_INVALID, "%lld steps and %lld tracks need %.2f MiB of staging; the static scratch of "
                    "pic_run_pulse_static holds %zu MiB (pic_run_pulse has no such limit)", (long long)n_steps,
                    (long long)n_tracks, need / 1048576.0, kStaticScratchBytes >> 20);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != s->device) return fail(PIC_E_INVALID, "the handle lives on device %d, current is %d", s->device, dev);
    unsigned char *base = nullptr;
    int rc = static_scratch(dev, &base);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipStreamSynchronize(s->gstream));
    TrackView tv;
    tv.per_step = reinterpret_cast<double *>(base);
    int64_t *offsets_d = reinterpret_cast<int64_t *>(base + rec);
    double *xs_d = reinterpret_cast<double *>(base + rec + offb);
    double *ys_d = reinterpret_cast<double *>(base + rec + offb + trk);
    tv.offsets_d = offsets_d;
    tv.offsets_h = step_offsets;
    tv.xs = xs_d;
    tv.ys = ys_d;
    tv.n_steps = n_steps;
    HIP_TRY(hipMemcpyAsync(offsets_d, step_offsets, (size_t)(n_steps + 1) * 8, hipMemcpyHostToDevice, s->stream));
    if (n_tracks > 0) {
        HIP_TRY(hipMemcpyAsync(xs_d, xs, (size_t)n_tracks * 8, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(ys_d, ys, (size_t)n_tracks * 8, hipMemcpyHostToDevice, s->stream));
    }
    // Nothing here must start at zero: every row of the record is written
    // whole by the fold of its step. The gathers read the offsets and centres
    // only after the copies above (same stream; the side stream's first
    // gather is queued inside run_steps_on after it synchronises both).
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    rc = run_steps_on(s, tv, 0, n_steps, true);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(per_step_out, tv.per_step, (size_t)n_steps * 32, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PIC_OK;
}

int pic_deposit(pic_sim *s, const double *xs, const double *ys, int64_t m, double *inserted) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (m < 0 || (m > 0 && (!xs || !ys))) return fail(PIC_E_INVALID, "bad track arrays");
    double host4[4] = {0, 0, 0, 0};
    if (m > 0) {
        double *dx = nullptr, *dy = nullptr;
        int64_t *doff = nullptr;
        const int64_t off[2] = {0, m};
        if (hipMalloc(&dx, m * 8) != hipSuccess || hipMalloc(&dy, m * 8) != hipSuccess ||
            hipMalloc(&doff, sizeof off) != hipSuccess) {
            if (dx) hipFree(dx);
            if (dy) hipFree(dy);
            return fail(PIC_E_NOMEM, "hipMalloc for %lld tracks failed", (long long)m);
        }
        int rc = PIC_OK;
        if (hipMemcpyAsync(dx, xs, m * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(dy, ys, m * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(doff, off, sizeof off, hipMemcpyHostToDevice, s->stream) != hipSuccess)
            rc = fail(PIC_E_HIP, "track upload failed");
        if (!rc) rc = enqueue_gather(s, dx, dy, doff, 0, 1);
        if (!rc) {
            const int64_t n = s->p.no_xy * s->p.no_xy * s->p.no_z;
            const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192));
            const int no_xy = (int)s->p.no_xy, pitch = (int)s->pitch, nze = (int)s->p.nze, no_z = (int)s->p.no_z;
            if (s->prec == PIC_FP64)
                hipLaunchKernelGGL(pic::broadcast_kernel<double>, dim3(blocks), dim3(256), 0, s->stream,
                                   (double *)s->cur_p, (double *)s->cur_n, s->dens, no_xy, pitch, nze, no_z);
            else
                hipLaunchKernelGGL(pic::broadcast_kernel<float>, dim3(blocks), dim3(256), 0, s->stream,
                                   (float *)s->cur_p, (float *)s->cur_n, s->dens, no_xy, pitch, nze, no_z);
            if (hipGetLastError() != hipSuccess) rc = fail(PIC_E_HIP, "broadcast launch failed");
        }
        if (!rc) rc = enqueue_deposit_reduce(s, s->scratch4);
        if (!rc && (hipMemcpyAsync(host4, s->scratch4, 32, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                    hipStreamSynchronize(s->stream) != hipSuccess))
            rc = fail(PIC_E_HIP, "deposit readback failed");
        hipStreamSynchronize(s->stream);
        hipFree(dx);
        hipFree(dy);
        hipFree(doff);
        if (rc) return rc;
    }
    if (inserted) *inserted = host4[0];
    return PIC_OK;
}

int pic_step(pic_sim *s, double out3[3]) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    int rc;
    if ((rc = flush_pending(s))) return rc;
    if ((rc = enqueue_sweep(s, 0, false, nullptr, nullptr, 0, s->scratch4))) return rc;
    if ((rc = flush_pending(s))) return rc;
    swap_buffers(s);
    double host4[4];
    HIP_TRY(hipMemcpyAsync(host4, s->scratch4, 32, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (out3) {
        out3[0] = host4[1];
        out3[1] = host4[2];
        out3[2] = host4[3];
    }
    return PIC_OK;
}

int pic_read_fields(pic_sim *s, double *pos, double *neg) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (s->p.wall == PIC_WALL_REFLECTING) {
        // the sweep reads the reflecting ring through the interior; store it now
        const int64_t n = (4 * s->p.no_xy - 4) * s->pitch;
        const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096));
        if (s->prec == PIC_FP64)
            hipLaunchKernelGGL(pic::ring_mirror_kernel<double>, dim3(blocks), dim3(256), 0, s->stream,
                               (double *)s->cur_p, (double *)s->cur_n, (int)s->p.no_xy, (int)s->pitch);
        else
            hipLaunchKernelGGL(pic::ring_mirror_kernel<float>, dim3(blocks), dim3(256), 0, s->stream,
                               (float *)s->cur_p, (float *)s->cur_n, (int)s->p.no_xy, (int)s->pitch);
        HIP_LAUNCH_CHECK();
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    void *src[2] = {s->cur_p, s->cur_n};
    double *dst[2] = {pos, neg};
    std::vector<float> tmp;
    for (int q = 0; q < 2; ++q) {
        if (!dst[q]) continue;
        if (s->prec == PIC_FP64) {
            HIP_TRY(hipMemcpy(dst[q], src[q], s->n_vox * 8, hipMemcpyDeviceToHost));
        } else {
            tmp.resize(s->n_vox);
            HIP_TRY(hipMemcpy(tmp.data(), src[q], s->n_vox * 4, hipMemcpyDeviceToHost));
            const int64_t rows = s->p.no_xy * s->p.no_xy, nzb = s->p.nzb;
            for (int64_t r = 0; r < rows; ++r)
                for (int64_t k = 0; k < nzb; ++k) dst[q][r * nzb + k] = (double)tmp[r * s->pitch + k];
        }
    }
    return PIC_OK;
}

int pic_write_fields(pic_sim *s, const double *pos, const double *neg) {
    if (!s || !pos || !neg) return fail(PIC_E_INVALID, "NULL argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const double *src[2] = {pos, neg};
    void *dst[2][2] = {{s->cur_p, s->nxt_p}, {s->cur_n, s->nxt_n}};
    std::vector<float> tmp;
    for (int q = 0; q < 2; ++q) {
        for (int b = 0; b < 2; ++b) {
            if (s->prec == PIC_FP64) {
                HIP_TRY(hipMemcpy(dst[q][b], src[q], s->n_vox * 8, hipMemcpyHostToDevice));
            } else {
                tmp.assign(s->n_vox, 0.0f);  // row padding stays zero
                const int64_t rows = s->p.no_xy * s->p.no_xy, nzb = s->p.nzb;
                for (int64_t r = 0; r < rows; ++r)
                    for (int64_t k = 0; k < nzb; ++k) tmp[r * s->pitch + k] = (float)src[q][r * nzb + k];
                HIP_TRY(hipMemcpy(dst[q][b], tmp.data(), s->n_vox * 4, hipMemcpyHostToDevice));
            }
        }
    }
    return PIC_OK;
}

int pic_synchronize(pic_sim *s) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipStreamSynchronize(s->gstream));  // the side stream's gathers (a run waits for them itself)
    return PIC_OK;
}

int pic_set_tiling(pic_sim *s, int rows_per_tile, int planes_per_chunk, int values_per_lane) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (rows_per_tile < 0 || planes_per_chunk < 0) return fail(PIC_E_INVALID, "negative tiling");
    if (values_per_lane < 0 || values_per_lane > 4 || values_per_lane == 3)
        return fail(PIC_E_INVALID, "values_per_lane must be 0, 1, 2 or 4");
    if (values_per_lane == 4 && s->prec != PIC_FP32) return fail(PIC_E_INVALID, "4 values per lane needs FP32 fields");
    if (values_per_lane == 2 && s->prec == PIC_FP64 && rows_per_tile > 2)
        return fail(PIC_E_INVALID, "two FP64 values per lane take rows_per_tile 0, 1 or 2");
    if (values_per_lane == 4 && rows_per_tile > 3)
        return fail(PIC_E_INVALID, "four values per lane takes rows_per_tile 0, 1, 2 or 3");
    if (rows_per_tile && rows_per_tile != 1 && rows_per_tile != 2 && rows_per_tile != 3 && rows_per_tile != 4 &&
        rows_per_tile != 8)
        return fail(PIC_E_INVALID, "rows_per_tile must be 0, 1, 2, 3, 4 or 8");
    if (rows_per_tile == 3 && (s->prec != PIC_FP32 || (values_per_lane != 0 && values_per_lane != 4)))
        return fail(PIC_E_INVALID, "rows_per_tile 3 needs FP32 fields with four values per lane");
    if (rows_per_tile == 8 && (s->prec != PIC_FP32 || values_per_lane != 1))
        return fail(PIC_E_INVALID, "rows_per_tile 8 needs FP32 fields and values_per_lane 1");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int old_tj = s->force_tj, old_chunk = s->force_chunk, old_v = s->force_v;
    s->force_v = values_per_lane;
    s->force_tj = rows_per_tile;
    s->force_chunk = planes_per_chunk;
    s->shape_valid = false;
    const int rc = ensure_partials(s, sweep_shape(s).blocks());
    if (rc) {
        s->force_tj = old_tj;
        s->force_chunk = old_chunk;
        s->force_v = old_v;
        s->shape_valid = false;
    }
    return rc;
}

int pic_set_steps_per_pass(pic_sim *s, int steps, int planes_per_chunk) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (steps < 0 || steps > 2) return fail(PIC_E_INVALID, "steps per pass must be 0, 1 or 2");
    if (planes_per_chunk < 0) return fail(PIC_E_INVALID, "negative planes_per_chunk");
    if (steps == 2) {
        if (const char *why = pass2_unsupported(s)) return fail(PIC_E_INVALID, "%s", why);
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int old_steps = s->steps_per_pass, old_chunk = s->pass2_chunk;
    s->steps_per_pass = steps;
    if (planes_per_chunk > 0) s->pass2_chunk = planes_per_chunk;
    const int rc = pass2_unsupported(s) ? PIC_OK : ensure_pair_partials(s, pass2_shape(s).blocks());
    if (rc) {
        s->steps_per_pass = old_steps;
        s->pass2_chunk = old_chunk;
    }
    return rc;
}

int pic_set_tile_group(pic_sim *s, int tiles_per_block) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (tiles_per_block < 0 || tiles_per_block > 8) return fail(PIC_E_INVALID, "tiles_per_block must be 0..8");
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->force_tpb = tiles_per_block;
    s->shape_valid = false;
    return ensure_partials(s, sweep_shape(s).blocks());
}

int pic_get_tiling(pic_sim *s, int *rows_per_tile, int *planes_per_chunk, int *values_per_lane) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    const SweepShape sh = sweep_shape(s);
    if (rows_per_tile) *rows_per_tile = sh.tj;
    if (planes_per_chunk) *planes_per_chunk = sh.chunk;
    if (values_per_lane) *values_per_lane = sh.v;
    return PIC_OK;
}

int pic_enable_timing(pic_sim *s, int enable) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (enable < 0) return fail(PIC_E_INVALID, "enable must be 0, 1 or a number of sweep launches per event pair");
    s->timing = enable != 0;
    s->timing_group = enable > 1 ? enable : 1;
    return PIC_OK;
}

int pic_set_graphs(pic_sim *s, int enable) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    s->graphs = enable != 0;
    return PIC_OK;
}

int pic_get_timing(pic_sim *s, double *sweep_ms, int64_t *sweep_launches, double *deposit_ms, int64_t *deposit_launches,
                   double *other_ms) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    HIP_TRY(hipStreamSynchronize(s->stream));
    double t[3] = {0, 0, 0};
    int64_t n[3] = {0, 0, 0};
    for (size_t q = 0; q + 1 < s->ev_used; q += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_pool[q], s->ev_pool[q + 1]));
        const int kind = s->ev_kind[q];
        t[kind] += ms;
        n[kind] += s->ev_count[q];
    }
    if (sweep_ms) *sweep_ms = t[0];
    if (sweep_launches) *sweep_launches = n[0];
    if (deposit_ms) *deposit_ms = t[1];
    if (deposit_launches) *deposit_launches = n[1];
    if (other_ms) *other_ms = t[2];
    return PIC_OK;
}

int pic_reset(pic_sim *s) {
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    s->pending = pic_sim::Pending{};
    s->pair_open = false;
    void *bufs[4] = {s->cur_p, s->cur_n, s->nxt_p, s->nxt_n};
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(s->n_vox, 256), 16384));
    for (void *b : bufs) {
        if (s->prec == PIC_FP64)
            hipLaunchKernelGGL(pic::fill_kernel<double>, dim3(blocks), dim3(256), 0, s->stream, (double *)b, s->n_vox);
        else
            hipLaunchKernelGGL(pic::fill_kernel<float>, dim3(blocks), dim3(256), 0, s->stream, (float *)b, s->n_vox);
        HIP_LAUNCH_CHECK();
    }
    if (s->per_step) {
        const int64_t n = std::max<int64_t>(s->n_steps, 1) * 4;
        hipLaunchKernelGGL(pic::fill_kernel<double>, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096))),
                           dim3(256), 0, s->stream, s->per_step, n);
        HIP_LAUNCH_CHECK();
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PIC_OK;
}

int pic_batch_create(pic_sim *const *sims, int n, pic_batch **out) {
    if (!out) return fail(PIC_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!sims || n < 1) return fail(PIC_E_INVALID, "a batch needs at least one handle");
    for (int c = 0; c < n; ++c) {
        if (!sims[c]) return fail(PIC_E_INVALID, "handle %d is NULL", c);
        for (int d = 0; d < c; ++d)
            if (sims[d] == sims[c]) return fail(PIC_E_INVALID, "handle %d appears twice", c);
        if (sims[c]->prec != sims[0]->prec) return fail(PIC_E_INVALID, "handles of one batch share a precision");
        if (sims[c]->p.wall != sims[0]->p.wall) return fail(PIC_E_INVALID, "handles of one batch share a wall type");
        if (sims[c]->device != sims[0]->device) return fail(PIC_E_INVALID, "handles of one batch share a device");
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != sims[0]->device) return fail(PIC_E_INVALID, "the handles live on device %d, current is %d", sims[0]->device, dev);
    pic_batch *b = new pic_batch();
    b->sims.assign(sims, sims + n);
    b->gdesc_build.resize(n);
    b->prec = sims[0]->prec;
    b->reflect = sims[0]->p.wall == PIC_WALL_REFLECTING;
    auto cleanup = [&](int code) {
        pic_batch_destroy(b);
        return code;
    };
    int rc = batch_tiling(b);
    if (rc) return cleanup(rc);
    int64_t xy_max = 0;
    for (int c = 0; c < n; ++c) xy_max = std::max(xy_max, sims[c]->p.no_xy * sims[c]->p.no_xy);
    if (hipMalloc(&b->zero_dens, (size_t)xy_max * 8) != hipSuccess ||
        hipMemset(b->zero_dens, 0, (size_t)xy_max * 8) != hipSuccess)
        return cleanup(fail(PIC_E_NOMEM, "allocation of the batch's zero density failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->slot_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->slot_ev[1], hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(PIC_E_HIP, "stream or event creation failed"));
    *out = b;
    return PIC_OK;
}

int pic_batch_destroy(pic_batch *b) {
    if (!b) return PIC_OK;
    for (pic_batch *&l : b->lanes) {
        pic_batch_destroy(l);
        l = nullptr;
    }
    if (b->stream) hipStreamSynchronize(b->stream);
    free_batch_slots(b);
    if (b->zero_dens) hipFree(b->zero_dens);
    for (hipEvent_t e : b->slot_ev)
        if (e) hipEventDestroy(e);
    if (b->done) hipEventDestroy(b->done);
    for (hipEvent_t e : b->ev_pool) hipEventDestroy(e);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return PIC_OK;
}

int pic_batch_set_tiling(pic_batch *b, int rows_per_tile, int planes_per_chunk) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    if (rows_per_tile != 0 && rows_per_tile != 1 && rows_per_tile != 2 && rows_per_tile != 4)
        return fail(PIC_E_INVALID, "batched rows_per_tile must be 0, 1, 2 or 4");
    if (rows_per_tile == 4 && b->prec == PIC_FP64) return fail(PIC_E_INVALID, "batched FP64 sweeps take 1 or 2 rows per tile");
    if (planes_per_chunk < 0) return fail(PIC_E_INVALID, "negative planes_per_chunk");
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (pic_batch *&l : b->lanes) {  // rebuilt with the new tiling at the next split run
        pic_batch_destroy(l);
        l = nullptr;
    }
    b->lane_of.clear();
    b->n_lanes = 0;
    const int old_tj = b->force_tj, old_chunk = b->force_chunk;
    b->force_tj = rows_per_tile;
    b->force_chunk = planes_per_chunk;
    int rc = batch_tiling(b);
    if (rc) {
        b->force_tj = old_tj;
        b->force_chunk = old_chunk;
        batch_tiling(b);
    }
    return rc;
}

int pic_batch_get_tiling(pic_batch *b, int *rows_per_tile, int *planes_per_chunk, int *values_per_lane,
                         int *windows_per_block) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    if (rows_per_tile) *rows_per_tile = b->cands.empty() ? 0 : b->cands[0].first;
    if (planes_per_chunk) *planes_per_chunk = b->cands.empty() ? 0 : b->cands[0].second;
    if (values_per_lane) *values_per_lane = b->v;
    if (windows_per_block) *windows_per_block = b->wpb;
    return PIC_OK;
}

int pic_batch_run(pic_batch *b, int64_t first, int64_t count) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    const int n = (int)b->sims.size();
    int64_t longest = 0;
    for (pic_sim *s : b->sims) longest = std::max(longest, s->n_steps);
    if (first < 0 || count < 0 || first + count > longest)
        return fail(PIC_E_INVALID, "steps [%lld, %lld) outside the longest uploaded run (%lld)", (long long)first,
                    (long long)(first + count), (long long)longest);
    // lanes drive launches from two host threads: not while timing
    if (!b->is_lane && !b->timing) {
        std::vector<int> part;
        const int n_lanes = lane_split(b, part);
        if (n_lanes > 1) return run_lanes(b, part, n_lanes, first, count);
    }
    // run order: longest first, so the configurations still running at any
    // step are a prefix and the grid covers exactly their blocks
    std::vector<int> ord(n);
    for (int c = 0; c < n; ++c) ord[c] = c;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int x, int y) { return b->sims[x]->n_steps > b->sims[y]->n_steps; });
    int64_t gbatch = 128;
    for (pic_sim *s : b->sims) gbatch = std::min(gbatch, s->batch);
    // every handle's queued work is done and nothing still reads the tables
    for (pic_sim *s : b->sims) HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->slot_busy[0] = b->slot_busy[1] = false;
    int rc = ensure_batch_slots(b, gbatch);
    if (rc) return rc;
    for (pic_sim *s : b->sims) {  // earlier runs were flushed
        s->pending = pic_sim::Pending{};
        s->pair_open = false;
    }
    b->ev_used = 0;
    const dim3 block((unsigned)(64 * b->wpb));
    std::vector<int> active, cand_of;
    int k = 0;
    // one segment per gather batch: gathers first, then the steps' sweeps
    for (int64_t s0 = first; s0 < first + count && rc == PIC_OK; s0 += gbatch, k ^= 1) {
        const int64_t s1 = std::min(first + count, s0 + gbatch);
        // built in ordinary host memory, copied into the pinned slot in one pass
        b->descs_build.resize((size_t)(s1 - s0) * n);
        b->offs_build.resize((size_t)(s1 - s0) * (n + 1));
        pic::SweepArgs *descs = b->descs_build.data();
        int *offs = b->offs_build.data();
        active.assign(s1 - s0, 0);
        cand_of.assign(s1 - s0, 0);
        // descriptors of every (step, configuration) of the segment, built in
        // step order so each handle's buffer swaps and pending steps advance
        for (int64_t t = s0; t < s1 && rc == PIC_OK; ++t) {
            int na = 0;
            while (na < n && t < b->sims[ord[na]]->n_steps) ++na;
            size_t ci = 0;  // first candidate tiling that keeps the device busy
            for (; ci + 1 < b->cands.size(); ++ci) {
                int64_t w = 0;
                for (int q = 0; q < na; ++q) w += b->shapes[ci][ord[q]].blocks() * b->wpb;
                if (w >= b->waves_min) break;
            }
            cand_of[t - s0] = (int)ci;
            int *o = offs + (t - s0) * (n + 1);
            o[0] = 0;
            for (int q = 0; q < na && rc == PIC_OK; ++q) {
                pic_sim *s = b->sims[ord[q]];
                const SweepShape &sh = b->shapes[ci][ord[q]];
                const int n_tiles = (int)gather_blocks(s);
                const int64_t xy = s->p.no_xy * s->p.no_xy;
                const int64_t slot = t % dens_slots(s->batch);
                const bool dep = s->offsets[t + 1] > s->offsets[t];
                rc = make_sweep_args(s, sh, t, dep, s->dens + slot * xy, s->dep_part + slot * n_tiles, n_tiles,
                                     s->per_step + 4 * t, descs[(t - s0) * n + q]);
                if (!dep) descs[(t - s0) * n + q].dens = b->zero_dens;  // the batched kernel always deposits
                swap_buffers(s);
                const int64_t nxt = (int64_t)o[q] + sh.blocks();
                if (nxt > INT32_MAX / 2) rc = fail(PIC_E_INVALID, "batch grid too large");
                o[q + 1] = (int)nxt;
            }
            active[t - s0] = na;
        }
        if (rc) break;
        if (b->slot_busy[k]) HIP_TRY(hipEventSynchronize(b->slot_ev[k]));  // host slot k copied out
        const size_t dbytes = (size_t)((s1 - s0) * n) * sizeof(pic::SweepArgs);
        const size_t obytes = (size_t)((s1 - s0) * (n + 1)) * sizeof(int);
        memcpy(b->descs_h[k], descs, dbytes);
        memcpy(b->offs_h[k], offs, obytes);
        HIP_TRY(hipMemcpyAsync(b->descs_d[k], b->descs_h[k], dbytes, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->offs_d[k], b->offs_h[k], obytes, hipMemcpyHostToDevice, b->stream));
        // one launch gathers the segment's steps of every configuration with tracks
        int ng = 0;
        std::vector<int> &goffs = b->goffs_build;
        goffs.assign(n + 1, 0);
        for (int q = 0; q < n; ++q) {
            pic_sim *s = b->sims[ord[q]];
            const int64_t e = std::min(s1, s->n_steps);
            if (s0 >= e) break;
            if (s->offsets[e] > s->offsets[s0]) {
                b->gdesc_build[ng] = make_gather_args(s, s->xs, s->ys, s->offsets_d, s0);
                const int64_t nxt = goffs[ng] + gather_blocks(s) * (e - s0);
                if (nxt > INT32_MAX / 2) return fail(PIC_E_INVALID, "batch gather grid too large");
                goffs[ng + 1] = (int)nxt;
                ++ng;
            }
        }
        if (ng > 0) {
            memcpy(b->gdesc_h[k], b->gdesc_build.data(), ng * sizeof(pic::GatherArgs));
            memcpy(b->goffs_h[k], goffs.data(), (ng + 1) * sizeof(int));
            HIP_TRY(hipMemcpyAsync(b->gdesc_d[k], b->gdesc_h[k], ng * sizeof(pic::GatherArgs), hipMemcpyHostToDevice,
                                   b->stream));
            HIP_TRY(hipMemcpyAsync(b->goffs_d[k], b->goffs_h[k], (ng + 1) * sizeof(int), hipMemcpyHostToDevice,
                                   b->stream));
        }
        HIP_TRY(hipEventRecord(b->slot_ev[k], b->stream));
        b->slot_busy[k] = true;
        // the segment's launches: the gather, then one sweep per step
        b->seg.clear();
        if (ng > 0)
            b->seg.push_back(KLaunch::make((const void *)pic::batch_gather_kernel, dim3((unsigned)goffs[ng]), dim3(256), 1,
                                           batch_table(b->gdesc_d[k], b->goffs_d[k], ng)));
        for (int64_t t = s0; t < s1 && rc == PIC_OK; ++t) {
            const int na = active[t - s0];
            const int *o = offs + (t - s0) * (n + 1);
            const dim3 grid((unsigned)(o[na] + na));  // na fold blocks, then the sweep blocks
            const auto table = batch_table(b->descs_d[k] + (t - s0) * n, b->offs_d[k] + (t - s0) * (n + 1), na);
            const int tj = b->cands[cand_of[t - s0]].first;
            const void *f = batch_fn(b->prec, b->v, tj, b->reflect, b->cands[cand_of[t - s0]].second);
            if (!f) rc = fail(PIC_E_INVALID, "no batched sweep kernel for %d rows per tile", tj);
            else b->seg.push_back(KLaunch::make(f, grid, block, 0, table));
        }
        if (rc) break;
        if (b->graphs && !b->timing) {
            rc = submit(b->seg, b->sims[0]->device, b->stream);
        } else {
            for (const KLaunch &l : b->seg) {
                record_batch(b, l.kind);
                if ((rc = launch_now(l, b->stream))) break;
                record_batch(b, l.kind);
            }
        }
    }
    for (pic_sim *s : b->sims)
        if (!rc) rc = flush_pending(s, b->stream);
    // later calls on a handle's own stream wait for the batch
    if (hipEventRecord(b->done, b->stream) != hipSuccess) return rc ? rc : fail(PIC_E_HIP, "hipEventRecord failed");
    for (pic_sim *s : b->sims)
        if (hipStreamWaitEvent(s->stream, b->done, 0) != hipSuccess && !rc) rc = fail(PIC_E_HIP, "hipStreamWaitEvent failed");
    return rc;
}

int pic_batch_enable_timing(pic_batch *b, int enable) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    b->timing = enable != 0;
    return PIC_OK;
}

int pic_batch_set_graphs(pic_batch *b, int enable) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    b->graphs = enable != 0;
    for (pic_batch *l : b->lanes)
        if (l) l->graphs = b->graphs;
    return PIC_OK;
}

int pic_batch_get_timing(pic_batch *b, double *sweep_ms, int64_t *sweep_launches, double *gather_ms,
                         int64_t *gather_launches) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    HIP_TRY(hipStreamSynchronize(b->stream));
    double t[2] = {0, 0};
    int64_t n[2] = {0, 0};
    for (size_t q = 0; q + 1 < b->ev_used; q += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev_pool[q], b->ev_pool[q + 1]));
        t[b->ev_kind[q]] += ms;
        ++n[b->ev_kind[q]];
    }
    if (sweep_ms) *sweep_ms = t[0];
    if (sweep_launches) *sweep_launches = n[0];
    if (gather_ms) *gather_ms = t[1];
    if (gather_launches) *gather_launches = n[1];
    return PIC_OK;
}

int pic_batch_synchronize(pic_batch *b) {
    if (!b) return fail(PIC_E_INVALID, "batch is NULL");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return PIC_OK;
}

int pic_adjoint_create(pic_sim *s, int64_t checkpoint_every, pic_adjoint **out) {
    return pic_adjoint_create_levels(s, checkpoint_every, 0, out);
}

int pic_adjoint_create_levels(pic_sim *s, int64_t checkpoint_every, int64_t segments_per_outer, pic_adjoint **out) {
    if (!out) return fail(PIC_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!s) return fail(PIC_E_INVALID, "sim is NULL");
    if (s->prec != PIC_FP64) return fail(PIC_E_INVALID, "the adjoint needs FP64 fields");
    if (s->p.no_xy < 4) return fail(PIC_E_INVALID, "the adjoint needs no_xy >= 4");
    if (s->n_steps <= 0) return fail(PIC_E_INVALID, "upload the tracks before creating the adjoint");
    if (checkpoint_every < 1) return fail(PIC_E_INVALID, "checkpoint_every must be positive");
    if (segments_per_outer < 0) return fail(PIC_E_INVALID, "segments_per_outer must be 0 (automatic) or positive");
    if (s->n_vox * 8 >= INT32_MAX)  // the adjoint kernel addresses each array through a 32-bit buffer offset
        return fail(PIC_E_INVALID, "the adjoint needs arrays below 2 GB (this grid has %lld voxels)", (long long)s->n_vox);
    pic_adjoint *ad = new pic_adjoint();
    ad->s = s;
    // a segment's densities must stay in the handle's gather slots until its
    // adjoint steps are done
    ad->K = std::min<int64_t>(checkpoint_every, s->batch);
    ad->T = s->n_steps;
    ad->nseg = ceil_div(ad->T, ad->K);
    // Two levels when there are many segments: the forward keeps every M-th
    // segment's starting state, and the reverse pass recomputes one outer
    // span of M segments at a time to recover the others -- one more forward
    // pass in all, for about 2 sqrt(nseg) states held instead of nseg
    // (3 pulses of 48^2 x 206 at K = 64: 1659 segments, M = 41).
    if (segments_per_outer == 0)
        ad->M = ad->nseg > 64 ? (int64_t)std::ceil(std::sqrt((double)ad->nseg)) : 1;
    else
        ad->M = std::min(segments_per_outer, ad->nseg);
    ad->n_outer = ceil_div(ad->nseg, ad->M);
    ad->chunk = adjoint_chunk(s);
    const int64_t xy = s->p.no_xy * s->p.no_xy, tiles = gather_blocks(s), nblk = adjoint_blocks(ad);
    const size_t pair = 2 * (size_t)s->n_vox * 8;
    // the segment store is swept in place: kFieldSlack past its last array, as
    // past every carrier array (a 16-byte lane may read one value past a row)
    const size_t sizes[] = {ad->n_outer * pair, (size_t)(ad->K + 1) * pair + pic::kFieldSlack, 2 * pair, (size_t)ad->K * xy * 8,
                            (size_t)ad->K * xy * 8, (size_t)ad->K * tiles * 2 * 8,
                            (size_t)nblk * 2 * pic::kAdjNV * 8, (size_t)pic::kAdjNV * 8,
                            (size_t)ad->T * 2 * 8, 64, (size_t)xy * 8, ad->M > 1 ? (size_t)ad->M * pair : 0};
    double **ptrs[] = {&ad->ckpt, &ad->store, &ad->mu, &ad->dgf, &ad->dh, &ad->dpart,
                       &ad->acc, &ad->grad, &ad->dinj_rec, &ad->scratch, &ad->zeros, &ad->inner};
    size_t total = 0;
    for (size_t b : sizes) total += b;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        delete ad;
        return fail(PIC_E_HIP, "hipMemGetInfo failed");
    }
    if ((double)total > 0.9 * (double)fr) {
        delete ad;
        return fail(PIC_E_NOMEM, "the adjoint needs %.3f GiB (checkpoints every %lld steps), %.3f GiB free",
                    total / 1073741824.0, (long long)ad->K, fr / 1073741824.0);
    }
    for (size_t q = 0; q < sizeof sizes / sizeof sizes[0]; ++q) {
        if (sizes[q] == 0) continue;
        if (hipMalloc(ptrs[q], sizes[q]) != hipSuccess) {
            pic_adjoint_destroy(ad);
            return fail(PIC_E_NOMEM, "hipMalloc for the adjoint failed");
        }
        if (hipMemset(*ptrs[q], 0, sizes[q]) != hipSuccess) {
            pic_adjoint_destroy(ad);
            return fail(PIC_E_HIP, "hipMemset for the adjoint failed");
        }
    }
    for (hipEvent_t &e : ad->ev)
        if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) {
            pic_adjoint_destroy(ad);
            return fail(PIC_E_HIP, "hipEventCreate failed");
        }
    ad->bytes = total;
    *out = ad;
    return PIC_OK;
}

int pic_adjoint_destroy(pic_adjoint *ad) {
    if (!ad) return PIC_OK;
    if (ad->s && ad->s->stream) hipStreamSynchronize(ad->s->stream);
    double *bufs[] = {ad->ckpt, ad->store, ad->mu, ad->dgf, ad->dh, ad->dpart, ad->acc, ad->grad,
                      ad->dinj_rec, ad->scratch, ad->zeros, ad->inner};
    for (double *b : bufs)
        if (b) hipFree(b);
    for (hipEvent_t e : ad->ev)
        if (e) hipEventDestroy(e);
    delete ad;
    return PIC_OK;
}

int pic_adjoint_run(pic_adjoint *ad, double sums[2], double d_rec[11], double d_inj[2]) {
    if (!ad) return fail(PIC_E_INVALID, "adjoint is NULL");
    pic_sim *s = ad->s;
    if (s->n_steps != ad->T) return fail(PIC_E_INVALID, "the handle's tracks changed since the adjoint was created");
    const int64_t T = ad->T, K = ad->K, nv = s->n_vox;
    int rc = pic_reset(s);
    if (rc) return rc;
    // forward, a checkpoint at the start of every M-th segment
    auto save_state = [&](double *c) -> int {
        HIP_TRY(hipMemcpyAsync(c, s->cur_p, nv * 8, hipMemcpyDeviceToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(c + nv, s->cur_n, nv * 8, hipMemcpyDeviceToDevice, s->stream));
        return PIC_OK;
    };
    HIP_TRY(hipEventRecord(ad->ev[0], s->stream));
    for (int64_t k = 0; k < ad->nseg && rc == PIC_OK; ++k) {
        if (k % ad->M == 0) rc = save_state(ad->ckpt + (k / ad->M) * 2 * nv);
        if (!rc) rc = pic_run_steps(s, k * K, std::min(K, T - k * K));
    }
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ad->ev[1], s->stream));
    // reverse sweep
    hipLaunchKernelGGL(pic::fill_kernel<double>, dim3(1024), dim3(256), 0, s->stream, ad->mu, 4 * nv);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(pic::fill_kernel<double>, dim3(256), dim3(256), 0, s->stream, ad->dinj_rec, 2 * T);
    HIP_LAUNCH_CHECK();
    const int64_t nblk = adjoint_blocks(ad);
    hipLaunchKernelGGL(pic::fill_kernel<double>, dim3(256), dim3(256), 0, s->stream, ad->acc,
                       nblk * 2 * pic::kAdjNV);
    HIP_LAUNCH_CHECK();
    int cur = 0;
    for (int64_t k = ad->nseg - 1; k >= 0 && rc == PIC_OK; --k) {
        const int64_t s0 = k * K, e = std::min(T, s0 + K);
        const double *start = ad->ckpt + k * 2 * nv;
        if (ad->M > 1) {
            // entering an outer span from its end: rerun it forward from its
            // checkpoint, keeping each segment's starting state (the same
            // launches as the forward pass, so the same states bit for bit)
            const int64_t o = k / ad->M, k0 = o * ad->M;
            if (k == std::min(ad->nseg, k0 + ad->M) - 1) {
                HIP_TRY(hipMemcpyAsync(s->cur_p, ad->ckpt + o * 2 * nv, nv * 8, hipMemcpyDeviceToDevice, s->stream));
                HIP_TRY(hipMemcpyAsync(s->cur_n, ad->ckpt + o * 2 * nv + nv, nv * 8, hipMemcpyDeviceToDevice,
                                       s->stream));
                for (int64_t q = k0; q <= k && rc == PIC_OK; ++q) {
                    rc = save_state(ad->inner + (q - k0) * 2 * nv);
                    if (!rc && q < k) rc = pic_run_steps(s, q * K, std::min(K, T - q * K));
                }
                if (rc) break;
            }
            start = ad->inner + (k - k0) * 2 * nv;
        }
        HIP_TRY(hipMemcpyAsync(ad->store, start, 2 * nv * 8, hipMemcpyDeviceToDevice, s->stream));
        if (s->offsets[e] > s->offsets[s0]) {
            rc = enqueue_gather(s, s->xs, s->ys, s->offsets_d, s0, e - s0, s->stream, s->offsets.data());
            if (!rc) rc = adjoint_deriv_gather(ad, s0, e - s0);
        }
        for (int64_t t = s0; t + 1 < e && rc == PIC_OK; ++t) rc = adjoint_recompute_sweep(ad, t, t - s0);
        for (int64_t t = e - 1; t >= s0 && rc == PIC_OK; --t) {
            rc = adjoint_step(ad, t, t - s0, cur);
            cur ^= 1;
        }
    }
    if (rc) return rc;
    hipLaunchKernelGGL(pic::fold_accumulators_kernel<pic::kAdjNV>, dim3(1), dim3(64), 0, s->stream, ad->acc,
                       (int)nblk, ad->grad);
    HIP_LAUNCH_CHECK();
    HIP_TRY(hipEventRecord(ad->ev[2], s->stream));
    // per-step records to the host, summed there in step order
    std::vector<double> per(T * 4), gr(pic::kAdjNV), di(T * 2);
    HIP_TRY(hipMemcpyAsync(per.data(), s->per_step, T * 4 * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(gr.data(), ad->grad, pic::kAdjNV * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(di.data(), ad->dinj_rec, T * 2 * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipEventElapsedTime(&ad->forward_ms, ad->ev[0], ad->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ad->backward_ms, ad->ev[1], ad->ev[2]));
    double si = 0.0, sr = 0.0;
    for (int64_t t = 0; t < T; ++t) {
        si += per[4 * t + 0];
        sr += per[4 * t + 1];
    }
    if (sums) {
        sums[0] = si;
        sums[1] = sr;
    }
    if (d_rec)
        for (int c = 0; c < pic::kAdjNV; ++c) d_rec[c] = gr[c];
    if (d_inj)
        for (int c = 0; c < 2; ++c) {
            double acc = 0.0;
            for (int64_t t = 0; t < T; ++t) acc += di[2 * t + c];
            d_inj[c] = acc;
        }
    return PIC_OK;
}

int pic_adjoint_get_info(pic_adjoint *ad, int64_t *checkpoint_every, size_t *device_bytes, double *forward_ms,
                         double *backward_ms) {
    if (!ad) return fail(PIC_E_INVALID, "adjoint is NULL");
    if (checkpoint_every) *checkpoint_every = ad->K;
    if (device_bytes) *device_bytes = ad->bytes;
    if (forward_ms) *forward_ms = ad->forward_ms;
    if (backward_ms) *backward_ms = ad->backward_ms;
    return PIC_OK;
}

int pic_adjoint_get_levels(pic_adjoint *ad, int64_t *segments_per_outer, int64_t *outer_checkpoints,
                           int64_t *segments) {
    if (!ad) return fail(PIC_E_INVALID, "adjoint is NULL");
    if (segments_per_outer) *segments_per_outer = ad->M;
    if (outer_checkpoints) *outer_checkpoints = ad->n_outer;
    if (segments) *segments = ad->nseg;
    return PIC_OK;
}

}  // extern "C"
