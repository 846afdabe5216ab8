// ekg_engine.hpp -- device-resident state behind the C ABI (include/ekg.h).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ekg.h"
#include "ekg_kernels.hpp"

namespace ekg {

// Thrown inside the library, turned into an error code at the ABI.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);

// Grow-only device buffer.  A buffer may instead point into memory it does
// not own (adopt): release() then forgets it without freeing.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    void reserve(size_t n);
    void adopt(void* q, size_t n);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Device workspace of one launch group on one stream (grow-only buffers).
struct Workspace {
    DevBuf hearts, reqs, starts, layer_k, lead_pos, tgrid, thl, lc, xy, G, aux, sums, bnd, pre_rng,
        pre_ab, partial, csums, ecg;
    std::vector<double> h_k, h_pos, h_tgrid;
    std::vector<float> h_thl;
    std::vector<char> h_tables;    // hearts + reqs + starts as last uploaded
    std::vector<char> h_grids;     // tgrid + thl as last uploaded
    // captured launch sequence of the last group (graph mode)
    hipGraphExec_t graph = nullptr;
    std::vector<char> graph_key;
    // second stream of a group's launches (the chunk, prefix and reduction
    // kernels beside the class-window kernel) and its fork and join events;
    // created by ensure_side, on the device that is current
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void ensure_side();
    // Fixed workspace: when `arena` is set, every buffer of a group is carved
    // from it in order (no device allocation); a group that does not fit
    // fails with EKG_ENOMEM.  Contents are not kept between groups.
    char* arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
    // reserve() for an owning workspace, a piece of the arena otherwise
    void need(DevBuf& b, size_t n);
    ~Workspace();
};
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Workspace ws;
    Workspace ws_fixed;  // the allocation-free entry point's view of the device's static workspace
    bool profiling = false;
    int kernel = EKG_KERNEL_AUTO;  // contraction path of the fused kernel
    bool graphs = false;           // replay repeated launch groups from a captured hipGraph
    int64_t graph_launches = 0, graph_captures = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;  // fused launches not yet timed
    std::vector<hipEvent_t> free_events;
    int64_t fused_launches = 0;
    double fused_ms = 0.0;
    DevBuf nb_table;  // [4][26][4] int8 {d0, d1, d2, cube bit} of the four dipole neighbourhoods
    DevBuf crit_tg, crit_off, crit_stats, crit_out;  // criteria: targets, offsets, target statistics, output
    std::vector<double> crit_h;          // targets, offsets and comparison length as last prepared
    std::mutex mu;    // serialises ABI calls on this context
    void bind() const;
    hipEvent_t get_event();
    void collect_timings();
};

// Per-heart layout for the ECG kernels: myocardial voxels sorted by
// (layer, activation time, grid index), cut into layer-pure chunks.  With
// activation classes on, each layer lists first its voxels outside classes
// (in that order, cut into chunks) and then its classes, each a run of
// voxels with one activation time of at least class_min voxels, cut into
// class chunks.
struct EcgLayout {
    bool valid = false;
    int64_t V = 0;
    int n_chunks = 0;
    int n_bricks = 0;
    int max_layer = 0;
    int chunk_target = 0;          // voxels per chunk the layout was cut with
    int class_min = -1;            // smallest class it was built with (0: no classes)
    int n_cchunks = 0, n_classes = 0;
    std::vector<int4> h_chunks, h_cchunks, h_classes;
    std::vector<double> h_class_a;
    std::vector<uint2> h_geo;      // per voxel of the layout: packed grid coordinates, 3x3x3 myocardium mask
    // Per dipole neighbourhood (EKG_NHOOD_*), built on first use: the class
    // windows' entries as lane-major groups of 16 and their pieces (see
    // ensure_cells).  n_ccells counts (class, cell) pairs with a non-zero
    // dipole vector, n_centries (window, cell) pairs.
    bool cells_built[4] = {false, false, false, false};
    int64_t n_ccells[4] = {0, 0, 0, 0}, n_centries[4] = {0, 0, 0, 0}, n_cgroups[4] = {0, 0, 0, 0};
    int n_cpieces[4] = {0, 0, 0, 0};
    DevBuf cgrp[4], crow[4], cpieces[4], cwin[4];
    std::vector<int> h_layer_chunk;
    std::vector<double> h_layer_t0;
    DevBuf vox_chunk, vox_a64, vox_a, vox_geo, chunks, bricks, brick_off, brick_vox, halo_off, halo_cells,
        layer_chunk, layer_t0, cchunks, classes, class_a;
    ekgk::HeartDev dev{};          // device pointers and sizes for the kernels
};
struct Heart {
    Ctx* ctx = nullptr;
    int device = 0;  // ctx->device, kept so that the heart can be destroyed after its context
    int n0 = 0, n1 = 0, n2 = 0;
    int64_t nvox = 0, n_myo = 0;
    int max_layer = 0;
    std::vector<uint8_t> h_layer, h_seed;
    DevBuf d_layer, d_seed, d_act;
    std::vector<double> h_act;
    bool excited = false;
    EcgLayout layout;
    // tuning chosen by ekg_heart_autotune (0 = defaults)
    int tune_chunk = 0;   // target voxels per chunk
    int tune_waves = 0;   // waves per block of the matrix-core fused kernel
    int tune_phase = 0;   // leads per lead-field phase (2 or 4)
    int tune_kernel = 0;  // EKG_KERNEL_* override
    int tune_class = 0;   // smallest activation class (voxels) of the matrix-core path; 0 default, -1 none
    std::mutex mu;
};

// One validated request of a launch group.
struct ReqSpec {
    Heart* h = nullptr;
    int L = 0, nl = 0, nhood = 0;
    double start = 0, length = 0, dt = 0;
    int64_t T = 0, Tn = 0, Tp = 0;
    int n_tiles = 0;
    const double* k = nullptr;    // host coefficients [nl][9] (host groups)
    const double* pos = nullptr;  // host lead positions [L][3] (host groups)
};

// A launch group: requests sharing the neighbourhood and the kernel path.
struct GroupPlan {
    std::vector<ReqSpec> reqs;
    int LP = 4, nhood = 0, waves = 2, lf_phase = 2;
    bool mfma = true;
};

int nhood_offsets(int type, int8_t dif[26][3]);
int64_t num_steps(double length, double dt);
// Validates one descriptor against its heart (shape only when k/pos are device data).
ReqSpec make_spec(Heart& h, const ekg_sim_desc& d, bool host_values);
// Builds a group from specs (all with the same neighbourhood).
GroupPlan plan_group(Ctx& ctx, std::vector<ReqSpec> reqs);
// Builds (or keeps) the heart's layout for the matrix-core path (mfma, with
// activation classes as tuned) or the vector-ALU path (no classes).
void ensure_layout(Heart& h, bool mfma);
// Whether ensure_layout (and, on the matrix-core path, ensure_cells for
// nhood) would find everything built already.
bool layout_ready(const Heart& h, bool mfma, int nhood);
// Cell lists of the heart's class chunks for dipole neighbourhood `nhood`.
void ensure_cells(Heart& h, int nhood);
// Whether the heart's simulations take the matrix-core path (as plan_group decides).
bool heart_mfma(const Ctx& ctx, const Heart& h);
size_t group_device_bytes(const GroupPlan& g);
// Enqueue a whole group on `stream`.  With d_k/d_pos null the host values of
// the specs are packed and uploaded; otherwise request r reads
// d_k[r*nl*9 ...] and d_pos[r*L*3 ...] (uniform device batches).  ECGs go to
// d_ecg (null: the workspace's ecg buffer), request r at the packed offset
// sum_{q<r} L_q T_q.
// allow_graph false: eager launches even in graph mode.
void enqueue_group(Ctx& ctx, GroupPlan& g, Workspace& ws, hipStream_t stream, const double* d_k,
                   const double* d_pos, double* d_ecg, bool time_fused, bool allow_graph = true);
void simulate(Heart& h, int batch, const ekg_sim_desc* descs, double* ecg);
// Times the tunable knobs on `batch` copies of `sample` and keeps the fastest
// on the heart: the contraction path, then for the matrix-core path every
// (chunk size, waves per block) pair, for the vector-ALU path the chunk size
// and the lead-field phase width.
ekg_tune autotune(Heart& h, const ekg_sim_desc& sample, int batch);
void simulate_device(Heart& h, const ekg_batch_dev& d);
// simulate_device with every buffer in the device's static workspace: no
// device allocation, eager launches; the heart's layout must be built.
void simulate_device_fixed(Heart& h, const ekg_batch_dev& d);
// Criteria of device ECGs on the context's stream (targets uploaded from the host).
void criteria_device(Ctx& ctx, const ekg_criteria_desc& c, const double* d_ecg, int batch, int n_leads,
                     int64_t n_steps, double* d_out);
void simulate_criteria(Heart& h, int batch, const ekg_sim_desc* descs, const ekg_criteria_desc& c,
                       double* criteria);

void excite(Heart& h, bool is_int, const double* transfer, int rows, int cols,
            const uint8_t* delay);

}  // namespace ekg
