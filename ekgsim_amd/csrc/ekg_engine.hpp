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

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Device workspace of one batch of simulations on one stream.
struct Workspace {
    DevBuf layer_k, lead_pos, tgrid, thl, rtab, lc, G, aux, gsum, partial, ecg;  // G: voxel records
    std::vector<double> h_tgrid, h_k, h_pos;
    std::vector<float> h_thl;
    double grid_dt = 0.0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Workspace ws;
    bool profiling = false;
    int kernel = EKG_KERNEL_AUTO;  // contraction path of the fused kernel
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;  // fused launches not yet timed
    std::vector<hipEvent_t> free_events;
    int64_t fused_launches = 0;
    double fused_ms = 0.0;
    DevBuf nb_table;  // [4][26][4] int8 {d0, d1, d2, cube bit} of the four dipole neighbourhoods
    std::mutex mu;    // serialises ABI calls on this context
    void bind() const;
    hipEvent_t get_event();
    void collect_timings();
};

// Per-heart layout for the ECG kernels: myocardial voxels sorted by
// (layer, activation time, grid index), cut into layer-pure chunks.
struct EcgLayout {
    bool valid = false;
    int64_t V = 0;
    int n_chunks = 0;
    int n_bricks = 0;
    int max_layer = 0;
    std::vector<int4> h_chunks;
    DevBuf vox_chunk, vox_a64, vox_a, chunks, bricks, brick_off, brick_vox, halo_off, halo_cells;
};

struct Heart {
    Ctx* ctx = nullptr;
    int n0 = 0, n1 = 0, n2 = 0;
    int64_t nvox = 0, n_myo = 0;
    int max_layer = 0;
    std::vector<uint8_t> h_layer, h_seed;
    DevBuf d_layer, d_seed, d_act;
    std::vector<double> h_act;
    bool excited = false;
    EcgLayout layout;
    std::mutex mu;
};

// Host-side description of one batch, validated.
struct BatchPlan {
    int B = 0, L = 0, LP = 0, n_layers = 0, nhood = 0;
    bool mfma = true;
    int64_t T = 0, Tn = 0, Tp = 0;
    int n_tiles = 0, n_groups = 0;
    double start = 0, length = 0, dt = 0;
};

int nhood_offsets(int type, int8_t dif[26][3]);
const void* ctx_nb_table(const Ctx& ctx, int nhood);
int64_t num_steps(double length, double dt);
BatchPlan plan_shape(const Heart& h, BatchPlan p);
bool use_mfma(const Ctx& ctx, int n_leads);
BatchPlan plan_batch(const Heart& h, int batch, const ekg_sim_desc* descs);
void ensure_layout(Heart& h);
// Copy the per-request coefficients and lead positions of host descriptors
// into ws.layer_k / ws.lead_pos (and size ws.ecg).
void upload_batch(const BatchPlan& p, const ekg_sim_desc* descs, Workspace& ws, hipStream_t stream);
// Enqueue the whole batch on `stream`: d_layer_k [B][n_layers][9] and
// d_lead_pos [B][L][3] in device memory, ECGs to d_ecg [B][L][T] FP64.
void enqueue_batch(Ctx& ctx, Heart& h, const BatchPlan& p, const double* d_layer_k,
                   const double* d_lead_pos, double* d_ecg, Workspace& ws, hipStream_t stream,
                   bool time_fused);
size_t batch_device_bytes(const Heart& h, const BatchPlan& p);
void simulate(Heart& h, int batch, const ekg_sim_desc* descs, double* ecg);
void simulate_device(Heart& h, const ekg_batch_dev& d);

void excite(Heart& h, bool is_int, const double* transfer, int rows, int cols,
            const uint8_t* delay);

}  // namespace ekg
