// ekg_kernels.hpp -- launch interface of the gfx950 kernels (ekg_ca.hip,
// ekg_ecg.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stdint.h>

namespace ekgk {

// Spare voxel rows per request in the matrix-core layout (groups of 4 voxels
// plus a one-group look-ahead past the last chunk).
constexpr int MFMA_PAD = 8;

// Largest chunk (myocardial voxels) of a layout: the chunk kernel keeps a
// whole chunk's G rows in LDS.
constexpr int CHUNK_MAX = 512;

// Bound tiles: the chunk bounds {z, f}, the prefix tables and the reduction
// work in tiles of 2^BT_SHIFT steps -- 8 on the matrix-core path (its
// transition tiles are two 4-step quads; measured at 128^3 x 256 requests
// without classes: 16-step tiles 6.29 ms, 8 5.88 ms, 4 6.40 ms per chunk-kernel
// launch), 64 on the vector-ALU path (one step per lane of a wave).
constexpr int BT_SHIFT_MFMA = 3;
constexpr int BT_SHIFT_VALU = 6;
// Activation classes per class window, and groups of 16 window entries per
// piece (one class-window block per piece and request).  Six classes keep the
// kernel's accumulators within 64 VGPRs (eight waves per SIMD); eight share
// more lead weights but leave five waves and run slower.
constexpr int CLASS_WINDOW = 6;
// Largest class-window piece (groups of 16 entries); a window is cut into
// the fewest pieces of at most this size, all of about equal size.  Measured
// per 256 x 12 step on the 128^3 heart: 1.46 ms at 1024 against 1.48-1.49 ms
// with 512 (equal or fixed pieces), 1.49 ms at 1536 and 1.54 ms at 2048.
constexpr int CLASS_PIECE = 1024;

// ---- excitation sequence ---------------------------------------------------

template <bool kInt> struct CaT { using type = double; };
template <> struct CaT<true> { using type = int; };

hipError_t launch_ca_init(bool is_int, const uint8_t* layer, const uint8_t* seed, void* dist,
                          uint8_t* active, int n0, int n1, int n2, int nb1, int nb2, hipStream_t s);
hipError_t launch_ca_sweep(bool is_int, const uint8_t* layer, const uint8_t* delay,
                           const double* wtab, int wcols, void* dist, int n0, int n1, int n2,
                           int nb0, int nb1, int nb2, const uint8_t* act_in, uint8_t* act_out,
                           int* any_marked, hipStream_t s);
hipError_t launch_ca_finalize(bool is_int, const void* dist, double* act, int64_t n, hipStream_t s);

// ---- ECG --------------------------------------------------------------------

// Per (request, layer) constants of the Wohlfart AP in the exp2 domain:
//   V(tau) = S * (alpha*exp2(c4*tau) + beta) * exp2(c5*tau) * R(t) + k0,
//   S = 1/(1 + exp2(c1*tau)),  c_i = -k_i*log2(e),  alpha = k2(1-k3), beta = k2*k3.
struct LayerConst {
    float c1, c4, c5, alpha, beta;
    float skip;  // 1: voxels far before their upstroke may be skipped (k4 + k5 <= k1 / 4)
    double k0;
};
static_assert(sizeof(LayerConst) == 32, "LayerConst layout");

// One heart of a launch group, as the kernels see it.
struct HeartDev {
    const uint8_t* grid_layer;   // [n0*n1*n2] layer per grid voxel
    const int4* bricks;          // [n_bricks] {z0, y0, x0, 0} of 8^3 bricks holding myocardium
    const int* brick_off;        // [n_bricks + 1] offsets into brick_vox
    const int2* brick_vox;       // {sorted voxel index, local cell lz*64 + ly*8 + lx}
    const int* halo_off;         // [n_bricks + 1] offsets into halo_cells
    const uint16_t* halo_cells;  // myocardial cells of each brick + halo (index in the 10^3 box)
    const int* vox_chunk;        // [V] chunk of each sorted voxel
    const double* vox_a64;       // [V] activation time
    const float2* vox_a;         // [V] activation time (hi, lo)
    const uint2* vox_geo;        // [V] {z | y << 10 | x << 20, 3x3x3 myocardium mask (bit (dz+1)*9+(dy+1)*3+dx+1)}
    const int4* chunks;          // [n_chunks] {layer, v_begin, v_end, t_ref as float bits}
    const int* layer_chunk;      // [max_layer + 2]: chunks of layer l are [layer_chunk[l], layer_chunk[l+1])
    const double* layer_t0;      // [max_layer + 1]: reference time of layer l
    // Activation classes: runs of one layer and one activation time, whose
    // voxels all have the same AP.  Their voxels are cut into class chunks
    // (index n_chunks + k of the per-chunk sums); the chunks above hold the
    // other voxels.
    const int4* cchunks;         // [n_cchunks] {layer, v_begin, v_end, t_ref as float bits}
    const int4* classes;         // [n_classes] {layer, first class chunk, end class chunk, 0}
    const double* class_a;       // [n_classes] activation time of the class
    // class windows for the group's neighbourhood: CLASS_WINDOW classes each
    // (window w: classes [w CLASS_WINDOW, (w+1) CLASS_WINDOW)), their entries
    // (cells with a non-zero dipole vector for a class of the window) in
    // groups of 16, group q in rows crow[q].x, ... of 64 floats of cgrp
    // {coordinate, |c|^2, then a coefficient row per class of the mask
    // crow[q].y}; the groups cut into pieces {window, first group, end group},
    // those of window w being [cwin[w], cwin[w+1])
    const float* cgrp;
    const int2* crow;
    const int4* cpieces;
    const int* cwin;
    int n0, n1, n2;
    int V, n_chunks, n_bricks, max_layer;
    int n_cchunks, n_classes;
};

// One request of a launch group.  Offsets index the group's packed arrays.
struct ReqDev {
    int heart, L, nl, T, Tn, Tp, n_tiles;
    int bt_shift;     // log2 of the bound tiles of the chunk bounds and prefix tables (16 or 64 steps)
    int64_t k_off;    // layer_k[k_off + 9*(l-1) + q]
    int64_t pos_off;  // lead_pos[pos_off + 3*i + c]
    int64_t ecg_off;  // ecg[ecg_off + i*T + j]
    int64_t t_off;    // tgrid/thl[t_off + j], j < Tp
    int64_t lay_off;  // lc[lay_off + l-1]
    int64_t xy_off;   // xy[xy_off + (l-1)*Tp + j] = {X_l(t_j), Y_l(t_j)}
    int64_t g_off;    // record row of sorted voxel v: g_off + v
    int64_t c_off;    // per-chunk arrays: sums[(c_off + c)*3*LP ...], bnd[c_off + c]
    int64_t p_off;    // partial[p_off + (c*LP + i)*Tp + j]
    int64_t pre_off;  // prefix tables: [(pre_off + (l-1)*(n_bt+1) + bound tile)], n_bt = Tp >> bt_shift
    int64_t cls_off;  // class sums: csums[(cls_off + class)*3*LP ...]
};

// A launch group: requests of any hearts, lead counts and time grids that
// share the dipole neighbourhood, the padded lead count LP and the kernel
// path.  Work-item starts are exclusive prefix sums over requests (B + 1).
struct GroupArgs {
    const HeartDev* hearts;
    const ReqDev* reqs;
    int B, LP, nhood, mfma, mfma_waves;
    int lf_phase;              // leads per lead-field phase (2 or 4)
    int chunk_max;             // largest chunk target of the group's hearts (chunk kernel LDS rows)
    int cs_max;                // most activation classes of one request
    const int64_t* lf_start;   // lead-field blocks (bricks) per request
    const int64_t* ch_start;   // chunks per request
    const int64_t* xy_start;   // nl * Tp per request
    const int64_t* lay_start;  // nl per request
    const int64_t* out_start;  // L * T per request
    const int64_t* cs_start;   // activation classes per request
    const int64_t* cl_start;   // 64-step tiles of the class kernel per request (0 without classes)
    const int64_t* red_start;  // reduction items per request: L x (T rounded up to a bound tile)
    int64_t n_lf, n_ch, n_xy, n_lay, n_out, n_cc, n_cs, n_cl, n_red;  // totals
    int ch_uniform;            // chunks per request when all requests have the same count, else 0
    int one_heart;             // every request on one heart (tables hold it as heart 0)
    // class-window blocks: runs of consecutive requests of one heart
    // {first request, requests, pieces of the heart}, blocks of run u
    // [cc_rstart[u], cc_rstart[u+1]) = pieces x requests
    const int4* cc_runs;
    const int64_t* cc_rstart;
    int n_runs;
    int cl_per_req;            // class-kernel blocks per request when equal for all and B % 8 == 0, else 0
    const double* layer_k;
    const double* lead_pos;
    const double* tgrid;       // FP64 sample times
    const float2* thl;         // FP32 hi/lo of the same times
    LayerConst* lc;
    double2* xy;
    double* rec;               // 16-double G row per voxel (leads zero-padded)
    float4* aux;               // {P, Q, a_hi, a_lo} per voxel
    double* sums;              // per (request, chunk): [3][LP] sum G, sum G P', sum G Q'
    int2* bnd;                 // per (request, chunk): {first non-zero, first saturated} bound tile
    int2* pre_rng;             // per (request, layer, bound tile): transition chunks [lo, hi)
    double* pre_ab;            // per (request, layer, bound tile): [2][LP]; tile n_bt: [LP] sum G
    double* partial;
    double* csums;             // per (request, class): [3][LP] sum G, sum G P', sum G Q'
    double* ecg;
};

hipError_t launch_ap_tables(const GroupArgs& a, hipStream_t s);
hipError_t launch_lead_field(const GroupArgs& a, hipStream_t s);
hipError_t launch_chunk_sums(const GroupArgs& a, hipStream_t s);
hipError_t launch_layer_prefix(const GroupArgs& a, hipStream_t s);
hipError_t launch_ecg_fused(const GroupArgs& a, hipStream_t s);
// Matrix-core path: lead field, chunk sums and transition tiles in one kernel.
hipError_t launch_ecg_chunk(const GroupArgs& a, hipStream_t s);
hipError_t launch_ecg_reduce(const GroupArgs& a, hipStream_t s);
// Resolves every matrix-core-path kernel once (see ekg_ecg.hip).
void prime_kernels();
// Activation classes: the lead-weight sums of the class chunks, their sums
// per class, then the classes' terms added to the reduced ECG.
hipError_t launch_class_windows(const GroupArgs& a, hipStream_t s);
hipError_t launch_class_sums(const GroupArgs& a, hipStream_t s);
hipError_t launch_ecg_class(const GroupArgs& a, hipStream_t s);

// ---- fixed workspace of the allocation-free entry point ----------------------

// Bytes of the static device workspace (ekg_scratch.hip), per GPU.
constexpr size_t SCRATCH_BYTES = (size_t)256 << 20;
// Device address of `device`'s copy of the workspace (resolved once per
// device with that device current, then cached).
hipError_t scratch_ptr(int device, void** out);

// ---- criteria -----------------------------------------------------------------

// stats: per target {mean, variance, b . b, first maximum} over the first len samples
hipError_t launch_target_stats(const double* targets, int n_targets, int64_t target_len, int64_t len, double* stats,
                               hipStream_t s);
hipError_t launch_criteria(int mode, int peaks, int raw, const double* ecg, int B, int L, int64_t T,
                           const double* targets, const double* offsets, const double* stats, int n_crit,
                           int64_t target_len, double* out, hipStream_t s);

}  // namespace ekgk
