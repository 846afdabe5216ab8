// ekg_kernels.hpp -- launch interface of the gfx950 kernels (ekg_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stdint.h>

namespace ekgk {

// Spare voxel rows per request in the matrix-core layout (groups of 4 voxels
// plus a one-group look-ahead past the last chunk).
constexpr int MFMA_PAD = 8;

template <bool kInt> struct CaT { using type = double; };
template <> struct CaT<true> { using type = int; };

// Per (request, layer) constants of the Wohlfart AP in the exp2 domain:
//   V(tau) = S * (alpha*exp2(c4*tau) + beta) * exp2(c5*tau) * R(t) + k0,
//   S = 1/(1 + exp2(c1*tau)),  c_i = -k_i*log2(e),  alpha = k2(1-k3), beta = k2*k3.
struct LayerConst {
    float c1, c4, c5, alpha, beta;
    float skip;  // 1: voxels far before their upstroke may be skipped (k4 + k5 <= k1 / 4)
    double k0;
};
static_assert(sizeof(LayerConst) == 32, "LayerConst layout");

struct LeadFieldArgs {
    const uint8_t* grid_layer;   // [n0*n1*n2]
    int n0, n1, n2;
    const int4* bricks;          // [n_bricks] {z0, y0, x0, 0} of bricks holding myocardium
    const int* brick_off;        // [n_bricks + 1] offsets into brick_vox
    const int2* brick_vox;       // {sorted voxel index, local cell lz*64 + ly*8 + lx}
    const int* halo_off;         // [n_bricks + 1] offsets into halo_cells
    const uint16_t* halo_cells;  // myocardial cells of each brick + halo (index in the 10^3 box)
    int n_bricks;
    const int* vox_chunk;        // [V] chunk of each sorted voxel
    const double* vox_a64;       // [V] activation time
    const float2* vox_a;         // [V] activation time (hi, lo)
    const int4* chunks;          // [n_chunks] {layer, v_begin, v_end, t_ref as float bits}
    int n_chunks;
    int nhood;                   // EKG_NHOOD_* (dipole neighbourhood)
    const double* lead_pos;      // [n_req][n_leads][3]
    int n_leads;
    const double* layer_k;       // [n_req][n_layers][9]
    int n_layers;
    int64_t V;
    int LP;                      // padded lead count (4, 8, 12, 16)
    int n_req;
    double* rec;                 // VALU: [n_req][V + 1][LP + 2] voxel records (G, P/Q, a)
                                 // MFMA: [n_req][V + MFMA_PAD][16] G rows
    float4* aux;                 // MFMA: [n_req][V + MFMA_PAD] {P, Q, a_hi, a_lo}
    double* gsum;                // [n_req][n_chunks][3][LP]: sum G, sum G P, sum G Q
    bool mfma;
};

struct FusedArgs {
    const float2* vox_a;         // [V] activation time (hi, lo)
    const double* rec;           // see LeadFieldArgs::rec
    const float4* aux;           // MFMA only
    const double* gsum;          // [n_req][n_chunks][3][LP]: sum G, sum G P, sum G Q
    const int4* chunks;          // [n_chunks]
    const double* vox_a64;       // [V] activation time, FP64
    const LayerConst* lc;        // [n_req][n_layers]
    const double* layer_k;       // [n_req][n_layers][9]
    const double* rtab;          // [n_req][n_layers][Tp]
    const double* tgrid;         // [Tp] time
    const float2* thl;           // [Tp] time (hi, lo)
    double* partial;             // [n_req][n_chunks][LP][Tp]
    int64_t V;
    int n_layers, n_chunks, Tp, n_tiles, n_groups, LP, n_req, n_leads;
    bool mfma;
};

hipError_t launch_ca_init(bool is_int, const uint8_t* layer, const uint8_t* seed, void* dist,
                          uint8_t* active, int n0, int n1, int n2, int nb1, int nb2, hipStream_t s);
hipError_t launch_ca_sweep(bool is_int, const uint8_t* layer, const uint8_t* delay,
                           const double* wtab, int wcols, void* dist, int n0, int n1, int n2,
                           int nb0, int nb1, int nb2, const uint8_t* act_in, uint8_t* act_out,
                           int* any_marked, hipStream_t s);
hipError_t launch_ca_finalize(bool is_int, const void* dist, double* act, int64_t n, hipStream_t s);
hipError_t launch_lead_field(const LeadFieldArgs& a, hipStream_t s);
hipError_t launch_ap_tables(const double* layer_k, int n_layers, const double* tgrid, int Tp,
                            int n_req, double* rtab, LayerConst* lc, hipStream_t s);
hipError_t launch_ecg_fused(const FusedArgs& a, hipStream_t s);
hipError_t launch_ecg_reduce(const double* partial, double* ecg, int n_req, int n_chunks, int LP,
                             int n_leads, int T, int Tn, int Tp, hipStream_t s);

}  // namespace ekgk
