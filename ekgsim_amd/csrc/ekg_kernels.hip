// ekg_kernels.hip -- CDNA4 (gfx950) kernels of the ekgsim simulation core.
//
//   ca_*            excitation sequence: min-plus relaxation over LDS-staged
//                   8^3 voxel bricks with a 1-voxel halo, bit-exact with the
//                   reference's priority-queue flood (simulator.cpp:212-286)
//   lead_field      per-voxel lead weights G_i(u) in FP64 (regrouped
//                   simulator.cpp:509-537) plus per-chunk sums of G
//   ap_tables       per (request, layer, t) repolarisation factor and
//                   per-layer FP32 constants of the Wohlfart AP (Wohlfart.h:195-203)
//   ecg_fused       AP evaluation fused with the lead contraction: one
//                   wave = 64 consecutive time steps of one voxel chunk,
//                   voxel data in SGPRs, FP32 AP, FP64 FMA accumulation
//   ecg_reduce      fixed-order FP64 sum of the per-chunk partial ECGs
//
// Every launch is checked on the host against the buffer extents it assumes
// (ekg_engine.cpp) before it is issued.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ekg_kernels.hpp"

namespace ekgk {

// ---------------------------------------------------------------------------
// Excitation sequence
// ---------------------------------------------------------------------------

constexpr int CA_B = 8;                      // brick edge (voxels)
constexpr int CA_H = CA_B + 2;               // brick edge with halo
constexpr int CA_N = CA_H * CA_H * CA_H;     // 1000 halo cells
constexpr int CA_THREADS = CA_B * CA_B * CA_B;

template <class T> __device__ __forceinline__ T ca_inf();
template <> __device__ __forceinline__ double ca_inf<double>() { return __builtin_inf(); }
template <> __device__ __forceinline__ int ca_inf<int>() { return 0x7fffffff; }

// dist := 1 at seeds, +inf elsewhere; flag the bricks holding seeds.
template <class T>
__global__ void ca_init_kernel(const uint8_t* __restrict__ layer, const uint8_t* __restrict__ seed,
                               T* __restrict__ dist, uint8_t* __restrict__ active, int n0, int n1,
                               int n2, int nb1, int nb2)
{
    int64_t n = (int64_t)n0 * n1 * n2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool s = seed[i] != 0 && layer[i] != 0;
        dist[i] = s ? (T)1 : ca_inf<T>();
        if (s) {
            int64_t x = i % n2, r = i / n2;
            int64_t y = r % n1, z = r / n1;
            active[((z / CA_B) * nb1 + y / CA_B) * nb2 + x / CA_B] = 1;
        }
    }
}

// One relaxation sweep over the active bricks.  Within a brick the 512
// threads (one per voxel) iterate Jacobi min-plus updates on the LDS copy
// until no voxel improves (__syncthreads_or is the block-wide ballot); the
// improved values are written back and, when a brick-boundary voxel
// improved, the neighbouring bricks are flagged for the next sweep.
//
// kInt = false: FP64, lag(u->v) = wtab[(layer(u)*wcols + layer(v))*4 + |d|^2]
//               with wtab = transfer * sqrt(|d|^2) computed on the host in
//               FP64, so each candidate is one IEEE add, the same rounding as
//               `active.time + lag` in simulator.cpp:241.
// kInt = true:  int32, lag(u->v) = delay[v].
// Values only decrease and each is the value of a real path, so any order
// of relaxation reaches the least fixpoint, which is what the reference's
// Dijkstra computes when every lag is >= 0 (checked on the host).
template <bool kInt>
__global__ __launch_bounds__(CA_THREADS) void ca_sweep_kernel(
    const uint8_t* __restrict__ layer, const uint8_t* __restrict__ delay,
    const double* __restrict__ wtab, int wcols,
    typename CaT<kInt>::type* __restrict__ dist, int n0, int n1, int n2, int nb0, int nb1,
    int nb2, const uint8_t* __restrict__ act_in, uint8_t* __restrict__ act_out,
    int* __restrict__ any_marked)
{
    using T = typename CaT<kInt>::type;
    __shared__ T s_d[CA_N];
    __shared__ uint8_t s_l[CA_N];
    __shared__ int s_mark[27];

    const int brick = blockIdx.x;
    if (!act_in[brick]) return;
    const int bz = brick / (nb1 * nb2);
    const int br = brick - bz * nb1 * nb2;
    const int by = br / nb2, bx = br - by * nb2;
    const int z0 = bz * CA_B, y0 = by * CA_B, x0 = bx * CA_B;

    for (int i = threadIdx.x; i < CA_N; i += blockDim.x) {
        int hz = i / (CA_H * CA_H), hy = (i / CA_H) % CA_H, hx = i % CA_H;
        int gz = z0 - 1 + hz, gy = y0 - 1 + hy, gx = x0 - 1 + hx;
        uint8_t lh = 0;
        T dh = ca_inf<T>();
        if (gz >= 0 && gz < n0 && gy >= 0 && gy < n1 && gx >= 0 && gx < n2) {
            const int64_t g = ((int64_t)gz * n1 + gy) * n2 + gx;
            lh = layer[g];
            dh = dist[g];
        }
        s_l[i] = lh;
        s_d[i] = dh;
    }
    if (threadIdx.x < 27) s_mark[threadIdx.x] = 0;
    __syncthreads();

    const int tz = threadIdx.x / (CA_B * CA_B);
    const int ty = (threadIdx.x / CA_B) % CA_B;
    const int tx = threadIdx.x % CA_B;
    const int gz = z0 + tz, gy = y0 + ty, gx = x0 + tx;
    const bool inside = gz < n0 && gy < n1 && gx < n2;
    const int me = ((tz + 1) * CA_H + (ty + 1)) * CA_H + (tx + 1);
    const int lv = inside ? s_l[me] : 0;
    const int64_t g = ((int64_t)gz * n1 + gy) * n2 + gx;
    const T orig = s_d[me];
    T cur = orig;
    int dly = 0;
    if (kInt && lv) dly = delay[g];

    for (;;) {
        T cand = cur;
        if (lv) {
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dz == 0 && dy == 0 && dx == 0) continue;
                        const int nb = me + (dz * CA_H + dy) * CA_H + dx;
                        const int lu = s_l[nb];
                        if (lu == 0) continue;
                        const T du = s_d[nb];
                        if (du == ca_inf<T>()) continue;
                        T c;
                        if constexpr (kInt) {
                            c = du + dly;
                        } else {
                            const int k = dz * dz + dy * dy + dx * dx;
                            c = du + wtab[(lu * wcols + lv) * 4 + k];
                        }
                        cand = c < cand ? c : cand;
                    }
        }
        __syncthreads();
        const int ch = cand < cur;
        if (ch) {
            cur = cand;
            s_d[me] = cur;
        }
        if (!__syncthreads_or(ch)) break;
    }

    if (lv && cur < orig) {
        dist[g] = cur;
        const int oz0 = tz == 0 ? -1 : 0, oz1 = tz == CA_B - 1 ? 1 : 0;
        const int oy0 = ty == 0 ? -1 : 0, oy1 = ty == CA_B - 1 ? 1 : 0;
        const int ox0 = tx == 0 ? -1 : 0, ox1 = tx == CA_B - 1 ? 1 : 0;
        for (int oz = oz0; oz <= oz1; ++oz)
            for (int oy = oy0; oy <= oy1; ++oy)
                for (int ox = ox0; ox <= ox1; ++ox)
                    if (oz | oy | ox) s_mark[((oz + 1) * 3 + (oy + 1)) * 3 + (ox + 1)] = 1;
    }
    __syncthreads();
    if (threadIdx.x < 27 && s_mark[threadIdx.x]) {
        const int oz = threadIdx.x / 9 - 1, oy = (threadIdx.x / 3) % 3 - 1, ox = threadIdx.x % 3 - 1;
        const int nz = bz + oz, ny = by + oy, nx = bx + ox;
        if (nz >= 0 && nz < nb0 && ny >= 0 && ny < nb1 && nx >= 0 && nx < nb2) {
            act_out[(nz * nb1 + ny) * nb2 + nx] = 1;
            *any_marked = 1;
        }
    }
}

// Unreached (still +inf) and empty voxels -> 0.0 (simulator.cpp:219 marks
// "not excited" with 0); the result is the FP64 activation-time grid.
template <class T>
__global__ void ca_finalize_kernel(const T* __restrict__ dist, double* __restrict__ act, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T d = dist[i];
        act[i] = (d == ca_inf<T>()) ? 0.0 : (double)d;
    }
}

// ---------------------------------------------------------------------------
// Lead field
// ---------------------------------------------------------------------------

// f(e) = e * (1/(l*sqrt(l))), l = |e|^2: the lead weight m of simulator.cpp:531-533
// for e = p - (u+1).  1/(l*sqrt(l)) = rsq(l)^3 with two Newton steps on the
// hardware FP64 reciprocal square root (quadratic convergence: a few ulp).
__device__ __forceinline__ void lead_f(double e0, double e1, double e2, double& m0, double& m1,
                                       double& m2)
{
    double l = e0 * e0;
    l = __builtin_fma(e1, e1, l);
    l = __builtin_fma(e2, e2, l);
    double r = __builtin_amdgcn_rsq(l);
    const double h = 0.5 * l;
    r = r * __builtin_fma(-h * r, r, 1.5);
    r = r * __builtin_fma(-h * r, r, 1.5);
    const double s = r * r * r;
    m0 = e0 * s;
    m1 = e1 * s;
    m2 = e2 * s;
}

constexpr int LB = 8;                        // lead-field brick edge
constexpr int LH = LB + 2;                   // with halo
constexpr int LN = LH * LH * LH;             // halo cells

// One block per (8^3 brick holding myocardium, request).  For each lead i
// the lead weight m_i of every myocardial cell of the brick and its 1-voxel
// halo is computed once into LDS; each myocardial voxel u of the brick then
// gathers
//   G_i(u) = sum_{d in N, u+d myocardial} (m_i(u+d) + m_i(u)) . d
//          = sum_d m_i(u+d) . d + m_i(u) . s_u,   s_u = sum of those d,
// the coefficient of V_u(t) in the reference's dipole sum regrouped by cell
// (simulator.cpp:509-537).  The voxel's record (sorted index v) receives
//   VALU layout (RS = LP + 2 doubles at rec[(b*(V+1) + v)*RS]):
//     [0, LP) G zero-padded, LP: (P_u, Q_u) floats, LP+1: (a_hi, a_lo) floats;
//   MFMA layout: G rows of 16 doubles at rec[(b*(V+MFMA_PAD) + v)*16] and
//     float4 {P_u, Q_u, a_hi, a_lo} at aux[b*(V+MFMA_PAD) + v];
// with P_u = k2(1-k3) e^{(k4+k5)(a_u - t_ref)}, Q_u = k2 k3 e^{k5 (a_u - t_ref)},
// t_ref the reference time of the voxel's chunk (saturated-path AP factors).
// Dipole neighbourhoods at compile time, in Neighbourhood::create order
// (simulator.h:325-390): NH = 0 2D4, 1 2D8, 2 3D4, 3 3D8.
template <int NH> struct Nhood;
template <> struct Nhood<0> {
    static constexpr int n = 4;
    static constexpr int8_t d[4][3] = {{0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};
};
template <> struct Nhood<1> {
    static constexpr int n = 8;
    static constexpr int8_t d[8][3] = {{0, -1, -1}, {0, -1, 0}, {0, -1, 1}, {0, 0, -1},
                                       {0, 0, 1},   {0, 1, -1}, {0, 1, 0},  {0, 1, 1}};
};
template <> struct Nhood<2> {
    static constexpr int n = 8;
    static constexpr int8_t d[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                                       {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
};
template <> struct Nhood<3> {
    static constexpr int n = 26;
    static constexpr int8_t d[26][3] = {
        {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1}, {-1, 1, -1},
        {-1, 1, 0},   {-1, 1, 1},  {0, -1, -1}, {0, -1, 0},  {0, -1, 1}, {0, 0, -1}, {0, 0, 1},
        {0, 1, -1},   {0, 1, 0},   {0, 1, 1},   {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1},
        {1, 0, 0},    {1, 0, 1},   {1, 1, -1},  {1, 1, 0},   {1, 1, 1}};
};

// One block of 512 threads per (8^3 brick holding myocardium, request).
// For two leads at a time the lead weight m_i of every myocardial cell of the
// brick and its 1-voxel halo (a compact per-brick list) is computed once into
// LDS (component arrays, zero outside the myocardium); thread t < n then
// gathers for the brick's t-th myocardial voxel u
//   G_i(u) = sum_{d in N} M_i(u+d) . d + m_i(u) . s_u,
//   M_i = m_i on myocardium and 0 elsewhere,  s_u = sum of d with u+d myocardial,
// which is sum_{d in N, u+d myo} (m_i(u+d) + m_i(u)) . d -- the coefficient of
// V_u(t) in the reference's dipole sum regrouped by cell (simulator.cpp:509-537).
// The voxel's record (sorted index v) receives
//   VALU layout (RS = LP + 2 doubles at rec[(b*(V+1) + v)*RS]):
//     [0, LP) G zero-padded, LP: (P_u, Q_u) floats, LP+1: (a_hi, a_lo) floats;
//   MFMA layout: G rows of 16 doubles at rec[(b*(V+MFMA_PAD) + v)*16] and
//     float4 {P_u, Q_u, a_hi, a_lo} at aux[b*(V+MFMA_PAD) + v];
// with P_u = k2(1-k3) e^{(k4+k5)(a_u - t_ref)}, Q_u = k2 k3 e^{k5 (a_u - t_ref)},
// t_ref the reference time of the voxel's chunk (saturated-path AP factors).
template <int LP, bool kMfma, int NH>
__global__ __launch_bounds__(512) void lead_field_kernel(
    const uint8_t* __restrict__ grid_layer, int n0, int n1, int n2,
    const int4* __restrict__ bricks, const int* __restrict__ brick_off,
    const int2* __restrict__ brick_vox, const int* __restrict__ halo_off,
    const uint16_t* __restrict__ halo_cells, const int* __restrict__ vox_chunk,
    const double* __restrict__ vox_a64, const float2* __restrict__ vox_a,
    const int4* __restrict__ chunks, const double* __restrict__ lead_pos, int n_leads,
    const double* __restrict__ layer_k, int n_layers, int64_t V, double* __restrict__ rec,
    float4* __restrict__ aux)
{
    using N = Nhood<NH>;
    constexpr int RS = kMfma ? 16 : LP + 2;
    constexpr int NG = kMfma ? 16 : LP;
    __shared__ uint8_t s_myo[LN];
    __shared__ double s_m[2][3][LN];
    const int br = blockIdx.x;
    const int b = blockIdx.y;
    const int t = threadIdx.x;
    const int4 bi = bricks[br];
    const int e0 = brick_off[br], nv = brick_off[br + 1] - e0;
    const int c0 = halo_off[br], nc = halo_off[br + 1] - c0;
    for (int q = t; q < LN; q += blockDim.x) {
        s_myo[q] = 0;
#pragma unroll
        for (int w = 0; w < 2; ++w) s_m[w][0][q] = s_m[w][1][q] = s_m[w][2][q] = 0.0;
    }
    __syncthreads();
    for (int q = t; q < nc; q += blockDim.x) s_myo[halo_cells[c0 + q]] = 1;
    __syncthreads();
    int h = 0, v = -1;
    double sd0 = 0.0, sd1 = 0.0, sd2 = 0.0;
    if (t < nv) {
        const int2 bv = brick_vox[e0 + t];
        v = bv.x;
        const int lz = bv.y >> 6, ly = (bv.y >> 3) & 7, lx = bv.y & 7;
        h = ((lz + 1) * LH + (ly + 1)) * LH + (lx + 1);
#pragma unroll
        for (int q = 0; q < N::n; ++q) {
            const double f = s_myo[h + (N::d[q][0] * LH + N::d[q][1]) * LH + N::d[q][2]];
            sd0 += f * N::d[q][0];
            sd1 += f * N::d[q][1];
            sd2 += f * N::d[q][2];
        }
    }
    const double* pos = lead_pos + (int64_t)b * n_leads * 3;
    const int64_t rows = V + (kMfma ? MFMA_PAD : 1);
    // the voxel's whole record is assembled in registers and stored once
    double row[RS];
#pragma unroll
    for (int i = 0; i < RS; ++i) row[i] = 0.0;
#pragma unroll
    for (int i = 0; i < LP; i += 2) {
        if (i >= n_leads) break;
        const bool two = i + 1 < n_leads;
        for (int q = t; q < nc; q += blockDim.x) {
            const int cell = halo_cells[c0 + q];
            const int gz = bi.x - 1 + cell / (LH * LH), gy = bi.y - 1 + (cell / LH) % LH,
                      gx = bi.z - 1 + cell % LH;
            const double z = (double)(gz + 1), y = (double)(gy + 1), x = (double)(gx + 1);
            double m0, m1, m2;
            lead_f(pos[3 * i] - z, pos[3 * i + 1] - y, pos[3 * i + 2] - x, m0, m1, m2);
            s_m[0][0][cell] = m0;
            s_m[0][1][cell] = m1;
            s_m[0][2][cell] = m2;
            if (two) {
                lead_f(pos[3 * i + 3] - z, pos[3 * i + 4] - y, pos[3 * i + 5] - x, m0, m1, m2);
                s_m[1][0][cell] = m0;
                s_m[1][1][cell] = m1;
                s_m[1][2][cell] = m2;
            }
        }
        __syncthreads();
        if (t < nv) {
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                double g = s_m[w][0][h] * sd0 + s_m[w][1][h] * sd1 + s_m[w][2][h] * sd2;
#pragma unroll
                for (int q = 0; q < N::n; ++q) {
                    const int hn = h + (N::d[q][0] * LH + N::d[q][1]) * LH + N::d[q][2];
                    if (N::d[q][0]) g += s_m[w][0][hn] * N::d[q][0];
                    if (N::d[q][1]) g += s_m[w][1][hn] * N::d[q][1];
                    if (N::d[q][2]) g += s_m[w][2][hn] * N::d[q][2];
                }
                if (w == 0 || two) row[i + w] = g;
            }
        }
        __syncthreads();
    }
    if (t < nv) {
        const int4 ch = chunks[vox_chunk[v]];
        const double* k = layer_k + ((int64_t)b * n_layers + (ch.x - 1)) * 9;
        const double da = vox_a64[v] - (double)__int_as_float(ch.w);
        const float2 pq{(float)(k[2] * (1.0 - k[3]) * exp((k[4] + k[5]) * da)),
                        (float)(k[2] * k[3] * exp(k[5] * da))};
        const float2 av = vox_a[v];
        double* gv = rec + ((int64_t)b * rows + v) * RS;
        if constexpr (kMfma) {
            aux[(int64_t)b * rows + v] = float4{pq.x, pq.y, av.x, av.y};
        } else {
            row[LP] = __builtin_bit_cast(double, pq);
            row[LP + 1] = __builtin_bit_cast(double, av);
        }
#pragma unroll
        for (int i = 0; i < RS; i += 2)
            *reinterpret_cast<double2*>(gv + i) = double2{row[i], row[i + 1]};
    }
}

// Per (chunk, request) and lead, over the chunk's voxels u:
//   sum G_u,   sum G_u P_u,   sum G_u Q_u     (FP64, P and Q as in lead_field)
// k0 enters the ECG as k0 * sum G (constant in time); once every voxel of the
// chunk is past its upstroke (sigmoid 1), the chunk adds exactly
//   X(t) * sum G P + Y(t) * sum G Q
// (the fully saturated tiles of the fused kernels).  Fixed summation order:
// a strided pass per thread, a shuffle tree per wave, then wave order.
template <int LP, bool kMfma>
__global__ __launch_bounds__(256) void chunk_sums_kernel(
    const int4* __restrict__ chunks, const double* __restrict__ vox_a64,
    const double* __restrict__ layer_k, int n_layers, int n_leads, int64_t V,
    const double* __restrict__ rec, int n_chunks, double* __restrict__ sums)
{
    constexpr int RS = kMfma ? 16 : LP + 2;
    __shared__ double s_wave[4][3 * LP];
    const int c = blockIdx.x;
    const int b = blockIdx.y;
    const int4 ch = chunks[c];
    const double* k = layer_k + ((int64_t)b * n_layers + (ch.x - 1)) * 9;
    const double tref = (double)__int_as_float(ch.w);
    const double alpha = k[2] * (1.0 - k[3]), beta = k[2] * k[3], k45 = k[4] + k[5], k5 = k[5];
    const double* recb = rec + (int64_t)b * (V + (kMfma ? MFMA_PAD : 1)) * RS;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double* out = sums + ((int64_t)b * n_chunks + c) * 3 * LP;
    if constexpr (kMfma) {
        // 16-double rows: a wave loads 4 whole rows per instruction (lane =
        // 16 * row slot + lead), so every load is one coalesced 512-byte
        // read.  For the wave's 16 rows, lanes 0-15 compute P and lanes 16-31
        // Q in one exp (lanes 32-63 repeat them), broadcast by shuffles.
        const int col = lane & 15, slot = lane >> 4;
        const bool is_q = (lane >> 4) & 1;
        const double coef = is_q ? beta : alpha, rate = is_q ? k5 : k45;
        double sg = 0.0, sp = 0.0, sq = 0.0;
        for (int v0 = ch.y + 16 * w; v0 < ch.z; v0 += 64) {
            const int ve = v0 + col;
            const double e = ve < ch.z ? coef * exp(rate * (vox_a64[ve] - tref)) : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * r + slot;
                const int v = v0 + row;
                const double g = v < ch.z ? recb[(int64_t)v * 16 + col] : 0.0;
                const double P = __shfl(e, row, 64), Q = __shfl(e, 16 + row, 64);
                sg += g;
                sp = __builtin_fma(g, P, sp);
                sq = __builtin_fma(g, Q, sq);
            }
        }
        // combine the 4 row slots of the wave (fixed order), then the waves
        sg += __shfl_down(sg, 16, 64);
        sp += __shfl_down(sp, 16, 64);
        sq += __shfl_down(sq, 16, 64);
        sg += __shfl_down(sg, 32, 64);
        sp += __shfl_down(sp, 32, 64);
        sq += __shfl_down(sq, 32, 64);
        if (lane < LP) {
            s_wave[w][lane] = sg;
            s_wave[w][LP + lane] = sp;
            s_wave[w][2 * LP + lane] = sq;
        }
    } else {
        double part[3 * LP];
#pragma unroll
        for (int i = 0; i < 3 * LP; ++i) part[i] = 0.0;
        for (int v = ch.y + threadIdx.x; v < ch.z; v += blockDim.x) {
            const double da = vox_a64[v] - tref;
            const double P = alpha * exp(k45 * da), Q = beta * exp(k5 * da);
#pragma unroll
            for (int i = 0; i < LP; ++i)
                if (i < n_leads) {
                    const double g = recb[(int64_t)v * RS + i];
                    part[i] += g;
                    part[LP + i] = __builtin_fma(g, P, part[LP + i]);
                    part[2 * LP + i] = __builtin_fma(g, Q, part[2 * LP + i]);
                }
        }
#pragma unroll
        for (int i = 0; i < 3 * LP; ++i) {
            double x = part[i];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
            if (lane == 0) s_wave[w][i] = x;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * LP; i += blockDim.x) {
        double sum = 0.0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) sum += s_wave[ww][i];
        out[i] = sum;
    }
}

// ---------------------------------------------------------------------------
// AP tables
// ---------------------------------------------------------------------------

// Per (request b, layer l, padded step j): the repolarisation factor
//   R(t_j) = 1 - (1 + exp(-k7 (t_j - k8) + log(2^(k7/k6) - 1)))^(-k6/k7)
// of Wohlfart.h:200-201, in FP64.  For a cell activated at a the reference
// evaluates it at tau = t - a with k8 - a (simulator.cpp:154-158), and
// tau - (k8 - a) = t - k8, so the factor depends on the layer and t only.
// Thread j == 0 also writes the per-layer FP32 constants.
__global__ void ap_tables_kernel(const double* __restrict__ layer_k, int n_layers,
                                 const double* __restrict__ tgrid, int Tp, int n_req,
                                 double* __restrict__ rtab, LayerConst* __restrict__ lc)
{
    const int64_t n = (int64_t)n_req * n_layers * Tp;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(idx % Tp);
        const int64_t bl = idx / Tp;
        const double* k = layer_k + bl * 9;
        const double t = tgrid[j];
        rtab[idx] =
            1.0 - pow(1.0 + exp(-k[7] * (t - k[8]) + log(pow(2.0, k[7] / k[6]) - 1.0)), -(k[6] / k[7]));
        if (j == 0) {
            const double LOG2E = 1.4426950408889634074;
            LayerConst c;
            c.c1 = (float)(-k[1] * LOG2E);
            c.c4 = (float)(-k[4] * LOG2E);
            c.c5 = (float)(-k[5] * LOG2E);
            c.alpha = (float)(k[2] * (1.0 - k[3]));
            c.beta = (float)(k[2] * k[3]);
            c.skip = (k[1] > 0.0 && k[4] + k[5] <= 0.25 * k[1]) ? 1.f : 0.f;
            c.k0 = k[0];
            lc[bl] = c;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused AP evaluation + lead contraction
// ---------------------------------------------------------------------------

// Sigmoid bounds in the exp2 domain.  With x = c1*tau, c1 = -k1 log2(e):
//   x >= SKIP_X for every lane  -> S < 2^-64: with k4 + k5 <= k1/4 (the
//                                  LayerConst skip flag) the cell's AP is below
//                                  k2 e^-33 and the voxel is skipped;
//   x <= SAT_X for every lane   -> exp2(x) < 2^-24, 1 + exp2(x) rounds to 1,
//                                  S = rcp(1) = 1 exactly in FP32: no sigmoid;
//   x <= SAT64_X                -> 1 + 2^x rounds to 1 in FP64 as well: the
//                                  reference's sigmoid is exactly 1.
constexpr float SKIP_X = 64.0f;
constexpr float SAT_X = -24.5f;
constexpr double SAT64_X = -54.0;
constexpr double LOG2E_D = 1.4426950408889634074;
// Largest lane factor kept; with S = 0 exactly in such lanes the product
// stays 0 instead of 0 * inf.
constexpr double XMAX = 1e30;

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// tau = t - a from hi/lo FP32 pairs.
__device__ __forceinline__ float tau_of(float th, float tl, float ah, float al)
{
    return (th - ah) + (tl - al);
}

// First voxel index in [lo, hi) whose predicate is false, for a predicate
// that is true on a prefix of the a-sorted range.  Wave-cooperative 64-ary
// search: each lane tests one probe, the ballot's count of true probes
// narrows the range to one stride; ranges up to 4096 take two rounds.  The
// result is wave-uniform.
template <class P>
__device__ __forceinline__ int first_false(int lo, int hi, P pred)
{
    const int lane = threadIdx.x & 63;
    while (hi - lo > 0) {
        const int n = hi - lo;
        const int stride = (n + 63) >> 6;
        const int v = lo + lane * stride;
        const bool t = v < hi && pred(v);
        const int c = __popcll(__ballot(t));  // probes lo, lo+stride, ... true for the first c
        if (stride == 1) return lo + c;
        // the first false lies in (lo + (c-1)*stride, lo + c*stride]
        const int nlo = c == 0 ? lo : lo + (c - 1) * stride + 1;
        const int nhi = lo + c * stride < hi ? lo + c * stride + 1 : hi;
        if (c == 0) return lo;
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

// Block = 4 waves; wave w of block (g, c, b) takes time tile 4g + w of
// chunk c of request b.  Lane = time step.  The cell AP of voxel u at t is
//   V = S(tau) * (P_u X(t) + Q_u Y(t)) + k0,  tau = t - a_u,
//   X(t) = e^{-(k4+k5)(t - t_ref)} R(t),  Y(t) = e^{-k5 (t - t_ref)} R(t),
// which is Wohlfart.h:197-202 with the exponentials split into a per-voxel
// factor (P, Q from lead_field) and a per-lane factor computed once per wave
// in FP64.  Voxels of a chunk are one layer, sorted by activation time, so
// for a tile they split into
//   [v0, s1)  sigmoid saturated in every lane: V = P X + Q Y (2 VALU ops),
//   [s1, s2)  general case,
//   [s2, v1)  exactly zero contribution (skipped),
// found by binary search on the tile's first and last lane (tau is monotone
// in t and in a).  When c1 >= 0 (no decaying sigmoid) the whole chunk takes
// the general path.  Lead sums are FP64 FMAs with the G row in SGPRs.
template <int LP>
__global__ __launch_bounds__(256) void ecg_fused_kernel(
    const float2* __restrict__ vox_a, const double* __restrict__ rec,
    const double* __restrict__ gsum,
    const int4* __restrict__ chunks, const double* __restrict__ vox_a64, const LayerConst* __restrict__ lc,
    const double* __restrict__ layer_k, const double* __restrict__ rtab,
    const double* __restrict__ tgrid, const float2* __restrict__ thl,
    double* __restrict__ partial, int64_t V, int n_layers, int n_chunks, int Tp, int n_tiles,
    int n_groups)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x / n_groups;
    const int tile = (blockIdx.x - c * n_groups) * 4 + w;
    const int b = blockIdx.y;
    if (tile >= n_tiles) return;

    const int4 ch = chunks[c];
    const int l = ch.x;  // layer, 1-based
    const int64_t bl = (int64_t)b * n_layers + (l - 1);
    const LayerConst kc = lc[bl];
    const int j = tile * 64 + lane;
    const double* sums = gsum + ((int64_t)b * n_chunks + c) * 3 * LP;
    double* out = partial + (((int64_t)b * n_chunks + c) * LP) * Tp + j;
    const double* k = layer_k + bl * 9;
    // per-lane time factors in FP64
    const double R = rtab[bl * Tp + j];
    const double dtr = tgrid[j] - (double)__int_as_float(ch.w);
    const double X64 = exp(-(k[4] + k[5]) * dtr) * R, Y64 = exp(-k[5] * dtr) * R;
    if (-k[1] * LOG2E_D * (tgrid[tile * 64] - vox_a64[ch.z - 1]) <= SAT64_X) {
        // every voxel of the chunk has a sigmoid of exactly 1 in every lane:
        // the chunk adds X(t) sum G P + Y(t) sum G Q (+ k0 sum G)
#pragma unroll
        for (int i = 0; i < LP; ++i)
            out[(int64_t)i * Tp] = __builtin_fma(X64, sums[LP + i], __builtin_fma(Y64, sums[2 * LP + i], kc.k0 * sums[i]));
        return;
    }
    const float2 tt = thl[j];
    const float2 tf = thl[tile * 64];
    const float2 tl = thl[tile * 64 + 63];
    // FP32 copies for the voxel loops
    const float X = (float)fmin(X64, XMAX);
    const float Y = (float)fmin(Y64, XMAX);
    constexpr int RS = LP + 2;
    const double* Rb = rec + (int64_t)b * (V + 1) * RS;

    double acc[LP];
#pragma unroll
    for (int i = 0; i < LP; ++i) acc[i] = 0.0;

    int s1 = ch.y, s2 = ch.z;
    if (kc.c1 < 0.0f) {
        // saturated prefix: max over lanes of c1*tau <= SAT_X (max at the first lane)
        s1 = first_false(ch.y, ch.z, [&](int v) {
            const float2 a = vox_a[v];
            return kc.c1 * tau_of(tf.x, tf.y, a.x, a.y) <= SAT_X;
        });
        // non-zero range ends where min over lanes of c1*tau >= SKIP_X (min at the last lane)
        if (kc.skip != 0.0f)
            s2 = first_false(s1, ch.z, [&](int v) {
                const float2 a = vox_a[v];
                return kc.c1 * tau_of(tl.x, tl.y, a.x, a.y) < SKIP_X;
            });
    }

    // Both loops keep the next voxel's record in flight while the current
    // one is consumed: one group of wave-uniform scalar loads per voxel,
    // issued before the FP64 FMAs of the previous voxel (the one-ahead load
    // of the last voxel reads the next record, which always exists).
    if (ch.y < s1) {
        const double* rp = Rb + (int64_t)ch.y * RS;
        double rn[RS];
#pragma unroll
        for (int i = 0; i < RS; ++i) rn[i] = rp[i];
        for (int v = ch.y; v < s1; ++v) {
            double r[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) r[i] = rn[i];
            rp += RS;
#pragma unroll
            for (int i = 0; i < RS; ++i) rn[i] = rp[i];
            const float2 pq = __builtin_bit_cast(float2, r[LP]);
            const float ap = __builtin_fmaf(pq.x, X, pq.y * Y);
            const double apd = (double)ap;
#pragma unroll
            for (int i = 0; i < LP; ++i) acc[i] = __builtin_fma(r[i], apd, acc[i]);
        }
    }
    if (s1 < s2) {
        const double* rp = Rb + (int64_t)s1 * RS;
        double rn[RS];
#pragma unroll
        for (int i = 0; i < RS; ++i) rn[i] = rp[i];
        for (int v = s1; v < s2; ++v) {
            double r[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) r[i] = rn[i];
            rp += RS;
#pragma unroll
            for (int i = 0; i < RS; ++i) rn[i] = rp[i];
            const float2 pq = __builtin_bit_cast(float2, r[LP]);
            const float2 a = __builtin_bit_cast(float2, r[LP + 1]);
            const float tau = tau_of(tt.x, tt.y, a.x, a.y);
            const float e1 = fast_exp2(kc.c1 * tau);
            const float ap = __builtin_fmaf(pq.x, X, pq.y * Y) * fast_rcp(1.0f + e1);
            const double apd = (double)ap;
#pragma unroll
            for (int i = 0; i < LP; ++i) acc[i] = __builtin_fma(r[i], apd, acc[i]);
        }
    }

    // k0 is constant in time: sum_u G_u (ap_u + k0) = sum_u G_u ap_u + k0 * sum_u G_u.
#pragma unroll
    for (int i = 0; i < LP; ++i) out[(int64_t)i * Tp] = __builtin_fma(kc.k0, sums[i], acc[i]);
}

// ---------------------------------------------------------------------------
// Fused AP evaluation + lead contraction on the matrix cores
// ---------------------------------------------------------------------------

typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));

// Same computation as ecg_fused_kernel with the lead contraction on
// v_mfma_f64_16x16x4_f64: D[lead i][time j] += sum_k G[voxel k][lead i] * V[voxel k][time j].
// gfx950 lane layout (checked by a one-hot probe): A(i,k) in lane 16k+i,
// B(k,j) in lane 16k+j, D(i,j) in lane 16(i%4)+j, register i/4.
// Block = up to 16 waves of one (chunk, request); a wave takes 64-step time
// tiles w, w+nw, ... and, per tile, splits the a-sorted chunk like the VALU
// kernel.  Per group of 4 voxels: one coalesced 512-byte load of the four G
// rows (A), one float4 {P, Q, a} per lane (voxel k = lane/16), 4 AP values per
// lane (time blocks m = 0..3, j = lane%16) and 4 MFMAs into 4 FP64
// accumulator tiles.  Groups that reach past the chunk end are masked with
// selects; the 4 spare rows per request keep their loads in bounds.
template <int LP>
__global__ __launch_bounds__(1024) void ecg_fused_mfma_kernel(
    const float2* __restrict__ vox_a, const double* __restrict__ G16,
    const float4* __restrict__ aux, const double* __restrict__ gsum,
    const int4* __restrict__ chunks, const double* __restrict__ vox_a64, const LayerConst* __restrict__ lc,
    const double* __restrict__ layer_k, const double* __restrict__ rtab,
    const double* __restrict__ tgrid, const float2* __restrict__ thl,
    double* __restrict__ partial, int64_t V, int n_layers, int n_chunks, int Tp, int n_tiles,
    int n_leads)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int c = blockIdx.x;
    const int b = blockIdx.y;
    const int4 ch = chunks[c];
    const int l = ch.x;
    const int64_t bl = (int64_t)b * n_layers + (l - 1);
    const LayerConst kc = lc[bl];
    const double* kk = layer_k + bl * 9;
    const double k45 = kk[4] + kk[5], k5 = kk[5];
    const double tref = (double)__int_as_float(ch.w);
    const int jl = lane & 15;   // time column
    const int kv = lane >> 4;   // voxel slot of the B operand / lead group of D
    const double* Gb = G16 + (int64_t)b * (V + MFMA_PAD) * 16;
    const float4* Ab = aux + (int64_t)b * (V + MFMA_PAD);
    const double* sums = gsum + ((int64_t)b * n_chunks + c) * 3 * LP;
    double* out = partial + ((int64_t)b * n_chunks + c) * LP * Tp;
    const double c1d = -kk[1] * LOG2E_D;
    const double a_last = vox_a64[ch.z - 1];

    for (int tile = w; tile < n_tiles; tile += nw) {
        if (c1d * (tgrid[tile * 64] - a_last) <= SAT64_X) {
            // every voxel's sigmoid is exactly 1 in every lane: the chunk adds
            // X(t) sum G P + Y(t) sum G Q (+ k0 sum G), one time step per lane
            const int j = tile * 64 + lane;
            const double R = rtab[bl * Tp + j];
            const double dtr = tgrid[j] - tref;
            const double X64 = exp(-k45 * dtr) * R, Y64 = exp(-k5 * dtr) * R;
#pragma unroll
            for (int i = 0; i < LP; ++i)
                if (i < n_leads)
                    out[(int64_t)i * Tp + j] =
                        __builtin_fma(X64, sums[LP + i], __builtin_fma(Y64, sums[2 * LP + i], kc.k0 * sums[i]));
            continue;
        }
        float X[4], Y[4], TH[4], TL[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int j = tile * 64 + 16 * m + jl;
            const double R = rtab[bl * Tp + j];
            const double dtr = tgrid[j] - tref;
            X[m] = (float)fmin(exp(-k45 * dtr) * R, XMAX);
            Y[m] = (float)fmin(exp(-k5 * dtr) * R, XMAX);
            const float2 t2 = thl[j];
            TH[m] = t2.x;
            TL[m] = t2.y;
        }
        const float2 tf = thl[tile * 64];
        const float2 tl = thl[tile * 64 + 63];
        int s1 = ch.y, s2 = ch.z;
        if (kc.c1 < 0.0f) {
            s1 = first_false(ch.y, ch.z, [&](int v) {
                const float2 a = vox_a[v];
                return kc.c1 * tau_of(tf.x, tf.y, a.x, a.y) <= SAT_X;
            });
            if (kc.skip != 0.0f)
                s2 = first_false(s1, ch.z, [&](int v) {
                    const float2 a = vox_a[v];
                    return kc.c1 * tau_of(tl.x, tl.y, a.x, a.y) < SKIP_X;
                });
        }
        mfma_f64x4 acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = mfma_f64x4{0.0, 0.0, 0.0, 0.0};

        // Groups of 4 voxels; the next group's A rows and {P, Q, a} are loaded
        // while the current group's MFMAs run (vector loads complete in order).
        // Every group starts below ch.z <= V, so the look-ahead group starts at
        // most at V + 3 and its rows stay below V + 7 < V + MFMA_PAD.
        int vg = ch.y;
        double a_n = Gb[(int64_t)vg * 16 + lane];
        float4 q_n = Ab[vg + kv];
        // groups of 4 voxels that are saturated in every lane of the tile
        for (; vg + 4 <= s1; vg += 4) {
            const double a_op = a_n;
            const float4 q = q_n;
            a_n = Gb[(int64_t)(vg + 4) * 16 + lane];
            q_n = Ab[vg + 4 + kv];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double bv = (double)__builtin_fmaf(q.x, X[m], q.y * Y[m]);
                acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_op, bv, acc[m], 0, 0, 0);
            }
        }
        // remaining groups up to the last non-zero voxel: general path, masked at the chunk end
        for (; vg < s2; vg += 4) {
            const bool ok = vg + kv < ch.z;
            const double a_op = ok ? a_n : 0.0;
            const float4 q = q_n;
            a_n = Gb[(int64_t)(vg + 4) * 16 + lane];
            q_n = Ab[vg + 4 + kv];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const float tau = tau_of(TH[m], TL[m], q.z, q.w);
                const float e1 = fast_exp2(kc.c1 * tau);
                const float ap = __builtin_fmaf(q.x, X[m], q.y * Y[m]) * fast_rcp(1.0f + e1);
                const double bv = ok ? (double)ap : 0.0;
                acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_op, bv, acc[m], 0, 0, 0);
            }
        }
        // D(i, j) sits in lane 16(i%4)+j, register i/4: lead i = 4r + lane/16
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 4 * r + kv;
            if (i < n_leads && i < LP) {
                const double g = sums[i];
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    out[(int64_t)i * Tp + tile * 64 + 16 * m + jl] = __builtin_fma(kc.k0, g, acc[m][r]);
            }
        }
    }
}

// ecg[b][i][t] = sum over chunks c (in order) of partial[b][c][i][t] for
// t < Tn, and 0 for Tn <= t < T: samples past the last iteration of the
// reference's neighbour loop (simulator.cpp:519) receive no dipole.
__global__ void ecg_reduce_kernel(const double* __restrict__ partial, double* __restrict__ ecg,
                                  int n_req, int n_chunks, int LP, int n_leads, int T, int Tn, int Tp)
{
    const int64_t n = (int64_t)n_req * n_leads * T;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(idx % T);
        const int64_t bi = idx / T;
        const int i = (int)(bi % n_leads);
        const int b = (int)(bi / n_leads);
        double s = 0.0;
        if (t < Tn) {
            const double* p = partial + ((int64_t)b * n_chunks * LP + i) * Tp + t;
            for (int c = 0; c < n_chunks; ++c) s += p[(int64_t)c * LP * Tp];
        }
        ecg[idx] = s;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int grid_for(int64_t n, int threads, int cap = 65536)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

hipError_t launch_ca_init(bool is_int, const uint8_t* layer, const uint8_t* seed, void* dist,
                          uint8_t* active, int n0, int n1, int n2, int nb1, int nb2, hipStream_t s)
{
    const int64_t n = (int64_t)n0 * n1 * n2;
    if (is_int)
        hipLaunchKernelGGL(ca_init_kernel<int>, dim3(grid_for(n, 256)), dim3(256), 0, s, layer, seed,
                           (int*)dist, active, n0, n1, n2, nb1, nb2);
    else
        hipLaunchKernelGGL(ca_init_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, layer,
                           seed, (double*)dist, active, n0, n1, n2, nb1, nb2);
    return hipGetLastError();
}

hipError_t launch_ca_sweep(bool is_int, const uint8_t* layer, const uint8_t* delay,
                           const double* wtab, int wcols, void* dist, int n0, int n1, int n2,
                           int nb0, int nb1, int nb2, const uint8_t* act_in, uint8_t* act_out,
                           int* any_marked, hipStream_t s)
{
    const int nbricks = nb0 * nb1 * nb2;
    if (is_int)
        hipLaunchKernelGGL(ca_sweep_kernel<true>, dim3(nbricks), dim3(CA_THREADS), 0, s, layer,
                           delay, wtab, wcols, (int*)dist, n0, n1, n2, nb0, nb1, nb2, act_in,
                           act_out, any_marked);
    else
        hipLaunchKernelGGL(ca_sweep_kernel<false>, dim3(nbricks), dim3(CA_THREADS), 0, s, layer,
                           delay, wtab, wcols, (double*)dist, n0, n1, n2, nb0, nb1, nb2, act_in,
                           act_out, any_marked);
    return hipGetLastError();
}

hipError_t launch_ca_finalize(bool is_int, const void* dist, double* act, int64_t n, hipStream_t s)
{
    if (is_int)
        hipLaunchKernelGGL(ca_finalize_kernel<int>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                           (const int*)dist, act, n);
    else
        hipLaunchKernelGGL(ca_finalize_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                           (const double*)dist, act, n);
    return hipGetLastError();
}

template <int LP, bool M, int NH>
static void lf_launch_nh(const LeadFieldArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((lead_field_kernel<LP, M, NH>), dim3(a.n_bricks, a.n_req), dim3(512), 0, s,
                       a.grid_layer, a.n0, a.n1, a.n2, a.bricks, a.brick_off, a.brick_vox,
                       a.halo_off, a.halo_cells, a.vox_chunk, a.vox_a64, a.vox_a, a.chunks,
                       a.lead_pos, a.n_leads, a.layer_k, a.n_layers, a.V, a.rec, a.aux);
}

template <int LP, bool M>
static void lf_launch_t(const LeadFieldArgs& a, hipStream_t s)
{
    switch (a.nhood) {
    case 0: lf_launch_nh<LP, M, 0>(a, s); break;
    case 1: lf_launch_nh<LP, M, 1>(a, s); break;
    case 2: lf_launch_nh<LP, M, 2>(a, s); break;
    default: lf_launch_nh<LP, M, 3>(a, s); break;
    }
    hipLaunchKernelGGL((chunk_sums_kernel<LP, M>), dim3(a.n_chunks, a.n_req), dim3(256), 0, s,
                       a.chunks, a.vox_a64, a.layer_k, a.n_layers, a.n_leads, a.V, a.rec, a.n_chunks,
                       a.gsum);
}

template <int LP>
static void lf_launch(const LeadFieldArgs& a, hipStream_t s)
{
    if (a.mfma) lf_launch_t<LP, true>(a, s);
    else lf_launch_t<LP, false>(a, s);
}

hipError_t launch_lead_field(const LeadFieldArgs& a, hipStream_t s)
{
    switch (a.LP) {
    case 4: lf_launch<4>(a, s); break;
    case 8: lf_launch<8>(a, s); break;
    case 12: lf_launch<12>(a, s); break;
    case 16: lf_launch<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ap_tables(const double* layer_k, int n_layers, const double* tgrid, int Tp,
                            int n_req, double* rtab, LayerConst* lc, hipStream_t s)
{
    const int64_t n = (int64_t)n_req * n_layers * Tp;
    hipLaunchKernelGGL(ap_tables_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, layer_k,
                       n_layers, tgrid, Tp, n_req, rtab, lc);
    return hipGetLastError();
}

template <int LP>
static void fused_launch(const FusedArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(ecg_fused_kernel<LP>, dim3(a.n_chunks * a.n_groups, a.n_req), dim3(256), 0,
                       s, a.vox_a, a.rec, a.gsum, a.chunks, a.vox_a64, a.lc, a.layer_k, a.rtab, a.tgrid,
                       a.thl, a.partial, a.V, a.n_layers, a.n_chunks, a.Tp, a.n_tiles, a.n_groups);
}

template <int LP>
static void fused_mfma_launch(const FusedArgs& a, hipStream_t s)
{
    const int nw = a.n_tiles < 16 ? a.n_tiles : 16;
    hipLaunchKernelGGL(ecg_fused_mfma_kernel<LP>, dim3(a.n_chunks, a.n_req), dim3(64 * nw), 0, s,
                       a.vox_a, a.rec, a.aux, a.gsum, a.chunks, a.vox_a64, a.lc, a.layer_k, a.rtab, a.tgrid,
                       a.thl, a.partial, a.V, a.n_layers, a.n_chunks, a.Tp, a.n_tiles, a.n_leads);
}

hipError_t launch_ecg_fused(const FusedArgs& a, hipStream_t s)
{
    if (a.mfma) {
        switch (a.LP) {
        case 4: fused_mfma_launch<4>(a, s); break;
        case 8: fused_mfma_launch<8>(a, s); break;
        case 12: fused_mfma_launch<12>(a, s); break;
        case 16: fused_mfma_launch<16>(a, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.LP) {
    case 4: fused_launch<4>(a, s); break;
    case 8: fused_launch<8>(a, s); break;
    case 12: fused_launch<12>(a, s); break;
    case 16: fused_launch<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ecg_reduce(const double* partial, double* ecg, int n_req, int n_chunks, int LP,
                             int n_leads, int T, int Tn, int Tp, hipStream_t s)
{
    const int64_t n = (int64_t)n_req * n_leads * T;
    hipLaunchKernelGGL(ecg_reduce_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, partial, ecg,
                       n_req, n_chunks, LP, n_leads, T, Tn, Tp);
    return hipGetLastError();
}

}  // namespace ekgk
