// ekg_ecg.hip -- CDNA4 (gfx950) kernels of the ECG path.
//
// A launch group holds requests of any cached hearts, lead counts and time
// grids (ekg_kernels.hpp: GroupArgs).  Work items are enumerated per request
// from exclusive prefix sums, so one launch of each kernel covers the group:
//
//   ap_tables     per (request, layer, step): X_l(t), Y_l(t) -- the
//                 time factors of the Wohlfart AP (Wohlfart.h:195-203) with
//                 the repolarisation factor R_l(t) -- in FP64, and the
//                 per-layer FP32 constants
//   lead_field    per (request, 8^3 brick): lead weights G_i(u) in FP64
//                 (the reference's dipole sum regrouped by cell,
//                 simulator.cpp:509-537), per-voxel AP factors P_u, Q_u
//   chunk_sums    per (request, chunk): sum G, sum G P', sum G Q' in FP64
//                 and the chunk's tile bounds [z, f): before z every voxel
//                 is far before its upstroke, from f on every sigmoid is 1
//   layer_prefix  per (request, layer, tile): prefix sums of the saturated
//                 chunks and the range of chunks still in transition
//   ecg_fused     per (request, chunk), transition tiles only: AP
//                 evaluation fused with the lead contraction (FP32 AP, FP64
//                 accumulation on the matrix cores or the vector ALUs)
//   ecg_reduce    per (request, lead, step): closed form of the saturated
//                 chunks + fixed-order FP64 sum of the transition partials
//
// Every launch is checked on the host against the buffer extents it assumes
// (ekg_engine.cpp) before it is issued.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ekg_kernels.hpp"

namespace ekgk {

// Request owning work item x: the largest r with start[r] <= x (requests
// without work items share their start with the next one).  A batch whose
// requests all hold the same number of items is answered by one guess and one
// check (two dependent loads instead of a log2(B)-deep chain); any other
// batch falls through to the binary search.
__device__ __forceinline__ int find_req(const int64_t* __restrict__ start, int B, int64_t x)
{
    if (B <= 1) return 0;
    const int64_t s0 = start[0], w = start[1] - s0, d = x - s0;
    if (w > 0 && d >= 0) {
        const int64_t g = (d < 0x7fffffff && w < 0x7fffffff) ? (int64_t)((uint32_t)d / (uint32_t)w) : d / w;
        if (g < B) {
            const int r = (int)g;
            if (start[r] <= x && (r + 1 == B || start[r + 1] > x)) return r;
        }
    }
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Sigmoid bounds in the exp2 domain.  With x = c1*tau, c1 = -k1 log2(e):
//   x >= SKIP_X for every lane  -> S < 2^-48: with k4 + k5 <= k1/4 (the
//                                  LayerConst skip flag) the cell's AP is below
//                                  k2 2^-36 < k2 e^-24 and the voxel is skipped;
//   x <= SAT_X for every lane   -> exp2(x) < 2^-24, 1 + exp2(x) rounds to 1,
//                                  S = rcp(1) = 1 exactly in FP32: no sigmoid;
//   x <= SATC_X (FP64 test on   -> every voxel's S is within 2^-24.5 of 1, as
//   the latest voxel of a          in the FP32 AP: the chunk (or group of 16
//   chunk or group at a tile's     voxels in a quad) takes the exact FP64
//   or quad's first step)          closed form X_l sum G P' + Y_l sum G Q'
//                                  instead of FP32 evaluation.
constexpr float SKIP_X = 48.0f;
constexpr float SAT_X = -24.5f;
constexpr double SATC_X = -24.5;
constexpr double LOG2E_D = 1.4426950408889634074;
// Largest FP32 lane factor kept; with S = 0 exactly in such lanes the
// product stays 0 instead of 0 * inf.
constexpr double XMAX = 1e30;

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// tau = t - a from hi/lo FP32 pairs.
__device__ __forceinline__ float tau_of(float th, float tl, float ah, float al)
{
    return (th - ah) + (tl - al);
}

// Offset of (lead i, step j) in a chunk's partial rows (LP x Tp doubles),
// stored tile-major: for each bound tile of 2^bs steps its LP rows of 2^bs
// steps, so that a tile's rows are one contiguous run (written and read in
// whole cache lines, whichever tiles a chunk is in transition for).
__device__ __forceinline__ int64_t prow(int i, int j, int lp, int bs)
{
    return ((int64_t)((j >> bs) * lp + i) << bs) + (j & ((1 << bs) - 1));
}

// First index in [lo, hi) whose predicate is false, for a predicate that is
// true on a prefix of the range.  Wave-cooperative 64-ary search: each lane
// tests one probe, the ballot's count of true probes narrows the range to one
// stride; ranges up to 4096 take two rounds.  The result is wave-uniform.
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
        if (c == 0) return lo;
        const int nlo = lo + (c - 1) * stride + 1;
        const int nhi = lo + c * stride < hi ? lo + c * stride + 1 : hi;
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

// First index in [0, n) where a monotone (false...true) predicate holds, n if none.
template <class P>
__device__ __forceinline__ int first_true(int n, P pred)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pred(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// AP tables
// ---------------------------------------------------------------------------

// Per (request, layer l, padded step j), with t0 the layer's reference time:
//   R(t) = 1 - (1 + exp(-k7 (t - k8) + log(2^(k7/k6) - 1)))^(-k6/k7)
//   X_l(t) = e^{-(k4+k5)(t - t0)} R(t),   Y_l(t) = e^{-k5 (t - t0)} R(t)
// R is the repolarisation factor of Wohlfart.h:200-201.  A cell activated at
// a is evaluated at tau = t - a with k8 - a (simulator.cpp:154-158), and
// tau - (k8 - a) = t - k8: R depends on the layer and t only.  Step j == 0
// also writes the per-layer FP32 constants.
__global__ void ap_tables_kernel(const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs,
                                 int B, const int64_t* __restrict__ xy_start, int64_t n_xy,
                                 const double* __restrict__ layer_k, const double* __restrict__ tgrid,
                                 double2* __restrict__ xy, LayerConst* __restrict__ lc)
{
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n_xy;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = find_req(xy_start, B, idx);
        const ReqDev R = reqs[r];
        const int64_t loc = idx - xy_start[r];
        const int l = (int)(loc / R.Tp) + 1;
        const int j = (int)(loc % R.Tp);
        const HeartDev& H = hearts[R.heart];
        const double* k = layer_k + R.k_off + 9 * (l - 1);
        const double t = tgrid[R.t_off + j];
        const double t0 = l <= H.max_layer ? H.layer_t0[l] : 0.0;
        const double Rt =
            1.0 - pow(1.0 + exp(-k[7] * (t - k[8]) + log(pow(2.0, k[7] / k[6]) - 1.0)), -(k[6] / k[7]));
        xy[R.xy_off + loc] = double2{exp(-(k[4] + k[5]) * (t - t0)) * Rt, exp(-k[5] * (t - t0)) * Rt};
        if (j == 0) {
            LayerConst c;
            c.c1 = (float)(-k[1] * LOG2E_D);
            c.c4 = (float)(-k[4] * LOG2E_D);
            c.c5 = (float)(-k[5] * LOG2E_D);
            c.alpha = (float)(k[2] * (1.0 - k[3]));
            c.beta = (float)(k[2] * k[3]);
            c.skip = (k[1] > 0.0 && k[4] + k[5] <= 0.25 * k[1]) ? 1.f : 0.f;
            c.k0 = k[0];
            lc[R.lay_off + l - 1] = c;
        }
    }
}

// ---------------------------------------------------------------------------
// Lead field
// ---------------------------------------------------------------------------

// f(e) = e / |e|^3: the lead weight m of simulator.cpp:531-533 for
// e = p - (u+1).  Computed in FP32 (the weights are kept in FP32): e from the
// electrode's FP32 hi/lo split, so it is within an ulp of the FP64 difference,
// then the hardware reciprocal square root with one Newton step.
__device__ __forceinline__ void lead_f32(float e0, float e1, float e2, float& m0, float& m1, float& m2)
{
    float l = e0 * e0;
    l = __builtin_fmaf(e1, e1, l);
    l = __builtin_fmaf(e2, e2, l);
    float r = __builtin_amdgcn_rsqf(l);
    r = r * __builtin_fmaf(-0.5f * l * r, r, 1.5f);
    const float s = r * r * r;
    m0 = e0 * s;
    m1 = e1 * s;
    m2 = e2 * s;
}

// s = 1/|e|^3 in FP32 with the hardware reciprocal square root alone.  With
// e in FP32, the lead weight's dipole term s (d . e) is an exact FP64 product
// (24 + at most 29 significant bits), so a cell's weight enters every G row
// that uses it with the same value.
__device__ __forceinline__ float lead_s32(float e0, float e1, float e2)
{
    float l = e0 * e0;
    l = __builtin_fmaf(e1, e1, l);
    l = __builtin_fmaf(e2, e2, l);
    const float r = __builtin_amdgcn_rsqf(l);
    return r * r * r;
}

constexpr int LB = 8;                        // lead-field brick edge
constexpr int LH = LB + 2;                   // with halo
constexpr int LN = LH * LH * LH;             // halo cells

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

// One block of 512 threads per (request, 8^3 brick holding myocardium).
// For WPH (2 or 4, autotuned) leads at a time the lead weight m_i of every
// myocardial cell of the brick and its 1-voxel halo (a compact per-brick list,
// one thread per cell looping over the leads) is computed in FP32 and kept in
// LDS (zero outside the myocardium); thread t < n then gathers, in FP64, for
// the brick's t-th myocardial voxel u
//   G_i(u) = sum_{d in N} M_i(u+d) . d + m_i(u) . s_u,
//   M_i = m_i on myocardium and 0 elsewhere,  s_u = sum of those d,
// which is sum_{d in N, u+d myo} (m_i(u+d) + m_i(u)) . d -- the coefficient of
// V_u(t) in the reference's dipole sum regrouped by cell (simulator.cpp:509-537).
// The FP32 weights change the ECG by a few 1e-7 relative (an error in m_i(w)
// enters only through V differences across the dipoles at w); G itself is
// summed and kept in FP64, since its own rounding multiplies V_u directly and
// would cost up to 3e-6 on small hearts.  The voxel's record (row g_off + v of
// the request) receives a 16-double G row (zero-padded) and
// aux = {P_u, Q_u, a_hi, a_lo} with P_u = k2(1-k3) e^{(k4+k5)(a_u - t_ref)},
// Q_u = k2 k3 e^{k5 (a_u - t_ref)}, t_ref the reference time of the voxel's chunk.
template <int LP, int NH, int WPH>
__global__ __launch_bounds__(512) void lead_field_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ lf_start, const double* __restrict__ lead_pos,
    const double* __restrict__ layer_k, double* __restrict__ rec, float4* __restrict__ aux)
{
    using N = Nhood<NH>;
    // Lead weights use the halo box with padded row and plane strides (RS, PS
    // cells): with 16-byte cells, a 10-cell row stride puts lanes of adjacent
    // rows on the same LDS banks.
    constexpr int RS = 12, PS = RS * LH, MLN = PS * LH;
    constexpr int HM = PS + RS + 1;       // largest backward neighbour offset
    constexpr int ROWS = WPH * MLN / 8;   // G rows of 128 bytes that fit in the lead-weight area
    __shared__ uint8_t s_myo[LN];
    __shared__ uint16_t s_cell[LN];
    __shared__ float4 s_m[WPH][MLN];      // {m_x, m_y, m_z, 0}: one 16-byte read per neighbour
    __shared__ float s_p[16][2][3];       // electrode {hi, lo} per component
    __shared__ int s_v[LB * LB * LB];     // sorted voxel index of each brick voxel
    const int r = find_req(lf_start, B, blockIdx.x);
    const ReqDev R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int br = (int)(blockIdx.x - lf_start[r]);
    const int t = threadIdx.x;
    const int4 bi = H.bricks[br];
    const int e0 = H.brick_off[br], nv = H.brick_off[br + 1] - e0;
    const int c0 = H.halo_off[br], nc = H.halo_off[br + 1] - c0;
    const int n_leads = R.L;
    for (int q = t; q < LN; q += blockDim.x) s_myo[q] = 0;
    for (int q = t; q < MLN; q += blockDim.x) {
#pragma unroll
        for (int w = 0; w < WPH; ++w) s_m[w][q] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    if (t < 3 * n_leads) {
        const double p = lead_pos[R.pos_off + t];
        const float hi = (float)p;
        s_p[t / 3][0][t % 3] = hi;
        s_p[t / 3][1][t % 3] = (float)(p - (double)hi);
    }
    __syncthreads();
    for (int q = t; q < nc; q += blockDim.x) {
        const int c = H.halo_cells[c0 + q];
        s_myo[c] = 1;
        s_cell[q] = (uint16_t)c;
    }
    __syncthreads();
    int hs = HM, v = -1;
    double sd0 = 0.0, sd1 = 0.0, sd2 = 0.0;
    if (t < nv) {
        const int2 bv = H.brick_vox[e0 + t];
        v = bv.x;
        s_v[t] = v;
        const int lz = bv.y >> 6, ly = (bv.y >> 3) & 7, lx = bv.y & 7;
        const int h = ((lz + 1) * LH + (ly + 1)) * LH + (lx + 1);
        hs = (lz + 1) * PS + (ly + 1) * RS + (lx + 1);
        int i0 = 0, i1 = 0, i2 = 0;
#pragma unroll
        for (int q = 0; q < N::n; ++q) {
            const int f = s_myo[h + (N::d[q][0] * LH + N::d[q][1]) * LH + N::d[q][2]];
            i0 += f * N::d[q][0];
            i1 += f * N::d[q][1];
            i2 += f * N::d[q][2];
        }
        sd0 = (double)i0;
        sd1 = (double)i1;
        sd2 = (double)i2;
    }
    double row[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) row[i] = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < LP; i0 += WPH) {
        if (i0 >= n_leads) break;
        const int nw = n_leads - i0 < WPH ? n_leads - i0 : WPH;
        for (int q = t; q < nc; q += blockDim.x) {
            const int cell = s_cell[q];
            const int cz = cell / (LH * LH), cy = (cell / LH) % LH, cx = cell % LH;
            // grid coordinate + 1 of the cell (the reference's voxel centre)
            const float fz = (float)(bi.x + cz), fy = (float)(bi.y + cy), fx = (float)(bi.z + cx);
#pragma unroll
            for (int w = 0; w < WPH; ++w) {
                if (w < nw) {
                    const float(*p)[3] = s_p[i0 + w];
                    float m0, m1, m2;
                    lead_f32((p[0][0] - fz) + p[1][0], (p[0][1] - fy) + p[1][1], (p[0][2] - fx) + p[1][2],
                             m0, m1, m2);
                    s_m[w][cz * PS + cy * RS + cx] = float4{m0, m1, m2, 0.0f};
                }
            }
        }
        __syncthreads();
        if (t < nv) {
#pragma unroll
            for (int w = 0; w < WPH; ++w) {
                if (w < nw) {
                    // non-negative constant offsets from the lowest neighbour
                    const float4* mb = &s_m[w][hs - HM];
                    const float4 mu = mb[HM];
                    double g = (double)mu.x * sd0 + (double)mu.y * sd1 + (double)mu.z * sd2;
#pragma unroll
                    for (int q = 0; q < N::n; ++q) {
                        const float4 mn = mb[HM + N::d[q][0] * PS + N::d[q][1] * RS + N::d[q][2]];
                        if (N::d[q][0]) g += (double)mn.x * N::d[q][0];
                        if (N::d[q][1]) g += (double)mn.y * N::d[q][1];
                        if (N::d[q][2]) g += (double)mn.z * N::d[q][2];
                    }
                    row[i0 + w] = g;
                }
            }
        }
        __syncthreads();
    }
    if (t < nv) {
        const int4 ch = H.chunks[H.vox_chunk[v]];
        const double* k = layer_k + R.k_off + 9 * (ch.x - 1);
        const double da = H.vox_a64[v] - (double)__int_as_float(ch.w);
        const float2 av = H.vox_a[v];
        aux[R.g_off + v] = float4{(float)(k[2] * (1.0 - k[3]) * exp((k[4] + k[5]) * da)),
                                  (float)(k[2] * k[3] * exp(k[5] * da)), av.x, av.y};
    }
    // The G rows leave through LDS (the lead-weight area is free after the
    // last gather): each 8-lane group then stores one whole 128-byte row per
    // instruction, instead of every lane touching its own line with each
    // 16-byte store (the rows of a brick are scattered in sorted order).
    double2* s_rows = reinterpret_cast<double2*>(&s_m[0][0]);
    for (int base = 0; base < nv; base += ROWS) {
        const int rr = t - base;
        if (t < nv && rr >= 0 && rr < ROWS) {
#pragma unroll
            for (int i = 0; i < 8; ++i) s_rows[rr * 8 + (i ^ (rr & 7))] = double2{row[2 * i], row[2 * i + 1]};
        }
        __syncthreads();
        const int nr = nv - base < ROWS ? nv - base : ROWS;
        for (int q = t; q < nr * 8; q += blockDim.x) {
            const int r8 = q >> 3, i = q & 7;
            reinterpret_cast<double2*>(rec + (R.g_off + s_v[base + r8]) * 16)[i] = s_rows[r8 * 8 + (i ^ (r8 & 7))];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Chunk sums and tile bounds
// ---------------------------------------------------------------------------

// Per (request, chunk) and lead, over the chunk's voxels u, in FP64 with the
// layer's reference time t0:
//   sum G_u,   sum G_u P'_u,   sum G_u Q'_u,
//   P'_u = k2(1-k3) e^{(k4+k5)(a_u - t0)},  Q'_u = k2 k3 e^{k5 (a_u - t0)}.
// Once every voxel of the chunk is past its upstroke (sigmoid 1), the chunk
// adds exactly X_l(t) sum G P' + Y_l(t) sum G Q' (+ k0 sum G).  Fixed
// summation order: coalesced row passes, a shuffle tree per wave, then wave
// order.  Thread 0 also finds the chunk's tile bounds {z, f}: tiles before z
// are skipped for every voxel (the fused kernel's skip test on the earliest
// voxel at the tile's last step), tiles from f on are saturated in FP64 for
// every voxel (test on the latest voxel at the tile's first step).
template <int LP>
__global__ __launch_bounds__(256) void chunk_sums_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ ch_start, const double* __restrict__ layer_k,
    const LayerConst* __restrict__ lc, const double* __restrict__ tgrid,
    const float2* __restrict__ thl, const double* __restrict__ rec, double* __restrict__ sums,
    int2* __restrict__ bnd)
{
    __shared__ double s_wave[4][3 * LP];
    const int r = find_req(ch_start, B, blockIdx.x);
    const ReqDev R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int c = (int)(blockIdx.x - ch_start[r]);
    const int4 ch = H.chunks[c];
    const double* k = layer_k + R.k_off + 9 * (ch.x - 1);
    const double t0 = H.layer_t0[ch.x];
    const double alpha = k[2] * (1.0 - k[3]), beta = k[2] * k[3], k45 = k[4] + k[5], k5 = k[5];
    const double* recb = rec + R.g_off * 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double* out = sums + (R.c_off + c) * 3 * LP;
    // 16-double rows: a wave loads 4 whole rows per instruction (lane = 16 *
    // row slot + lead), one coalesced 512-byte read.  For the wave's 16 rows,
    // lanes 0-15 compute P' and lanes 16-31 Q' in one exp (lanes 32-63 repeat
    // them), broadcast by shuffles.
    const int col = lane & 15, slot = lane >> 4;
    const bool is_q = (lane >> 4) & 1;
    const double coef = is_q ? beta : alpha, rate = is_q ? k5 : k45;
    double sg = 0.0, sp = 0.0, sq = 0.0;
    for (int v0 = ch.y + 16 * w; v0 < ch.z; v0 += 64) {
        const int ve = v0 + col;
        const double e = ve < ch.z ? coef * exp(rate * (H.vox_a64[ve] - t0)) : 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 4 * q + slot;
            const int v = v0 + row;
            const double g = v < ch.z ? recb[(int64_t)v * 16 + col] : 0.0;
            const double P = __shfl(e, row, 64), Q = __shfl(e, 16 + row, 64);
            sg += g;
            sp = __builtin_fma(g, P, sp);
            sq = __builtin_fma(g, Q, sq);
        }
    }
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
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * LP; i += blockDim.x) {
        double sum = 0.0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) sum += s_wave[ww][i];
        out[i] = sum;
    }
    if (threadIdx.x == 0) {
        const LayerConst kc = lc[R.lay_off + ch.x - 1];
        const double c1d = -k[1] * LOG2E_D;
        const double a_max = H.vox_a64[ch.z - 1];
        const float2 a_min = H.vox_a[ch.y];
        const double* tg = tgrid + R.t_off;
        const float2* th = thl + R.t_off;
        const int f = c1d < 0.0 ? first_true(R.n_tiles, [&](int tile) {
            return c1d * (tg[tile * 64] - a_max) <= SATC_X;
        }) : R.n_tiles;
        const int z = (kc.c1 < 0.0f && kc.skip != 0.0f) ? first_true(R.n_tiles, [&](int tile) {
            const float2 tl = th[tile * 64 + 63];
            return kc.c1 * tau_of(tl.x, tl.y, a_min.x, a_min.y) < SKIP_X;
        }) : 0;
        bnd[R.c_off + c] = int2{z < f ? z : f, f};
    }
}

// ---------------------------------------------------------------------------
// Layer prefix tables
// ---------------------------------------------------------------------------

// One block per (request, layer).  Within a layer the chunks are ordered by
// activation, so both tile bounds are non-decreasing along the chunk list:
// for tile tau the saturated chunks are a prefix [begin, lo) and the chunks in
// transition a range [lo, hi).  Writes per tile {lo, hi} and the prefix sums
// A = sum_{c < lo} sum G P', B = sum_{c < lo} sum G Q'; slot n_bt holds the
// layer's sum G (for k0).  The sums have a fixed order that depends only on
// the layout: the layer's chunks form consecutive segments of `seg` chunks
// (at most PRE_SEGS segments), each summed in chunk order; the segment sums
// are accumulated in order; a prefix at lo is the accumulated sum of the
// segments before lo's segment plus that segment's chunks before lo, in order.
constexpr int PRE_SEGS = 128;

template <int LP>
__global__ __launch_bounds__(256) void layer_prefix_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ lay_start, const double* __restrict__ sums,
    const int2* __restrict__ bnd, int2* __restrict__ pre_rng, double* __restrict__ pre_ab)
{
    constexpr int NC3 = 3 * LP;
    __shared__ double s_seg[PRE_SEGS + 1][NC3];   // [k][col]: sum of the segments before k
    const int r = find_req(lay_start, B, blockIdx.x);
    const ReqDev R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int l = (int)(blockIdx.x - lay_start[r]) + 1;
    const int t = threadIdx.x;
    const int cb = l <= H.max_layer ? H.layer_chunk[l] : 0;
    const int ce = l <= H.max_layer ? H.layer_chunk[l + 1] : 0;
    const int nc = ce - cb;
    const int seg = nc <= PRE_SEGS ? 1 : (nc + PRE_SEGS - 1) / PRE_SEGS;
    const int nseg = (nc + seg - 1) / seg;
    const int nbt = R.Tp >> R.bt_shift;   // bound tiles
    const int64_t slot = R.pre_off + (int64_t)(l - 1) * (nbt + 1);
    const int2* bd = bnd + R.c_off;
    const double* sc = sums + R.c_off * NC3;
    // segment sums (columns: sum G, sum G P', sum G Q' per lead)
    for (int it = t; it < nseg * NC3; it += blockDim.x) {
        const int k = it / NC3, col = it - k * NC3;
        const int c0 = cb + k * seg, c1 = c0 + seg < ce ? c0 + seg : ce;
        double acc = 0.0;
        for (int c = c0; c < c1; ++c) acc += sc[(int64_t)c * NC3 + col];
        s_seg[k + 1][col] = acc;
    }
    __syncthreads();
    if (t < NC3) {
        double run = 0.0;
        s_seg[0][t] = 0.0;
        for (int k = 1; k <= nseg; ++k) {
            run += s_seg[k][t];
            s_seg[k][t] = run;
        }
    }
    __syncthreads();
    // per (tile, lead): transition range by binary searches, prefix at lo
    for (int it = t; it < nbt * LP; it += blockDim.x) {
        const int tile = it / LP, i = it - tile * LP;
        int a0 = cb, a1 = ce;
        while (a0 < a1) {
            const int mid = (a0 + a1) >> 1;
            if (bd[mid].y <= tile)
                a0 = mid + 1;
            else
                a1 = mid;
        }
        const int lo = a0;
        a1 = ce;
        while (a0 < a1) {
            const int mid = (a0 + a1) >> 1;
            if (bd[mid].x <= tile)
                a0 = mid + 1;
            else
                a1 = mid;
        }
        const int hi = a0;
        const int k = (lo - cb) / seg;
        double A = s_seg[k][LP + i], Bq = s_seg[k][2 * LP + i];
        for (int c = cb + k * seg; c < lo; ++c) {
            A += sc[(int64_t)c * NC3 + LP + i];
            Bq += sc[(int64_t)c * NC3 + 2 * LP + i];
        }
        if (i == 0) pre_rng[slot + tile] = int2{lo, hi};
        pre_ab[(slot + tile) * 2 * LP + i] = A;
        pre_ab[(slot + tile) * 2 * LP + LP + i] = Bq;
    }
    if (t < LP) {
        pre_ab[(slot + nbt) * 2 * LP + t] = s_seg[nseg][t];
        pre_ab[(slot + nbt) * 2 * LP + LP + t] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Fused AP evaluation + lead contraction (transition tiles)
// ---------------------------------------------------------------------------

// The cell AP of voxel u at t is
//   V = S(tau) * (P_u X(t) + Q_u Y(t)) + k0,  tau = t - a_u,
//   X(t) = e^{-(k4+k5)(t - t_ref)} R(t),  Y(t) = e^{-k5 (t - t_ref)} R(t),
// Wohlfart.h:197-202 with the exponentials split into a per-voxel factor
// (P, Q from lead_field) and a per-step factor (from the FP64 X_l, Y_l
// tables, rescaled to the chunk's t_ref, rounded once to FP32).  Voxels of a
// chunk are one layer, sorted by activation time, so for a tile they split
// into [v0, s1) saturated (2 VALU ops), [s1, s2) general, [s2, v1) skipped,
// found by ballot searches.  k0 is added by the reduction (k0 sum G).
//
// VALU variant: block = 4 waves of one (request, chunk); a wave takes tiles
// z + w, z + w + 4, ... < f; lane = step; lead sums are FP64 FMAs with the
// G row in SGPRs.
template <int LP>
__global__ __launch_bounds__(256) void ecg_fused_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ ch_start, const double* __restrict__ rec, const float4* __restrict__ aux,
    const LayerConst* __restrict__ lc, const double* __restrict__ layer_k,
    const double2* __restrict__ xy, const float2* __restrict__ thl, const int2* __restrict__ bnd,
    double* __restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = find_req(ch_start, B, blockIdx.x);
    const ReqDev R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int c = (int)(blockIdx.x - ch_start[r]);
    const int2 tb = bnd[R.c_off + c];
    const int4 ch = H.chunks[c];
    const int l = ch.x;
    const LayerConst kc = lc[R.lay_off + l - 1];
    const double* k = layer_k + R.k_off + 9 * (l - 1);
    const double tref = (double)__int_as_float(ch.w);
    const double dtr = tref - H.layer_t0[l];
    const double ex = exp((k[4] + k[5]) * dtr), ey = exp(k[5] * dtr);
    const double* Rb = rec + R.g_off * 16;
    const float4* Ab = aux + R.g_off;
    const double2* XY = xy + R.xy_off + (int64_t)(l - 1) * R.Tp;
    const float2* th = thl + R.t_off;

    for (int tile = tb.x + w; tile < tb.y; tile += 4) {
        const int j = tile * 64 + lane;
        const double2 xyj = XY[j];
        const float X = (float)fmin(xyj.x * ex, XMAX);
        const float Y = (float)fmin(xyj.y * ey, XMAX);
        const float2 tt = th[j];
        const float2 tf = th[tile * 64];
        const float2 tl = th[tile * 64 + 63];
        double acc[LP];
#pragma unroll
        for (int i = 0; i < LP; ++i) acc[i] = 0.0;
        int s1 = ch.y, s2 = ch.z;
        if (kc.c1 < 0.0f) {
            // saturated prefix: max over lanes of c1*tau <= SAT_X (max at the first lane)
            s1 = first_false(ch.y, ch.z, [&](int v) {
                const float2 a = H.vox_a[v];
                return kc.c1 * tau_of(tf.x, tf.y, a.x, a.y) <= SAT_X;
            });
            // non-zero range ends where min over lanes of c1*tau >= SKIP_X (min at the last lane)
            if (kc.skip != 0.0f)
                s2 = first_false(s1, ch.z, [&](int v) {
                    const float2 a = H.vox_a[v];
                    return kc.c1 * tau_of(tl.x, tl.y, a.x, a.y) < SKIP_X;
                });
        }
        // voxel data is wave-uniform (scalar loads); the G row feeds FP64 FMAs
        for (int v = ch.y; v < s2; ++v) {
            const float4 q = Ab[v];
            float ap = __builtin_fmaf(q.x, X, q.y * Y);
            if (v >= s1) ap *= fast_rcp(1.0f + fast_exp2(kc.c1 * tau_of(tt.x, tt.y, q.z, q.w)));
            const double apd = (double)ap;
            const double* g = Rb + (int64_t)v * 16;
#pragma unroll
            for (int i = 0; i < LP; ++i) acc[i] = __builtin_fma(g[i], apd, acc[i]);
        }
        double* out = partial + R.p_off + ((int64_t)c * LP) * R.Tp;
#pragma unroll
        for (int i = 0; i < LP; ++i)
            if (i < R.L) out[prow(i, j, LP, R.bt_shift)] = acc[i];
    }
}

typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Chunk kernel (matrix-core path): lead field, chunk sums and transition tiles
// of one (request, chunk) in one block
// ---------------------------------------------------------------------------

// Bit of neighbour offset d in a voxel's 3x3x3 myocardium mask.
__device__ constexpr int cube_bit(int d0, int d1, int d2) { return (d0 + 1) * 9 + (d1 + 1) * 3 + (d2 + 1); }

constexpr int CH_LEADS = 2;   // leads per pass of the chunk kernel's lead field (divides LP)

// G row of one myocardial voxel u from its own cell and its myocardial
// neighbours in the dipole neighbourhood N, CH_LEADS leads per pass:
//   G_i(u) = s(u) (e(u) . s_u) + sum_{d in N, u+d myo} s(u+d) (e(u+d) . d),
// s_u the sum of those d.  A cell's weight is s = rsq(|e|^2 rounded to
// FP32)^3 with |e|^2 in FP64; the offsets e = p - c are exact in FP64 (c is an
// integer coordinate + 1).  The voxel's own |e(u)|^2 is summed from its
// offsets; a neighbour cell's follows from it and the dipole sum
// sigma = e(u) . d, which the dipole product needs anyway:
//   e(u+d) . d = sigma - |d|^2,   |e(u+d)|^2 = |e(u)|^2 + |d|^2 - 2 sigma,
// one FMA per cell.  The FP64 value differs from the cell's directly summed
// |e|^2 (its own evaluation, or the class-window kernel's matrix product) by a few
// 1e-16 relative, so after rounding to FP32 a cell gets the same weight from
// every voxel that evaluates it except where those FP64 values straddle an
// FP32 rounding boundary (about one in 1e8 evaluations, each a one-ulp weight
// difference on one dipole); the weight's rounding therefore cancels between
// the two ends of each dipole and only V differences across the dipole see
// it (a weight derived per evaluating voxel in FP32 arithmetic would not
// cancel: ~1e-5).
template <int NH>
struct VoxelField {
    double cc[3];          // grid coordinate + 1 per axis (the reference's voxel centre), FP64
    double sd0, sd1, sd2;  // s_u
    uint32_t mk;           // 3x3x3 myocardium mask

    __device__ __forceinline__ explicit VoxelField(uint2 geo)
    {
        using N = Nhood<NH>;
        const int gz = (int)(geo.x & 1023u), gy = (int)((geo.x >> 10) & 1023u), gx = (int)(geo.x >> 20);
        mk = geo.y;
        int i0 = 0, i1 = 0, i2 = 0;
#pragma unroll
        for (int qn = 0; qn < N::n; ++qn) {
            const int f = (int)((mk >> cube_bit(N::d[qn][0], N::d[qn][1], N::d[qn][2])) & 1u);
            i0 += f * N::d[qn][0];
            i1 += f * N::d[qn][1];
            i2 += f * N::d[qn][2];
        }
        sd0 = (double)i0;
        sd1 = (double)i1;
        sd2 = (double)i2;
        cc[0] = (double)(gz + 1);
        cc[1] = (double)(gy + 1);
        cc[2] = (double)(gx + 1);
    }

    // g[w] = G_{i0+w}(u); leads at or past L give 0.  s_pd[i][c]: the
    // electrode in FP64.
    __device__ __forceinline__ void leads(int i0, int L, const double (*s_pd)[3], double g[CH_LEADS]) const
    {
        using N = Nhood<NH>;
#pragma unroll
        for (int w2 = 0; w2 < CH_LEADS; ++w2) g[w2] = 0.0;
        if (i0 >= L) return;
        double e[CH_LEADS][3], lo[CH_LEADS];
#pragma unroll
        for (int w2 = 0; w2 < CH_LEADS; ++w2) {
            // leads past the request's count repeat its first lead of the pass
            const int ii = i0 + w2 < L ? i0 + w2 : i0;
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) e[w2][c3] = s_pd[ii][c3] - cc[c3];
            lo[w2] = __builtin_fma(e[w2][2], e[w2][2], __builtin_fma(e[w2][1], e[w2][1], e[w2][0] * e[w2][0]));
            // own cell: m(u) . s_u
            const float r = __builtin_amdgcn_rsqf((float)lo[w2]);
            g[w2] = (double)(r * r * r) * (e[w2][0] * sd0 + e[w2][1] * sd1 + e[w2][2] * sd2);
        }
        // Corner neighbourhood (3D4): the eight dipole sums are
        // d0 ((e0 + d0 d1 e1) + d0 d2 e2), four values and their negations
        // (IEEE sums are odd in their operands, so these are the bits of the
        // term-by-term sums below), formed once per lead ahead of the
        // per-neighbour branches.
        double s4[CH_LEADS][4];
        if constexpr (NH == 2) {
#pragma unroll
            for (int w2 = 0; w2 < CH_LEADS; ++w2) {
                const double p = e[w2][0] + e[w2][1], m = e[w2][0] - e[w2][1];
                s4[w2][0] = p + e[w2][2];
                s4[w2][1] = p - e[w2][2];
                s4[w2][2] = m + e[w2][2];
                s4[w2][3] = m - e[w2][2];
            }
        }
#pragma unroll
        for (int qn = 0; qn < N::n; ++qn) {
            if ((mk >> cube_bit(N::d[qn][0], N::d[qn][1], N::d[qn][2])) & 1u) {
                const int d0 = N::d[qn][0], d1 = N::d[qn][1], d2 = N::d[qn][2];
                const double dd = (double)(d0 * d0 + d1 * d1 + d2 * d2);
#pragma unroll
                for (int w2 = 0; w2 < CH_LEADS; ++w2) {
                    // sigma = e(u) . d (d has entries -1, 0, 1)
                    double sg = 0.0;
                    if constexpr (NH == 2) {
                        const int q4 = (d0 * d1 > 0 ? 0 : 2) + (d0 * d2 > 0 ? 0 : 1);
                        sg = d0 > 0 ? s4[w2][q4] : -s4[w2][q4];
                    } else {
                        if (d0) sg = d0 > 0 ? e[w2][0] : -e[w2][0];
                        if (d1) sg = d1 > 0 ? sg + e[w2][1] : sg - e[w2][1];
                        if (d2) sg = d2 > 0 ? sg + e[w2][2] : sg - e[w2][2];
                    }
                    const double l = __builtin_fma(-2.0, sg, lo[w2] + dd);
                    const float r = __builtin_amdgcn_rsqf((float)l);
                    g[w2] = __builtin_fma((double)(r * r * r), sg - dd, g[w2]);
                }
            }
        }
#pragma unroll
        for (int w2 = 0; w2 < CH_LEADS; ++w2)
            if (i0 + w2 >= L) g[w2] = 0.0;
    }
};

// Block = `nw` waves of one (request, chunk) of at most NMAX voxels; the
// chunk's G rows never leave the CU.
//  1. Each voxel u of the chunk (one per thread, strided) gets its G row from its
//     own cell and its myocardial neighbours directly: per lead the lead
//     weight m = e s of each of those cells (FP32 e and s, lead_s32) and the
//     FP64 sum of the exact products
//       G_i(u) = s(u) (e(u) . s_u) + sum_{d in N, u+d myo} s(u+d) (e(u+d) . d)
//     into LDS, with the voxel's {P, Q, a} and FP64 {P', Q'}.  Wave 0 then
//     finds the bounds {z, f} in bound tiles of BT steps (the chunk_sums_kernel
//     predicates at that granularity): before z every voxel is skipped, from
//     f on every voxel is saturated and the reduction takes the closed form
//     X_l(t) sum G P' + Y_l(t) sum G Q'.
//     Per group of 16 voxels (chunk order) wave 0 also finds the quads (4
//     steps) of the tiles [z, f) in which the group is evaluated, [qs, qe): before
//     qs its earliest voxel is skipped, from qe on its latest is saturated.
//  2. Group sums sum G, sum G P', sum G Q' (voxels in order) and their
//     exclusive prefix over the groups in order; the last is the chunk's,
//     to `sums`.
//  3. Transition tiles [z, f) of BT steps, spread over the waves.  In quad
//     Q the groups saturated by then take the FP64 closed form X_l(t) (group
//     prefix of sum G P') + Y_l(t) (... sum G Q'); each group with
//     qs <= Q < qe is evaluated: a lane computes the FP32 AP of one (voxel,
//     step) pair, S (P X + Q Y), and the lead contraction
//     D[lead i][step j] += sum_k G[voxel k][lead i] V[voxel k][step j] runs
//     on v_mfma_f64_4x4x4f64 with the four blocks over the group's voxels
//     (block b takes voxels 4b .. 4b+3), one product per group of 4 leads.
//     gfx950 lane layout (checked by one-hot probes): A(block, row, k) in
//     lane 16k + 4 block + row, B(block, k, col) in lane 16k + 4 block + col,
//     D(block, row, col) in lane 16 row + 4 block + col.  LDS rows past the
//     chunk end (to the next multiple of 16) hold zeros and contribute
//     nothing; the blocks' sums are added in a fixed order.
// The host keeps every chunk of a group at most NMAX voxels.
template <int LP, int NH, int NMAX>
__global__ __launch_bounds__(1024) void ecg_chunk_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ ch_start, int ch_uniform, int one_heart, const double* __restrict__ lead_pos,
    const double* __restrict__ layer_k, const LayerConst* __restrict__ lc,
    const double2* __restrict__ xy, const double* __restrict__ tgrid, const float2* __restrict__ thl,
    double* __restrict__ sums, int2* __restrict__ bnd, double* __restrict__ partial)
{
    using N = Nhood<NH>;
    __shared__ double s_G[NMAX * LP];       // G rows of the chunk's voxels
    __shared__ float4 s_aux[NMAX];          // {P, Q, a_hi, a_lo} relative to the chunk's t_ref
    __shared__ double2 s_pq[NMAX];          // {P', Q'} relative to the layer's t0
    __shared__ double s_grp[NMAX / 16 + 1][3 * LP];  // group sums: exclusive prefix over the groups
    __shared__ int2 s_qb[NMAX / 16];        // per group: first evaluated and first saturated quad
    __shared__ double s_pd[16][3];          // electrode in FP64
    __shared__ int s_tb[2];
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int nw = blockDim.x >> 6;
    // Chunk-major when every request has the same number of chunks (one
    // heart): the requests of a chunk run back to back and read its voxel
    // records from L2 (blocks go round-robin over the XCDs, so each XCD
    // fetches a chunk's records once).
    int r, c;
    if (ch_uniform > 0) {
        r = (int)(blockIdx.x % (unsigned)B);
        c = (int)(blockIdx.x / (unsigned)B);
    } else {
        r = find_req(ch_start, B, blockIdx.x);
        c = (int)(blockIdx.x - ch_start[r]);
    }
    // With every request on one heart its record is read beside the
    // request's, and the chunk record right after: one dependent load fewer
    // before the chunk's voxel records can be requested.
    const HeartDev* Hp = hearts;
    if (!one_heart) Hp = hearts + reqs[r].heart;
    const HeartDev& H = *Hp;
    const ReqDev R = reqs[r];
    // the chunk record through the constant address space (a scalar load)
    using cu64 = __attribute__((address_space(4))) const unsigned long long;
    const unsigned long long cw0 = ((cu64*)H.chunks)[2 * c], cw1 = ((cu64*)H.chunks)[2 * c + 1];
    const int4 ch = int4{(int)(unsigned)cw0, (int)(cw0 >> 32), (int)(unsigned)cw1, (int)(cw1 >> 32)};
    const int l = ch.x, v0 = ch.y, n = ch.z - ch.y;
    const int ng = (n + 15) >> 4;  // groups of 16 voxels
    // the records of the thread's first voxel through the global address
    // space, requested before the request's coefficients (they depend on the
    // chunk alone) and first used after the barrier below
    using gu64 = __attribute__((address_space(1))) const unsigned long long;
    unsigned long long geo0 = 0ull, a0 = 0ull, av0 = 0ull;
    if (t < n) {
        geo0 = ((gu64*)H.vox_geo)[v0 + t];
        a0 = ((gu64*)H.vox_a64)[v0 + t];
        av0 = ((gu64*)H.vox_a)[v0 + t];
    }
    // the layer's coefficients, reference time and the chunk's extreme
    // activation times as scalar loads, all issued together
    const cu64* kk = (const cu64*)(layer_k + R.k_off + 9 * (l - 1));
    const LayerConst kc = lc[R.lay_off + l - 1];
    const double t0 = __builtin_bit_cast(double, ((cu64*)H.layer_t0)[l]);
    const double tref = (double)__int_as_float(ch.w);
    const double k1v = __builtin_bit_cast(double, kk[1]), k2v = __builtin_bit_cast(double, kk[2]),
                 k3v = __builtin_bit_cast(double, kk[3]), k4v = __builtin_bit_cast(double, kk[4]),
                 k5v = __builtin_bit_cast(double, kk[5]);
    const double alpha = k2v * (1.0 - k3v), beta = k2v * k3v, k45 = k4v + k5v, k5 = k5v;
    const double c1d = -k1v * LOG2E_D;
    const double* tg = tgrid + R.t_off;
    const float2* th = thl + R.t_off;
    if (t < 3 * R.L) s_pd[t / 3][t % 3] = __builtin_bit_cast(double, ((gu64*)lead_pos)[R.pos_off + t]);
    // operands of the bound-tile tests (tiles of BT steps), loaded now and
    // used after the G rows: up to NPROBE probes per lane (64 NPROBE tiles)
    constexpr int BT = 1 << BT_SHIFT_MFMA;
    constexpr int QPT = BT / 4;  // quads (4 steps) per tile
    constexpr int NPROBE = BT >= 8 ? 2 : 4;
    const int nbt = R.Tp >> BT_SHIFT_MFMA;
    const double a_max = __builtin_bit_cast(double, ((cu64*)H.vox_a64)[ch.z - 1]);
    const unsigned long long amw = ((cu64*)H.vox_a)[ch.y];
    const float2 a_min = float2{__int_as_float((int)(unsigned)amw), __int_as_float((int)(amw >> 32))};
    double tg_probe[NPROBE];
    float2 th_probe[NPROBE];
#pragma unroll
    for (int k = 0; k < NPROBE; ++k) {
        tg_probe[k] = 0.0;
        th_probe[k] = float2{0.0f, 0.0f};
    }
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < NPROBE; ++k)
            if (lane + 64 * k < nbt) {
                tg_probe[k] = __builtin_bit_cast(double, ((gu64*)tg)[(lane + 64 * k) * BT]);
                const unsigned long long pw = ((gu64*)th)[(lane + 64 * k) * BT + BT - 1];
                th_probe[k] = float2{__int_as_float((int)(unsigned)pw), __int_as_float((int)(pw >> 32))};
            }
    }
    // factors from the layer's t0 to the chunk's t_ref: P = P' e^{(k4+k5)(t0 - t_ref)}, Q = Q' e^{k5 (t0 - t_ref)}
    const double eP = exp(k45 * (t0 - tref)), eQ = exp(k5 * (t0 - tref));
    __syncthreads();

    // 1. G rows and AP factors, one voxel per thread (strided)
    for (int q = t; q < n; q += blockDim.x) {
        const int v = v0 + q;
        const bool first = q == t;
        const VoxelField<NH> vf(first ? uint2{(unsigned)geo0, (unsigned)(geo0 >> 32)} : H.vox_geo[v]);
#pragma unroll 1
        for (int i0 = 0; i0 < LP; i0 += CH_LEADS) {
            double g[CH_LEADS];
            vf.leads(i0, R.L, s_pd, g);
#pragma unroll
            for (int w2 = 0; w2 < CH_LEADS; ++w2) s_G[q * LP + i0 + w2] = g[w2];
        }
        const double a = first ? __builtin_bit_cast(double, a0) : H.vox_a64[v];
        const float2 av = first ? float2{__int_as_float((int)(unsigned)av0), __int_as_float((int)(av0 >> 32))}
                                : H.vox_a[v];
        const double pp = alpha * exp(k45 * (a - t0)), qq = beta * exp(k5 * (a - t0));
        s_aux[q] = float4{(float)(pp * eP), (float)(qq * eQ), av.x, av.y};
        s_pq[q] = double2{pp, qq};
    }
    // Rows n .. (n rounded up to 16): zero G and P = Q = 0, so that the
    // contraction loads whole groups of 16 rows unclamped and they add 0.
    for (int q = n + t; q < ((n + 15) & ~15); q += blockDim.x) {
#pragma unroll
        for (int i = 0; i < LP; ++i) s_G[q * LP + i] = 0.0;
        s_aux[q] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    if (w == 0) {
        // tile bounds (the same predicates as chunk_sums_kernel), needed from
        // the transition tiles on: ballots over the prefetched probes, or
        // the wave-parallel search for requests with more tiles
        const bool sat_ok = c1d < 0.0, skip_ok = kc.c1 < 0.0f && kc.skip != 0.0f;
        int f = nbt, z = 0;
        if (nbt <= 64 * NPROBE) {
            // both predicates are monotone in the tile: the count of false
            // probes is the first true tile
            int cf = 0, cz = 0;
#pragma unroll
            for (int k = 0; k < NPROBE; ++k) {
                const bool in = lane + 64 * k < nbt;
                const bool fp = in && sat_ok && c1d * (tg_probe[k] - a_max) <= SATC_X;
                const bool zp = in && skip_ok &&
                                kc.c1 * tau_of(th_probe[k].x, th_probe[k].y, a_min.x, a_min.y) < SKIP_X;
                cf += __popcll(__ballot(fp));
                cz += __popcll(__ballot(zp));
            }
            if (sat_ok) f = nbt - cf;
            if (skip_ok) z = nbt - cz;
        } else {
            if (sat_ok)
                f = first_false(0, nbt, [&](int tile) { return !(c1d * (tg[tile * BT] - a_max) <= SATC_X); });
            if (skip_ok)
                z = first_false(0, nbt, [&](int tile) {
                    const float2 tl = th[tile * BT + BT - 1];
                    return !(kc.c1 * tau_of(tl.x, tl.y, a_min.x, a_min.y) < SKIP_X);
                });
        }
        const int zf = z < f ? z : f;
        if (lane == 0) {
            s_tb[0] = zf;
            s_tb[1] = f;
            bnd[R.c_off + c] = int2{zf, f};
        }
        // per group of 16 voxels (lane g): the quads of the tiles [zf, f) in
        // which it is evaluated, [qs, qe).  qs - QPT zf counts the quads whose
        // last step still skips the group's earliest voxel, qe - QPT zf those
        // whose first step does not yet saturate its latest voxel (the tile
        // predicates above per group; both hold on a prefix of the quads).
        if (lane < ng) {
            const float2 af = H.vox_a[v0 + 16 * lane];
            const double al = H.vox_a64[v0 + (16 * lane + 15 < n ? 16 * lane + 15 : n - 1)];
            int qs = QPT * zf, qe = QPT * zf;
            for (int Q = QPT * zf; Q < QPT * f; ++Q) {
                const float2 tl = th[4 * Q + 3];
                qs += (skip_ok && !(kc.c1 * tau_of(tl.x, tl.y, af.x, af.y) < SKIP_X)) ? 1 : 0;
                qe += (sat_ok && c1d * (tg[4 * Q] - al) <= SATC_X) ? 0 : 1;
            }
            s_qb[lane] = int2{qs, qe};
        }
    }
    __syncthreads();

    // 2. group and chunk sums.  Per group of 16 voxels (chunk order) and lead:
    // sum G, sum G P', sum G Q' over the group's voxels in order; then per
    // (sum, lead) the exclusive prefix over the groups in order: s_grp[g]
    // holds the sums of groups 0 .. g-1, s_grp[ng] the chunk's.
    for (int u = t; u < ng * LP; u += blockDim.x) {
        const int g = u / LP, i = u - g * LP;
        const int qa = 16 * g, qz = qa + 16 < n ? qa + 16 : n;
        double sG = 0.0, sP = 0.0, sQ = 0.0;
        for (int q = qa; q < qz; ++q) {
            const double gq = s_G[q * LP + i];
            const double2 pq = s_pq[q];
            sG += gq;
            sP = __builtin_fma(gq, pq.x, sP);
            sQ = __builtin_fma(gq, pq.y, sQ);
        }
        s_grp[g][i] = sG;
        s_grp[g][LP + i] = sP;
        s_grp[g][2 * LP + i] = sQ;
    }
    __syncthreads();
    if (t < 3 * LP) {
        double run = 0.0;
        for (int g = 0; g < ng; ++g) {
            const double x = s_grp[g][t];
            s_grp[g][t] = run;
            run += x;
        }
        s_grp[ng][t] = run;
        sums[(R.c_off + c) * 3 * LP + t] = run;
    }
    __syncthreads();

    // 3. transition tiles [z, f) of BT steps, spread over the waves.  A
    // tile is QPT quads of 4 steps; in quad Q group g is skipped (Q < qs),
    // evaluated (qs <= Q < qe) or saturated (Q >= qe), and the saturated
    // groups 0 .. nsq(Q)-1 take the FP64 closed form X_l s_grp[nsq].P' +
    // Y_l s_grp[nsq].Q'.  Evaluated groups run on v_mfma_f64_4x4x4f64 with
    // the blocks over the group's voxels: block b, inner index k holds voxel
    // 4b + k, so a lane computes the AP of one (voxel, step) pair of the quad
    // and feeds it to the LP/4 products of the lead groups; the blocks'
    // partial sums are added in block order.
    const int tz = s_tb[0], tf = s_tb[1];
    if (tf - tz <= w) return;  // no tile for this wave
    const double ex = exp(k45 * (tref - t0)), ey = exp(k5 * (tref - t0));
    constexpr int NLG = LP / 4;
    const int col = lane & 3;                             // step of the quad; lead row of the A operand
    const int vs = 4 * ((lane >> 2) & 3) + (lane >> 4);   // voxel of the group (A and B operands)
    const int2 qb = lane < ng ? s_qb[lane] : int2{0x7fffffff, 0x7fffffff};
    const double2* XY = xy + R.xy_off + (int64_t)(l - 1) * R.Tp;
    double* out = partial + R.p_off + (int64_t)c * LP * R.Tp;
    for (int sb = tz + w; sb < tf; sb += nw) {
        const int Qt = QPT * sb, j0 = BT * sb;
        float Xq[QPT], Yq[QPT], Th[QPT], Tl[QPT];
#pragma unroll
        for (int qd = 0; qd < QPT; ++qd) {
            const int j = j0 + 4 * qd + col;
            const double2 xyj = XY[j];
            Xq[qd] = (float)fmin(xyj.x * ex, XMAX);
            Yq[qd] = (float)fmin(xyj.y * ey, XMAX);
            const float2 tt = th[j];
            Th[qd] = tt.x;
            Tl[qd] = tt.y;
        }
        // groups saturated from the tile's first quad on, and groups not
        // skipped in its last quad: both prefixes of the group order
        const int g_lo = __popcll(__ballot(qb.y <= Qt));
        const int g_hi = __popcll(__ballot(qb.x < Qt + QPT));
        double acc[QPT][NLG];
#pragma unroll
        for (int qd = 0; qd < QPT; ++qd)
#pragma unroll
            for (int lg = 0; lg < NLG; ++lg) acc[qd][lg] = 0.0;
        for (int g = g_lo; g < g_hi; ++g) {
            const int qs = __builtin_amdgcn_readlane(qb.x, g), qe = __builtin_amdgcn_readlane(qb.y, g);
            const int v = 16 * g + vs;
            double ga[NLG];
#pragma unroll
            for (int lg = 0; lg < NLG; ++lg) ga[lg] = s_G[v * LP + 4 * lg + col];
            const float4 ax = s_aux[v];
#pragma unroll
            for (int qd = 0; qd < QPT; ++qd) {
                if (Qt + qd >= qs && Qt + qd < qe) {
                    float ap = __builtin_fmaf(ax.x, Xq[qd], ax.y * Yq[qd]);
                    ap *= fast_rcp(1.0f + fast_exp2(kc.c1 * tau_of(Th[qd], Tl[qd], ax.z, ax.w)));
                    const double bv = (double)ap;
#pragma unroll
                    for (int lg = 0; lg < NLG; ++lg)
                        acc[qd][lg] = __builtin_amdgcn_mfma_f64_4x4x4f64(ga[lg], bv, acc[qd][lg], 0, 0, 0);
                }
            }
        }
        // D(block, row, col) in lane 16 row + 4 block + col: lead 4 lg + row,
        // step col of the quad; blocks (0 + 1) + (2 + 3)
#pragma unroll
        for (int qd = 0; qd < QPT; ++qd) {
            const int nsq = __popcll(__ballot(qb.y <= Qt + qd));
            const int j = j0 + 4 * qd + col;
            const double2 xyj = XY[j];
            const double* cp = s_grp[nsq];
#pragma unroll
            for (int lg = 0; lg < NLG; ++lg) {
                double x = acc[qd][lg];
                x += __shfl_xor(x, 4, 64);
                x += __shfl_xor(x, 8, 64);
                const int i = 4 * lg + (lane >> 4);
                if ((lane & 12) == 0 && i < R.L)
                    out[prow(i, j, LP, BT_SHIFT_MFMA)] = __builtin_fma(xyj.y, cp[2 * LP + i], __builtin_fma(xyj.x, cp[LP + i], x));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Reduction
// ---------------------------------------------------------------------------

// ecg[r][i][t] for t < Tn: per layer (in order) the closed form of the
// chunks saturated by t's tile, X_l(t) A + Y_l(t) B, then k0 * sum G, then
// the partial rows of the chunks in transition in chunk order; 0 for
// Tn <= t < T: samples past the last iteration of the reference's neighbour
// loop (simulator.cpp:519) receive no dipole.
template <int LP>
__global__ void ecg_reduce_kernel(const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs,
                                  int B, const int64_t* __restrict__ red_start, int64_t n_red,
                                  const LayerConst* __restrict__ lc, const double2* __restrict__ xy,
                                  const int2* __restrict__ pre_rng, const double* __restrict__ pre_ab,
                                  const double* __restrict__ partial, double* __restrict__ ecg)
{
    // Items per request in the partial rows' order: bound tile, lead, step
    // of the tile, so that consecutive threads read consecutive doubles of a
    // tile's rows and share the tile's chunk range.
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n_red;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = find_req(red_start, B, idx);
        const ReqDev R = reqs[r];
        const HeartDev& H = hearts[R.heart];
        const int64_t loc = idx - red_start[r];
        const int bs = R.bt_shift;
        const int64_t blk = (int64_t)R.L << bs;
        const int tile = (int)(loc / blk);
        const int rem = (int)(loc - tile * blk);
        const int i = rem >> bs;
        const int j = (tile << bs) + (rem & ((1 << bs) - 1));
        if (j >= R.T) continue;
        double s = 0.0;
        if (j < R.Tn) {
            const int nbt = R.Tp >> bs;
            const int nl = R.nl < H.max_layer ? R.nl : H.max_layer;
            for (int l = 1; l <= nl; ++l) {
                const int64_t slot = R.pre_off + (int64_t)(l - 1) * (nbt + 1);
                const double2 xyj = xy[R.xy_off + (int64_t)(l - 1) * R.Tp + j];
                const double* ab = pre_ab + (slot + tile) * 2 * LP;
                s = __builtin_fma(xyj.x, ab[i], s);
                s = __builtin_fma(xyj.y, ab[LP + i], s);
                s = __builtin_fma(lc[R.lay_off + l - 1].k0, pre_ab[(slot + nbt) * 2 * LP + i], s);
                const int2 rg = pre_rng[slot + tile];
                const double* p = partial + R.p_off + prow(i, j, LP, bs);
                const int64_t st = (int64_t)LP * R.Tp;
                // eight independent loads in flight, added in chunk order
                int c = rg.x;
                for (; c + 8 <= rg.y; c += 8) {
                    double x[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(c + u) * st];
#pragma unroll
                    for (int u = 0; u < 8; ++u) s += x[u];
                }
                for (; c < rg.y; ++c) s += p[(int64_t)c * st];
            }
        }
        ecg[R.ecg_off + (int64_t)i * R.T + j] = s;
    }
}

// ---------------------------------------------------------------------------
// Activation classes
// ---------------------------------------------------------------------------

constexpr int CC_THREADS = 256;

// Class windows: one block per (piece of a class window, request).  Every
// voxel of a class has the same AP, so only the sum of the class's G rows is
// needed, and by regrouping the dipoles per cell it is
//   sum_u G(u) = sum_w s(w) (e(w) . D(w))
// over the cells w with a non-zero integer vector D(w) for the class
// (host-built per neighbourhood; cells inside the class cancel).  With
// e = p - c (c the cell's coordinate + 1) this is linear in the electrode p:
//   sum_w s(w) (e . D) = p . T + T3,  T_c = sum_w s(w) D_c(w),  T3 = sum_w s(w) K(w),
// K = -D . c.  A cell borders several classes, so the host lists each cell
// once per window of CLASS_WINDOW classes (an entry) with its D and K for
// each class of the window, and the weight s(w) is evaluated once per entry
// and lead.  The squared distance and the products run on
// v_mfma_f64_4x4x4f64 (4 blocks of 4 x 4 x 4):
//  1. |e|^2 = |p|^2 - 2 p . c + |c|^2 as an FP64 matrix product: rows are the
//     16 entries of a group (4 per block), columns the 4 leads of a group, the
//     inner dimension {c_z, c_y, c_x, 1} x {-2 p_z, -2 p_y, -2 p_x, |p|^2},
//     with |c|^2 (exact) as the accumulator input;
//  2. s = rsq(|e|^2 rounded to FP32)^3 in FP32 -- the chunk kernel's weight:
//     it rounds its own FP64 |e|^2 (from the exact FP64 offsets, a
//     neighbour's from the voxel's |e|^2 and dipole sum) the same way, so the
//     two kernels give a cell the same weight except where the two FP64
//     values (both within ~1e-14 relative of |e|^2) straddle an FP32
//     rounding boundary;
//  3. for each class i of the group's mask, T_i += s x {D_z, D_y, D_x, K}_i:
//     rows the leads, inner dimension the entries, one accumulator per class
//     of the window.
// gfx950 lane layout (checked by a one-hot probe): A(block, row, k) in lane
// 16k + 4 block + row, B(block, k, col) in lane 16k + 4 block + col,
// D(block, row, col) in lane 16 row + 4 block + col.  Product 1's result for
// (entry r of block b, lead col) lands in lane 16r + 4b + col, which is where
// product 3 takes its operand (block b, lead row, entry k): so a lane owns one
// entry of each group of 16 and one lead of each group of 4 throughout, and
// feeds product 1 coordinate (lane >> 4) of entry (b, lane & 3).  The host
// stores those per-lane operands of every group lane-major, so a lane loads
// coalesced words and converts them, with no index arithmetic or unpacking.
// Each wave takes groups of the piece strided over the waves, the blocks'
// partial sums are added in block order, the waves' in wave order, and
// sums[(c_off + n_chunks + p CLASS_WINDOW + i)*3*LP] receives the piece's
// {sum G, P' sum G, Q' sum G} of class i of its window, with the class's
// P' = k2(1-k3) e^{(k4+k5)(a - t0)} and Q' = k2 k3 e^{k5 (a - t0)} (zero for
// a class past the heart's last).
template <int LP>
__global__ __launch_bounds__(CC_THREADS) void ecg_class_window_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, const int4* __restrict__ runs,
    const int64_t* __restrict__ run_start, int n_runs, const double* __restrict__ lead_pos,
    const double* __restrict__ layer_k, double* __restrict__ sums)
{
    constexpr int NG = LP / 4;                 // groups of 4 leads
    constexpr int NW = CC_THREADS / 64;
    constexpr int CW = CLASS_WINDOW;
    __shared__ double s_t[CW][NW][LP][4];
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    // Piece-major within each run of requests of one heart (a run
    // {first request, requests, pieces}): the requests of a piece run back to
    // back and read its groups from L2 (blocks go round-robin over the XCDs,
    // so each XCD fetches a piece once per run).
    const int u = find_req(run_start, n_runs, blockIdx.x);
    const int4 run = runs[u];
    const int local = (int)(blockIdx.x - run_start[u]);
    const int k = local / run.y, r = run.x + (local - k * run.y);
    const ReqDev& R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int4 pc4 = H.cpieces[k];
    const int pw = __builtin_amdgcn_readfirstlane(pc4.x), g0 = __builtin_amdgcn_readfirstlane(pc4.y),
              g1 = __builtin_amdgcn_readfirstlane(pc4.z);
    const int rc = lane & 3;                                  // lead of a group (and product-3 column)
    const int kq = lane >> 4;                                 // product-1 inner index: coordinate kq
    // product-1 right operand: {-2 p_z, -2 p_y, -2 p_x, |p|^2} of the lane's
    // leads 4g + rc (leads past the request's count repeat lead 0; their sums
    // are zeroed)
    double pe[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int i = 4 * g + rc < R.L ? 4 * g + rc : 0;
        const double* p = lead_pos + R.pos_off + 3 * i;
        const double pp = __builtin_fma(p[2], p[2], __builtin_fma(p[1], p[1], p[0] * p[0]));
        pe[g] = kq < 3 ? -2.0 * p[kq < 3 ? kq : 0] : pp;
    }
    double acc[CW][NG];
#pragma unroll
    for (int i = 0; i < CW; ++i)
#pragma unroll
        for (int g = 0; g < NG; ++g) acc[i][g] = 0.0;
    // Each group's rows and class mask are wave-uniform: read through the
    // constant address space they become scalar loads, and the group's rows
    // through the global one take a scalar base and the lane's offset, so
    // neither needs vector address arithmetic.
    using ci64 = __attribute__((address_space(4))) const long long;
    ci64* const crow = (ci64*)H.crow;
    for (int q = g0 + w; q < g1; q += NW) {
        const long long cr = crow[q];  // {first row, class mask}
        const int row = (int)(cr & 0xffffffffLL), mask = (int)(cr >> 32);
        // the group's rows as a buffer resource: the lane's offset is a
        // fixed vector operand and each row's offset a scalar one (reads
        // past the group's rows would return 0, never fault)
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(H.cgrp + (int64_t)64 * row), 0, 256 * (2 + __builtin_popcount(mask)), 0x00020000);
        const int lo = 4 * lane;
        // product-1 left operand (coordinate kq of entry 4 block + rc, or 1)
        // and |c|^2 of entry 4 block + kq
        const double av = (double)__int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, lo, 0, 0));
        const double cc = (double)__int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, lo, 256, 0));
        float bf[CW];
#pragma unroll
        for (int i = 0, n = 2; i < CW; ++i)
            if ((mask >> i) & 1) bf[i] = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, lo, 256 * n++, 0));
        // A wave past its loads takes issue priority for the products and
        // weights of the group, and gives it back once its next loads are
        // due, so that the waves of a SIMD spread over the loop instead of
        // waiting on their loads together (measured: 1.0% per step).
        __builtin_amdgcn_s_setprio(1);
        double sw[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double e2 = __builtin_amdgcn_mfma_f64_4x4x4f64(av, pe[g], cc, 0, 0, 0);
            const float rr = __builtin_amdgcn_rsqf((float)e2);
            sw[g] = (double)(rr * rr * rr);
        }
#pragma unroll
        for (int i = 0; i < CW; ++i)
            if ((mask >> i) & 1) {
                // product-3 right operand: coefficient rc of entry 4 block + kq for class i
                const double bv = (double)bf[i];
#pragma unroll
                for (int g = 0; g < NG; ++g) acc[i][g] = __builtin_amdgcn_mfma_f64_4x4x4f64(sw[g], bv, acc[i][g], 0, 0, 0);
            }
        __builtin_amdgcn_s_setprio(0);
    }
    // acc[i][g] in lane 16 row + 4 block + col: lead 4g + row, coefficient
    // col, this wave's entries of one block; the four blocks in order
    const int base = lane & ~12;
#pragma unroll
    for (int i = 0; i < CW; ++i)
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double v0 = __shfl(acc[i][g], base), v1 = __shfl(acc[i][g], base | 4),
                         v2 = __shfl(acc[i][g], base | 8), v3 = __shfl(acc[i][g], base | 12);
            if ((lane & 12) == 0) s_t[i][w][4 * g + (lane >> 4)][lane & 3] = ((v0 + v1) + v2) + v3;
        }
    __syncthreads();
    for (int u = t; u < CW * LP; u += CC_THREADS) {
        const int i = u / LP, ld = u - i * LP;
        const int kc = CW * pw + i;
        double T[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            T[c] = 0.0;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) T[c] += s_t[i][ww][ld][c];
        }
        double sg = 0.0, pp = 0.0, qq = 0.0;
        if (kc < H.n_classes) {
            if (ld < R.L) {
                const double* p = lead_pos + R.pos_off + 3 * ld;
                sg = __builtin_fma(p[2], T[2], __builtin_fma(p[1], T[1], __builtin_fma(p[0], T[0], T[3])));
            }
            const int l = H.classes[kc].x;
            const double* kk = layer_k + R.k_off + 9 * (l - 1);
            const double da = H.class_a[kc] - H.layer_t0[l];
            pp = kk[2] * (1.0 - kk[3]) * exp((kk[4] + kk[5]) * da);
            qq = kk[2] * kk[3] * exp(kk[5] * da);
        }
        double* os = sums + (R.c_off + H.n_chunks + (int64_t)k * CW + i) * 3 * LP;
        os[ld] = sg;
        os[LP + ld] = pp * sg;
        os[2 * LP + ld] = qq * sg;
    }
}

// csums[(cls_off + k)*3*LP + v]: the sums of class k (k0 sum G, sum G P',
// sum G Q' per lead, k0 the class's layer's resting value), its window's
// pieces added in order.
template <int LP>
__global__ void class_sums_kernel(const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
                                  const int64_t* __restrict__ cs_start, const double* __restrict__ layer_k,
                                  const double* __restrict__ sums, double* __restrict__ csums)
{
    // grid rows over the requests (the request is block-uniform), columns
    // over a request's (class, column) items
    for (int r = blockIdx.y; r < B; r += gridDim.y) {
        const int64_t c0 = cs_start[r];
        const int64_t n = (cs_start[r + 1] - c0) * 3 * LP;
        if (n == 0) continue;
        const ReqDev& R = reqs[r];
        const HeartDev& H = hearts[R.heart];
        for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
             idx += (int64_t)gridDim.x * blockDim.x) {
            const int k = (int)(idx / (3 * LP));
            const int v = (int)(idx - (int64_t)k * (3 * LP));
            const int wi = k / CLASS_WINDOW, i = k - wi * CLASS_WINDOW;
            double s = 0.0;
            for (int p = H.cwin[wi]; p < H.cwin[wi + 1]; ++p)
                s += sums[(R.c_off + H.n_chunks + (int64_t)p * CLASS_WINDOW + i) * 3 * LP + v];
            if (v < LP) s *= layer_k[R.k_off + 9 * (H.classes[k].x - 1)];
            csums[(c0 + k) * 3 * LP + v] = s;
        }
    }
}

// ecg[r][i][t] += sum over the heart's activation classes of
//   S(t - a) (X_l(t) sum G P' + Y_l(t) sum G Q') + k0 sum G,
// with S(tau) = 1 / (1 + exp(-k1 tau)), the factor of the AP
// (Wohlfart.h:195-203) at the cell time tau = t - a (simulator.cpp:154-170),
// in FP32 as in the per-voxel AP (tau = t - a formed in FP64 and rounded once;
// 1 + 2^(c1 tau) rounds to 1 once the upstroke is 2^-24 from done, so S is
// exactly 1 from there on), times the FP64 sums.  Every voxel of a class has
// the same layer and activation time, hence the same AP, so the class's dipole
// terms factor through its sums.  Samples from Tn on receive nothing, as in
// the reduction.  One block of CL_WAVES waves per (request, 64 steps), one
// step per lane: wave w adds the classes [w n / CL_WAVES, (w+1) n / CL_WAVES)
// in order, two FMAs per lead, then the waves' sums are added in wave order.
// The time-independent sum over the classes of k0 sum G (csums holds k0 sum G)
// is formed once per block: thread (part p, lead i) adds classes p, p + NP,
// ..., then lead i's parts are added in order.
constexpr int CL_WAVES = 4;

template <int LP>
__global__ __launch_bounds__(64 * CL_WAVES) void ecg_class_kernel(
    const HeartDev* __restrict__ hearts, const ReqDev* __restrict__ reqs, int B,
    const int64_t* __restrict__ cl_start, int cl_per_req, const double* __restrict__ layer_k,
    const double2* __restrict__ xy, const double* __restrict__ tgrid, const double* __restrict__ csums,
    double* __restrict__ ecg)
{
    constexpr int NP = 64 * CL_WAVES / LP;  // parts of the k0 sum per lead
    __shared__ double s_acc[CL_WAVES - 1][LP][64];
    __shared__ double s_k0[NP][LP];
    // Blocks go round-robin over the 8 XCDs.  With cl_per_req blocks for
    // every request and B a multiple of 8, XCD x takes requests x, x + 8, ...
    // whole, so each request's class sums are fetched into one L2.
    int r, jb;
    if (cl_per_req > 0) {
        const int x = (int)(blockIdx.x & 7u), slot = (int)(blockIdx.x >> 3);
        r = 8 * (slot / cl_per_req) + x;
        jb = slot % cl_per_req;
    } else {
        r = find_req(cl_start, B, blockIdx.x);
        jb = (int)(blockIdx.x - cl_start[r]);
    }
    const ReqDev R = reqs[r];
    const HeartDev& H = hearts[R.heart];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int j0 = jb * 64;
    if (j0 >= R.Tn) return;  // the whole block
    const int j = j0 + lane;
    const bool live = j < R.Tn;
    const double t = tgrid[R.t_off + j];  // j < Tp: in the grid
    if ((int)threadIdx.x < NP * LP) {
        const int p = (int)threadIdx.x / LP, i = (int)threadIdx.x - p * LP;
        double c = 0.0;
#pragma unroll 4
        for (int k = p; k < H.n_classes; k += NP) c += csums[(R.cls_off + k) * 3 * LP + i];
        s_k0[p][i] = c;
    }
    double acc[LP];
#pragma unroll
    for (int i = 0; i < LP; ++i) acc[i] = 0.0;
    const int k0c = __builtin_amdgcn_readfirstlane((int)((int64_t)H.n_classes * w / CL_WAVES)),
              k1c = __builtin_amdgcn_readfirstlane((int)((int64_t)H.n_classes * (w + 1) / CL_WAVES));
    int lp = 0;
    double X = 0.0, Y = 0.0;
    float c1 = 0.0f;
    // the class records through the constant address space: scalar loads
    using ci32 = __attribute__((address_space(4))) const int;
    using ci64 = __attribute__((address_space(4))) const long long;
    ci32* const cls = (ci32*)H.classes;
    ci64* const cla = (ci64*)H.class_a;
    for (int kv = k0c; kv < k1c; ++kv) {
        const int k = __builtin_amdgcn_readfirstlane(kv);
        const int l = cls[4 * k];
        if (l != lp) {  // classes are ordered by layer
            const double2 xyj = xy[R.xy_off + (int64_t)(l - 1) * R.Tp + j];
            X = xyj.x;
            Y = xyj.y;
            c1 = (float)(-layer_k[R.k_off + 9 * (l - 1) + 1] * LOG2E_D);
            lp = l;
        }
        const float tau = (float)(t - __builtin_bit_cast(double, cla[k]));
        const double S = (double)fast_rcp(1.0f + fast_exp2(c1 * tau));
        const double SX = S * X, SY = S * Y;
        const double* cs = csums + (R.cls_off + k) * 3 * LP;
#pragma unroll
        for (int i = 0; i < LP; ++i) acc[i] = __builtin_fma(SY, cs[2 * LP + i], __builtin_fma(SX, cs[LP + i], acc[i]));
    }
    if (w > 0)
#pragma unroll
        for (int i = 0; i < LP; ++i) s_acc[w - 1][i][lane] = acc[i];
    __syncthreads();
    if (w == 0 && live) {
#pragma unroll
        for (int i = 0; i < LP; ++i) {
            double v = acc[i];
#pragma unroll
            for (int ww = 0; ww < CL_WAVES - 1; ++ww) v += s_acc[ww][i][lane];
            double c0 = 0.0;
            for (int p = 0; p < NP; ++p) c0 += s_k0[p][i];
            if (i < R.L) ecg[R.ecg_off + (int64_t)i * R.T + j] += v + c0;
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int grid_for(int64_t n, int threads, int cap = 1 << 20)
{
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

hipError_t launch_ap_tables(const GroupArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(ap_tables_kernel, dim3(grid_for(a.n_xy, 256)), dim3(256), 0, s, a.hearts, a.reqs,
                       a.B, a.xy_start, a.n_xy, a.layer_k, a.tgrid, a.xy, a.lc);
    return hipGetLastError();
}

template <int LP, int NH>
static void lf_launch(const GroupArgs& a, hipStream_t s)
{
    if (a.lf_phase == 2)
        hipLaunchKernelGGL((lead_field_kernel<LP, NH, 2>), dim3((unsigned)a.n_lf), dim3(512), 0, s, a.hearts,
                           a.reqs, a.B, a.lf_start, a.lead_pos, a.layer_k, a.rec, a.aux);
    else
        hipLaunchKernelGGL((lead_field_kernel<LP, NH, 4>), dim3((unsigned)a.n_lf), dim3(512), 0, s, a.hearts,
                           a.reqs, a.B, a.lf_start, a.lead_pos, a.layer_k, a.rec, a.aux);
}

template <int LP>
static void lf_launch_m(const GroupArgs& a, hipStream_t s)
{
    switch (a.nhood) {
    case 0: lf_launch<LP, 0>(a, s); break;
    case 1: lf_launch<LP, 1>(a, s); break;
    case 2: lf_launch<LP, 2>(a, s); break;
    default: lf_launch<LP, 3>(a, s); break;
    }
}

hipError_t launch_lead_field(const GroupArgs& a, hipStream_t s)
{
    if (a.n_lf == 0) return hipSuccess;
    if (a.lf_phase != 2 && a.lf_phase != 4) return hipErrorInvalidValue;
    switch (a.LP) {
    case 4: lf_launch_m<4>(a, s); break;
    case 8: lf_launch_m<8>(a, s); break;
    case 12: lf_launch_m<12>(a, s); break;
    case 16: lf_launch_m<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LP>
static void sums_launch(const GroupArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((chunk_sums_kernel<LP>), dim3((unsigned)a.n_ch), dim3(256), 0, s, a.hearts, a.reqs, a.B,
                       a.ch_start, a.layer_k, a.lc, a.tgrid, a.thl, a.rec, a.sums, a.bnd);
}

hipError_t launch_chunk_sums(const GroupArgs& a, hipStream_t s)
{
    if (a.n_ch == 0) return hipSuccess;
    switch (a.LP) {
    case 4: sums_launch<4>(a, s); break;
    case 8: sums_launch<8>(a, s); break;
    case 12: sums_launch<12>(a, s); break;
    case 16: sums_launch<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_layer_prefix(const GroupArgs& a, hipStream_t s)
{
    if (a.n_lay == 0) return hipSuccess;
    switch (a.LP) {
#define EKG_PREFIX(LPV)                                                                                  \
    case LPV:                                                                                            \
        hipLaunchKernelGGL(layer_prefix_kernel<LPV>, dim3((unsigned)a.n_lay), dim3(256), 0, s, a.hearts, \
                           a.reqs, a.B, a.lay_start, a.sums, a.bnd, a.pre_rng, a.pre_ab);               \
        break;
        EKG_PREFIX(4)
        EKG_PREFIX(8)
        EKG_PREFIX(12)
        EKG_PREFIX(16)
#undef EKG_PREFIX
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LP>
static void fused_launch(const GroupArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(ecg_fused_kernel<LP>, dim3((unsigned)a.n_ch), dim3(256), 0, s, a.hearts, a.reqs, a.B,
                       a.ch_start, a.rec, a.aux, a.lc, a.layer_k, a.xy, a.thl, a.bnd, a.partial);
}

hipError_t launch_ecg_fused(const GroupArgs& a, hipStream_t s)
{
    if (a.n_ch == 0) return hipSuccess;
    switch (a.LP) {
    case 4: fused_launch<4>(a, s); break;
    case 8: fused_launch<8>(a, s); break;
    case 12: fused_launch<12>(a, s); break;
    case 16: fused_launch<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LP, int NH, int NMAX>
static void chunk_launch3(const GroupArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((ecg_chunk_kernel<LP, NH, NMAX>), dim3((unsigned)a.n_ch), dim3(64 * a.mfma_waves), 0, s,
                       a.hearts, a.reqs, a.B, a.ch_start, a.ch_uniform, a.one_heart, a.lead_pos, a.layer_k, a.lc, a.xy,
                       a.tgrid, a.thl, a.sums, a.bnd, a.partial);
}

template <int LP, int NH>
static void chunk_launch2(const GroupArgs& a, hipStream_t s)
{
    // LDS rows for the group's largest chunk: the smaller the rows, the more
    // blocks share a CU
    if (a.chunk_max <= 128)
        chunk_launch3<LP, NH, 128>(a, s);
    else if (a.chunk_max <= 256)
        chunk_launch3<LP, NH, 256>(a, s);
    else
        chunk_launch3<LP, NH, 512>(a, s);
}

template <int LP>
static void chunk_launch1(const GroupArgs& a, hipStream_t s)
{
    switch (a.nhood) {
    case 0: chunk_launch2<LP, 0>(a, s); break;
    case 1: chunk_launch2<LP, 1>(a, s); break;
    case 2: chunk_launch2<LP, 2>(a, s); break;
    default: chunk_launch2<LP, 3>(a, s); break;
    }
}

hipError_t launch_ecg_chunk(const GroupArgs& a, hipStream_t s)
{
    if (a.n_ch == 0) return hipSuccess;
    // the chunk sums use 128 threads; every chunk fits the kernel's LDS rows
    if (a.mfma_waves < 2 || a.mfma_waves > 16 || a.chunk_max < 1 || a.chunk_max > CHUNK_MAX)
        return hipErrorInvalidValue;
    switch (a.LP) {
    case 4: chunk_launch1<4>(a, s); break;
    case 8: chunk_launch1<8>(a, s); break;
    case 12: chunk_launch1<12>(a, s); break;
    case 16: chunk_launch1<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_class_windows(const GroupArgs& a, hipStream_t s)
{
    if (a.n_cc == 0) return hipSuccess;
    if (a.n_cc > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (a.LP) {
#define EKG_CWIN(LPV)                                                                                        \
    case LPV:                                                                                                \
        hipLaunchKernelGGL(ecg_class_window_kernel<LPV>, dim3((unsigned)a.n_cc), dim3(CC_THREADS), 0, s, a.hearts, \
                           a.reqs, a.cc_runs, a.cc_rstart, a.n_runs, a.lead_pos, a.layer_k, a.sums);       \
        break;
        EKG_CWIN(4)
        EKG_CWIN(8)
        EKG_CWIN(12)
        EKG_CWIN(16)
#undef EKG_CWIN
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_class_sums(const GroupArgs& a, hipStream_t s)
{
    if (a.n_cs == 0) return hipSuccess;
    switch (a.LP) {
#define EKG_CSUMS(LPV)                                                                                  \
    case LPV:                                                                                           \
        hipLaunchKernelGGL(class_sums_kernel<LPV>,                                                     \
                           dim3(grid_for((int64_t)a.cs_max * 3 * LPV, 256), std::min(a.B, 65535)), dim3(256), \
                           0, s, a.hearts, a.reqs, a.B, a.cs_start, a.layer_k, a.sums, a.csums);         \
        break;
        EKG_CSUMS(4)
        EKG_CSUMS(8)
        EKG_CSUMS(12)
        EKG_CSUMS(16)
#undef EKG_CSUMS
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ecg_class(const GroupArgs& a, hipStream_t s)
{
    if (a.n_cl == 0) return hipSuccess;
    if (a.n_cl > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (a.LP) {
#define EKG_CLASS(LPV)                                                                                  \
    case LPV:                                                                                           \
        hipLaunchKernelGGL(ecg_class_kernel<LPV>, dim3((unsigned)a.n_cl), dim3(64 * CL_WAVES), 0, s, a.hearts,  \
                           a.reqs,                                                                        \
                           a.B, a.cl_start, a.cl_per_req, a.layer_k, a.xy, a.tgrid, a.csums, a.ecg);     \
        break;
        EKG_CLASS(4)
        EKG_CLASS(8)
        EKG_CLASS(12)
        EKG_CLASS(16)
#undef EKG_CLASS
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Every kernel of the matrix-core path resolved once by an attribute query,
// so that the first launch of an instantiation (a lead-slot width or
// neighbourhood a long-lived server has not used yet) does not pay the HIP
// runtime's lazy per-kernel load inside a request's latency.
template <int LP, int NH>
static void prime_nh(hipFuncAttributes& at)
{
    (void)hipFuncGetAttributes(&at, (const void*)ecg_chunk_kernel<LP, NH, 128>);
    (void)hipFuncGetAttributes(&at, (const void*)ecg_chunk_kernel<LP, NH, 256>);
    (void)hipFuncGetAttributes(&at, (const void*)ecg_chunk_kernel<LP, NH, 512>);
}

template <int LP>
static void prime_lp(hipFuncAttributes& at)
{
    prime_nh<LP, 0>(at);
    prime_nh<LP, 1>(at);
    prime_nh<LP, 2>(at);
    prime_nh<LP, 3>(at);
    (void)hipFuncGetAttributes(&at, (const void*)ecg_class_window_kernel<LP>);
    (void)hipFuncGetAttributes(&at, (const void*)class_sums_kernel<LP>);
    (void)hipFuncGetAttributes(&at, (const void*)layer_prefix_kernel<LP>);
    (void)hipFuncGetAttributes(&at, (const void*)ecg_reduce_kernel<LP>);
    (void)hipFuncGetAttributes(&at, (const void*)ecg_class_kernel<LP>);
}

void prime_kernels()
{
    hipFuncAttributes at;
    (void)hipFuncGetAttributes(&at, (const void*)ap_tables_kernel);
    prime_lp<4>(at);
    prime_lp<8>(at);
    prime_lp<12>(at);
    prime_lp<16>(at);
}

hipError_t launch_ecg_reduce(const GroupArgs& a, hipStream_t s)
{
    if (a.n_red == 0) return hipSuccess;
    switch (a.LP) {
#define EKG_REDUCE(LPV)                                                                                 \
    case LPV:                                                                                           \
        hipLaunchKernelGGL(ecg_reduce_kernel<LPV>, dim3(grid_for(a.n_red, 256)), dim3(256), 0, s,      \
                           a.hearts, a.reqs, a.B, a.red_start, a.n_red, a.lc, a.xy, a.pre_rng,          \
                           a.pre_ab, a.partial, a.ecg);                                                 \
        break;
        EKG_REDUCE(4)
        EKG_REDUCE(8)
        EKG_REDUCE(12)
        EKG_REDUCE(16)
#undef EKG_REDUCE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ekgk
