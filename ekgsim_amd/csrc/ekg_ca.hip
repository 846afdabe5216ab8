// ekg_ca.hip -- CDNA4 (gfx950) kernels of the excitation sequence:
// min-plus relaxation over LDS-staged 8^3 voxel bricks with a 1-voxel halo,
// bit-exact with the reference's priority-queue flood (simulator.cpp:212-286).
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

}  // namespace ekgk
