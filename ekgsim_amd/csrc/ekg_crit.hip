// ekg_crit.hip -- FP64 comparison criteria of simulated ECGs against targets,
// on the device right after the ECG reduction, so that a batch of
// evaluations returns only its criteria.
//
// Per (request, lead) the value of SimImplementation::calculateFitness
// (sim.cpp:600-702) for that lead: the comparison of the lead with its target
// at offset 0 (vectorMath.h:277-401), in the reference's loop order -- each
// thread runs the host restatement's loops (ekg_model.cpp::criterion) one for
// one with floating-point contraction off, so the values equal the host's bit
// for bit; the target's own sums come from target_stats_kernel, computed
// once per set of targets by the same loops -- and, when peaks are criteria,
// |argmax(sim) - argmax(target)|
// (first maxima, sim.cpp:662-670).  Values that are negative, zero,
// subnormal or not finite become 2e10 (sim.cpp:679-685) unless raw output is
// asked for.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ekg.h"
#include "ekg_kernels.hpp"

namespace ekgk {

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double range_sum(const double* v, int64_t a, int64_t b)
{
    double s = 0;
#pragma unroll 8
    for (int64_t i = a; i < b; ++i) s += v[i];
    return s;
}

// mean and variance in the order of vectorMath.h:145-186
__device__ __forceinline__ void mean_var(double& mean, double& var, const double* v, int64_t n)
{
    const double inv = 1.0 / double(n);
    mean = range_sum(v, 0, n) * inv;
    var = 0;
#pragma unroll 8
    for (int64_t i = 0; i < n; ++i) {
        const double d = v[i] - mean;
        var += d * d;
    }
    var *= inv;
}

// ts: the target's statistics over the comparison length (target_stats_kernel):
// {mean, variance, b . b, first maximum}
__device__ __forceinline__ double compare(int mode, const double* a, int64_t na, const double* b, int64_t nb, double off,
                                          const double* ts)
{
    const int64_t len = na < nb ? na : nb;
    switch (mode) {
    case EKG_CMP_CORRELATION: {
        double ma, va;
        mean_var(ma, va, a, len);
        const double mb = ts[0], vb = ts[1];
        double cov = 0;
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) cov += (a[i] - ma) * (b[i] - mb);
        cov /= double(len);
        return 1.0 - cov / (sqrt(va) * sqrt(vb));
    }
    case EKG_CMP_VECTOR_CORRELATION: {
        double ab = 0, aa = 0;
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) ab += a[i] * b[i];
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) aa += a[i] * a[i];
        const double bb = ts[2];
        return 1.0 - ab / (sqrt(aa) * sqrt(bb));
    }
    case EKG_CMP_RMS: {
        // sqrt of the mean of (a - b)^2; the variance the reference also
        // computes does not enter the value
        const double inv = 1.0 / double(len);
        double s = 0;
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) {
            const double d = a[i] - b[i];
            s += d * d;
        }
        return sqrt(s * inv);
    }
    case EKG_CMP_DEV_FROM_LINEAR: {
        double lo = 0, hi = 0;
        if (na > 0) {
            lo = hi = a[0];
            for (int64_t i = 1; i < na; ++i) {
                if (a[i] < lo)
                    lo = a[i];
                else if (a[i] > hi)
                    hi = a[i];
            }
        }
        const double scale = 1.0 / (hi - lo);
        const double inv = 1.0 / double(len);
        double s = 0;
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) s += (a[i] * scale + off) / (b[i] + off);
        const double mean = s * inv;
        double var = 0;
#pragma unroll 8
        for (int64_t i = 0; i < len; ++i) {
            const double d = (a[i] * scale + off) / (b[i] + off) - mean;
            var += d * d;
        }
        var *= inv;
        return var == 0 ? 1 / 1e-30 : sqrt(var);
    }
    default:
        return __builtin_nan("");
    }
}

// std::max_element: the first index whose value no later value exceeds (the
// running maximum is kept in a register; it compares the same values)
__device__ __forceinline__ int64_t first_max(const double* v, int64_t n)
{
    int64_t m = 0;
    double best = v[0];
#pragma unroll 8
    for (int64_t i = 1; i < n; ++i) {
        const double x = v[i];
        if (best < x) {
            best = x;
            m = i;
        }
    }
    return m;
}

}  // namespace

// Per target i, over the comparison length len: {mean, variance} as
// mean_var computes them, b . b in the order of the vector correlation, and
// the first maximum over the whole target -- the target's part of every
// comparison, the same for all requests.
__global__ void target_stats_kernel(const double* __restrict__ targets, int n_targets, int64_t target_len,
                                    int64_t len, double* __restrict__ stats)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_targets) return;
    const double* b = targets + (int64_t)i * target_len;
    double mb, vb;
    mean_var(mb, vb, b, len);
    double bb = 0;
#pragma unroll 8
    for (int64_t q = 0; q < len; ++q) bb += b[q] * b[q];
    double* o = stats + 4 * i;
    o[0] = mb;
    o[1] = vb;
    o[2] = bb;
    o[3] = (double)first_max(b, target_len);
}

// out[r*stride + c]: c = i (or 2i, 2i+1 with peaks) for lead i < n_crit; one
// thread per (request, lead), all following the same loops.
__global__ void criteria_kernel(int mode, int peaks, int raw, const double* __restrict__ ecg, int B, int L,
                                int64_t T, const double* __restrict__ targets, const double* __restrict__ offsets,
                                const double* __restrict__ stats, int n_crit, int64_t target_len,
                                double* __restrict__ out)
{
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)B * n_crit) return;
    const int r = (int)(idx / n_crit), i = (int)(idx % n_crit);
    const double* a = ecg + ((int64_t)r * L + i) * T;
    const double* b = targets + (int64_t)i * target_len;
    const double* ts = stats + 4 * i;
    double v = compare(mode, a, T, b, target_len, offsets[i], ts);
    if (!raw && (v < 0 || !isfinite(v) || !(fabs(v) >= 2.2250738585072014e-308))) v = 2e10;
    const int stride = peaks ? 2 * n_crit : n_crit;
    double* o = out + (int64_t)r * stride;
    if (peaks) {
        const int64_t sp = first_max(a, T), tp = (int64_t)ts[3];
        o[2 * i] = v;
        o[2 * i + 1] = (double)(sp > tp ? sp - tp : tp - sp);
    } else {
        o[i] = v;
    }
}

hipError_t launch_target_stats(const double* targets, int n_targets, int64_t target_len, int64_t len, double* stats,
                               hipStream_t s)
{
    if (n_targets < 1) return hipSuccess;
    hipLaunchKernelGGL(target_stats_kernel, dim3((unsigned)((n_targets + 63) / 64)), dim3(64), 0, s, targets,
                       n_targets, target_len, len, stats);
    return hipGetLastError();
}

hipError_t launch_criteria(int mode, int peaks, int raw, const double* ecg, int B, int L, int64_t T,
                           const double* targets, const double* offsets, const double* stats, int n_crit,
                           int64_t target_len, double* out, hipStream_t s)
{
    const int64_t n = (int64_t)B * n_crit;
    if (n == 0) return hipSuccess;
    if (n > (int64_t)1 << 30) return hipErrorInvalidValue;
    hipLaunchKernelGGL(criteria_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, mode, peaks, raw, ecg, B, L,
                       T, targets, offsets, stats, n_crit, target_len, out);
    return hipGetLastError();
}

}  // namespace ekgk
