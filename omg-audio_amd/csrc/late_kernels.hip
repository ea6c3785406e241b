// Late-field kernels for gfx950: the slots' running averages live in HBM;
// a fresh echogram is folded in elementwise, and every updated slot is
// reduced by one wave.  The sums run in the reference's order (sequential
// per band or per component), so they are bit-identical to the CPU.
#include "late_kernels.h"

#pragma clang fp contract(off)

namespace omg {

// Echogram::ema (crates/omg-core/src/tracer.rs:46-57): a += alpha * (o - a)
// on every bin and direction component.  One block per traced slot.
__global__ __launch_bounds__(256) void k_late_ema(LateArgs A)
{
    const LateJob J = A.jobs[blockIdx.x];
    if (!J.traced) return;
    const uint32_t nb = A.nbins * A.nbands, nd = A.nbins * 3u;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        J.avg_bins[i] += J.alpha * (J.cur_bins[i] - J.avg_bins[i]);
    for (uint32_t i = threadIdx.x; i < nd; i += blockDim.x)
        J.avg_dirs[i] += J.alpha * (J.cur_dirs[i] - J.avg_dirs[i]);
}

// The largest x in [0, total] whose quotient fl(x / total) is at most r:
// bisection over the ordered bit patterns of non-negative doubles.  The
// quotient never decreases with x, so `x <= bound` is exactly
// `x / total <= r` and the per-bin test needs no division.
__device__ double quotient_bound(double total, double r)
{
    uint64_t lo = 0, hi = (uint64_t)__double_as_longlong(total);
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (__longlong_as_double((long long)mid) / total <= r) lo = mid;
        else hi = mid;
    }
    return __longlong_as_double((long long)lo);
}

// One wave per slot.  Lane b < nbands: estimate_reverb's band b
// (tracer.rs:205-257) - the f64 backward integral, then the same
// integral again from the last bin down, keeping the lowest bins whose
// ratio to the total is at or below the -5 / -25 dB ratios (the ratio
// never increases with the bin, so those are the first bins the reference
// finds scanning forward), and the f32 late-energy sum.  Lane nbands:
// late_direction (tracer.rs:62-77).  Every sum runs in the reference's
// order; the loops are unrolled so that loads run ahead of the adds.
__global__ __launch_bounds__(64) void k_late_reduce(LateArgs A)
{
    const LateJob J = A.jobs[blockIdx.x];
    LateRed& R = A.out[blockIdx.x];
    const uint32_t lane = threadIdx.x, nb = A.nbands, n = A.nbins;
    const float* bins = J.avg_bins;
    if (lane < nb) {
        double total = 0.0;
#pragma unroll 8
        for (int64_t i = (int64_t)n - 1; i >= 0; i--) total += (double)bins[(size_t)i * nb + lane];
        int32_t t5 = -1, t25 = -1;
        double e5 = 0.0, e25 = 0.0;
        if (total > 0.0) {
            const double x5 = quotient_bound(total, A.r5), x25 = quotient_bound(total, A.r25);
            double e = 0.0;
#pragma unroll 8
            for (int64_t i = (int64_t)n - 1; i >= 0; i--) {
                e += (double)bins[(size_t)i * nb + lane];
                if (e <= x5) {
                    t5 = (int32_t)i;
                    e5 = e;
                }
                if (e <= x25) {
                    t25 = (int32_t)i;
                    e25 = e;
                }
            }
        }
        float late_e = 0.0f;
#pragma unroll 8
        for (uint32_t i = A.late_start; i < n; i++) late_e += bins[(size_t)i * nb + lane];
        R.total[lane] = total;
        R.edc5[lane] = e5;
        R.edc25[lane] = e25;
        R.t5[lane] = t5;
        R.t25[lane] = t25;
        R.level[lane] = fminf(sqrtf(late_e), 0.7f);
    } else if (lane == nb) {
        const float* dirs = J.avg_dirs;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, e = 0.0f;
#pragma unroll 8
        for (uint32_t i = A.dir_start; i < n; i++) {
            v0 += dirs[(size_t)i * 3];
            v1 += dirs[(size_t)i * 3 + 1];
            v2 += dirs[(size_t)i * 3 + 2];
            e += bins[(size_t)i * nb + 1];
        }
        const float len = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
        if (len < 1e-12f || e < 1e-12f) {
            R.dir[0] = 1.0f;
            R.dir[1] = 0.0f;
            R.dir[2] = 0.0f;
            R.aniso = 0.0f;
        } else {
            R.dir[0] = v0 / len;
            R.dir[1] = v1 / len;
            R.dir[2] = v2 / len;
            R.aniso = fminf(len / e, 1.0f);
        }
    }
}

hipError_t launch_late_ema(const LateArgs& A, hipStream_t s)
{
    if (A.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_late_ema, dim3(A.n), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_late_reduce(const LateArgs& A, hipStream_t s)
{
    if (A.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_late_reduce, dim3(A.n), dim3(64), 0, s, A);
    return hipGetLastError();
}

}  // namespace omg
