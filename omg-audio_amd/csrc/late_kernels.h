// Device side of the late-field state: EMA of fresh echograms into the
// slots' resident averages and the per-slot reductions behind
// estimate_reverb and late_direction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace omg {

constexpr uint32_t kLateMaxBands = 8;

// One slot update on the device.
struct LateJob {
    float* avg_bins;        // the slot's [nbins][nbands]
    float* avg_dirs;        // the slot's [nbins][3]
    const float* cur_bins;  // the fresh echogram, or null when not traced
    const float* cur_dirs;
    float alpha;
    uint32_t traced;
};

// Reductions of one slot's average, finished on the host.
struct LateRed {
    double total[kLateMaxBands];  // backward energy integral at bin 0
    double edc5[kLateMaxBands];   // integral at the -5 dB bin
    double edc25[kLateMaxBands];  // integral at the -25 dB bin
    int32_t t5[kLateMaxBands];    // first bin at or below -5 dB, or -1
    int32_t t25[kLateMaxBands];   // first bin at or below -25 dB, or -1
    float level[kLateMaxBands];   // min(sqrt(late energy), 0.7)
    float dir[3];
    float aniso;
};

struct LateArgs {
    const LateJob* jobs;
    uint32_t n;
    uint32_t nbins, nbands;
    uint32_t late_start;  // first bin of the late level (0.080 s)
    uint32_t dir_start;   // first bin of the late direction (0.08 s)
    double r5, r25;       // decay ratios at which 10 log10 reaches -5 / -25 dB
    LateRed* out;
};

hipError_t launch_late_ema(const LateArgs& A, hipStream_t s);
hipError_t launch_late_reduce(const LateArgs& A, hipStream_t s);

}  // namespace omg
