// mh_flare.h - the flare post-pass launcher (mh_flare.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace mh {

// Per-call constants, prepared on the host: clamped threshold, intensity,
// tint, and per scale the recurrence coefficient alpha = expf(-1 / radius)
// (the host's expf, as the reference computes it) and weight * 0.045.
struct FlareK {
    double intensity, thresh;
    double tint[3];
    float a1, a2, n1, n2;
};

// LDS bytes of one scanline's 15 float planes (3 sources, 12 recurrences).
inline size_t flare_lds_bytes(int w) { return (size_t)w * 15 * sizeof(float); }

// One workgroup (64 lanes) per row, grid-strided over rows. scratch == null:
// the row planes in LDS (flare_lds_bytes(w) must fit the device's limit);
// else grid * 15 * w floats of device scratch.
hipError_t launch_flare(unsigned char* rgb, int w, int h, const FlareK& k, float* scratch, int grid, hipStream_t st);

} // namespace mh
