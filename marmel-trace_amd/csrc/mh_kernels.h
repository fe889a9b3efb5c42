// mh_kernels.h - host-callable launchers of the render / query kernels.
// Each launcher is instantiated once per arithmetic type, in its own
// translation unit (mh_kernels_f32.hip is built with FMA contraction,
// mh_kernels_f64.hip without, so the double path rounds like the reference).
#pragma once

#include "marmel_hip.h"
#include "mh_scene_dev.h"

namespace mh {

template <typename R> struct WItem;

struct LaunchOut {
    float* linear;               // w*h*3 mean radiance (device)
    unsigned char* rgb;          // w*h*3 tone-mapped bytes (device)
    unsigned long long* rays;    // ray-query counter (device)
    int* overflow;               // Whitted stack overflow flag (device)
};

template <typename R>
hipError_t launch_pathtrace(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                            const LaunchOut& out, int block, hipStream_t st);

// Whitted: persistent grid of `threads` lanes, each with `cap` stack items
// of `item_bytes` at stack + lane * cap * item_bytes.
template <typename R>
hipError_t launch_whitted(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                          const LaunchOut& out, void* stack, int cap, int threads, hipStream_t st);

template <typename R> size_t whitted_item_bytes();

template <typename R>
hipError_t launch_closest(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, mh_hit* out,
                          hipStream_t st);

template <typename R>
hipError_t launch_any(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, uint8_t* out,
                      hipStream_t st);

} // namespace mh
