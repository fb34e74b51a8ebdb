// bench_plain.cpp -- the plain form of the batched workload, driven from C++
// for bench.py: one oibvh_scene per scene, each frame per scene the scene's
// model matrices (read in place from HBM, oibvh_scene_set_transforms with
// OIBVH_PTR_DEVICE_BORROW) and one oibvh_scene_frame (the reference's loop,
// main.cpp:284: every tree's refit, then Scene::detectCollision). Built as
// libbench_plain.so by oibvh_amd/facade_build.py; bench.py loads it with
// ctypes and times the calls with HIP events on the library's stream.
#include <cstdint>

#include "oibvh.h"

extern "C" {

// n_frames frames f = f0, f0 + 1, ... of every scene. Frame f's matrices of
// scene s start at mats + (f % n_cycle) * frame_stride + mat_off[s] (16
// floats per tree, in the scene's insertion order). flags as for
// oibvh_scene_frame (OIBVH_DETECT_ASYNC is added). Returns the first failing
// status (0 if none) and the scene in *failed.
int plain_frames(oibvh_scene* const* scenes, uint32_t n_scenes, const float* mats, const uint64_t* mat_off,
                 uint64_t frame_stride, uint32_t f0, uint32_t n_frames, uint32_t n_cycle, uint32_t entry,
                 uint32_t expand, uint32_t flags, uint32_t* failed)
{
    for (uint32_t k = 0; k < n_frames; k++) {
        const float* base = mats + (uint64_t)((f0 + k) % n_cycle) * frame_stride;
        for (uint32_t s = 0; s < n_scenes; s++) {
            oibvh_status st = oibvh_scene_set_transforms(scenes[s], base + mat_off[s], OIBVH_PTR_DEVICE_BORROW);
            if (st == OIBVH_OK) st = oibvh_scene_frame(scenes[s], nullptr, entry, expand, flags | OIBVH_DETECT_ASYNC);
            if (st != OIBVH_OK) {
                if (failed) *failed = s;
                return (int)st;
            }
        }
    }
    return 0;
}

}  // extern "C"
