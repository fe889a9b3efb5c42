// mh_build.h - device BVH build and device-side scene layout.
#pragma once

#include "marmel_hip.h"
#include "mh_scene_dev.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mh {

// Device allocations made by a build; the caller owns them (hipFree).
struct GpuBvh {
    std::vector<void*> owned;   // every allocation below, for the caller to free
    size_t bytes = 0;           // their total size
    // binary tree (node 0 = root) in the ABI record, and prim index per leaf slot
    mh_bvh_node* nodes = nullptr;
    int n_nodes = 0;
    int32_t* leaf_prims = nullptr;
    int n_leaf = 0;
    int max_depth = 0;
    int n_leaves = 0;
    // device layout
    DNode* dnodes = nullptr;
    int n_inner = 0;
    int root_ref = INT32_MIN;
    DNode4* dnodes4 = nullptr;
    int n_nodes4 = 0;
    int root4 = INT32_MIN;
    int depth4 = 0;
    P32* p32 = nullptr;
    P64* p64 = nullptr;
    DAttr<float>* a32 = nullptr;
    DAttr<double>* a64 = nullptr;
    P32* pl32 = nullptr;
    P64* pl64 = nullptr;
    int32_t* planes = nullptr;
    int n_planes = 0;
};

// Build the tree of n prims (device mh_prim array, file order; attrs device
// array or null) with binned SAH over `bins` bins on all three axes and
// leaves of at most max_leaf prims -- the decisions mt_scene_build_bvh makes
// for the same parameters -- then lay the scene out for the kernels (DNode,
// DNode4, leaf-ordered fp32 / fp64 records). Runs on `st`; returns 0 or a
// hipError_t; *err_depth is set when the tree exceeds 60 levels.
int gpu_build_scene(const mh_prim* d_prims, const mh_prim_attr* d_attrs, int n, bool any_attr, int max_leaf,
                    int bins, hipStream_t st, GpuBvh* out, int* err_depth);

} // namespace mh
