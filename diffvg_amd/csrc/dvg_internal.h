// dvg_internal.h -- host-side scene object of the library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/diffvg_amd.h"
#include "dvg_scene.h"

namespace dvg {

struct EdgeSample;

int set_error(const std::string &msg);  // records the message, returns 1

// Host-to-device copy / clear on the device's non-blocking upload stream (synchronous
// for the caller, independent of work in flight on other streams).
int upload(int device, void *dst, const void *src, size_t bytes);
int clear_device(int device, void *dst, size_t bytes);
// Fills the scene's per-segment tables that derive from its points (the cubic winding
// prep) on stream `st` (NULL: synchronously).
struct SceneHost;
int prep_segments(SceneHost &sc, hipStream_t st);

// Per-scene scratch reused across renders of the same or smaller size.
struct Workspace {
    size_t npix_cap = 0, nsamp_cap = 0, ntiles_cap = 0;
    float *weight = nullptr;
    float4 *colors = nullptr;
    EdgeSample *edges = nullptr;
    int *edge_tile = nullptr;
    int *edge_sorted = nullptr;
    int *tile_count = nullptr;
    int *tile_off = nullptr;
    int *edge_rank = nullptr;   // [nsamp] position of each boundary sample in its tile's bin
    int *chunk_off = nullptr;  // [ntiles + 1] boundary-sample chunk offsets per tile
    int reserve(size_t npix, size_t nsamp, size_t ntiles);
    void release();
};

// Device memory cache: blocks freed by destroyed scenes are kept per device and size
// class and handed to later scenes, so constructing a scene per forward call (as the
// autograd front end does) does not pay hipMalloc/hipFree.
void *cache_alloc(int device, size_t bytes, size_t *granted);
void cache_free(int device, void *p, size_t bytes);

// One render workspace per device, shared by all scenes.  A render holds it from the
// first to the last kernel it enqueues; its stream first waits for the event recorded
// by the previous user, so renders on different streams are ordered correctly.
Workspace &acquire_workspace(int device, hipStream_t st);
void release_workspace(int device, hipStream_t st);

// Topology tables and scratch of the device rebuild (dvg_scene_update), in the scene
// block next to the arrays it rewrites.
struct RefitTables {
    const int *seg_perm = nullptr;     // [nseg] per path: sorted-leaf positions in visit order
    const int *g_ids_desc = nullptr;   // [n_ids] group members in description order
    const int *g_perm = nullptr;       // [n_ids] per group: sorted-leaf positions in visit order
    const int *g_anchor = nullptr;     // [ng] shape whose first point anchors the SDF bound, or -1
    const int *g_sdf_flags = nullptr;  // [ng] bld::kSdfFindable | bld::kSdfApproxQuad
    float *thick_leaf0 = nullptr;      // [ns] radius of leaf 0 of a thickness path's segment BVH
    int *order = nullptr;              // [nseg] sort scratch
    int *gorder = nullptr;             // [n_ids] sort scratch
    const int *long_ids = nullptr;     // [n_long] paths of more than kRefitSegCap segments
    int n_long = 0;
    const int *big_gids = nullptr;     // [n_big] groups of more than kRefitGroupLane members
    int n_big = 0;
};

// the device rebuild sorts the members of a group of up to kRefitGroupLane shapes on one
// lane, of a larger group with a workgroup
constexpr int kRefitGroupLane = 64;

// the device rebuild gives a path of up to kRefitSegCap segments one wave, a longer one
// a workgroup
constexpr int kRefitSegCap = 512;

constexpr int kErrInts = 16;  // ints of SceneHost::err

struct SceneHost {
    int device = 0;
    bool device_set = false;      // a device was chosen and memory may be held on it
    int mode = DVG_MODE_TILED;
    int tile_w = 0;               // pixel tile width of the TILED schedule (0 = square default)
    int edge_spw = 32;            // boundary samples per wave of the edge kernel: 32 or 64
    int cw = 0, ch = 0, num_shapes = 0, num_groups = 0;
    int num_shape_params = 0, num_points_total = 0, num_color_params = 0;
    void *block = nullptr;        // scene arrays
    size_t block_bytes = 0, block_granted = 0;
    float *grad = nullptr;        // gradient arrays
    size_t grad_floats = 0, grad_granted = 0;
    size_t g_params = 0, g_points = 0, g_thick = 0, g_sw = 0, g_color = 0, g_xform = 0, g_filter = 0;
    // [0] render conditions, [1] conditions found by dvg_scene_update, [2] parameter
    // version on the device (advanced by every update), [4..9] and [10..15] the key of the
    // kept fragment lists (non-prefiltered, prefiltered) a forward last wrote
    int *err = nullptr;
    hipStream_t last_stream = nullptr;
    RefitTables refit;
    size_t top_offset = 0, top_bytes = 0;          // top-level BVH in the block
    size_t scratch_offset[2] = {0, 0}, scratch_bytes[2] = {0, 0};  // rebuild sort scratch in the block
    bool top_stale = false;       // the top-level BVH (PLAIN schedule) predates a device update
    uint64_t version = 0;         // advanced by every device update of the parameters
    // Fragment lists of the last non-prefiltered colour forward of the TILED schedule:
    // [kFragSlots][nsamp] 16-bit group ids (2 group + is_stroke) in compositing order,
    // then [nsamp] fragment counts.  The backward of the same render reads them instead
    // of shading its samples again.
    uint16_t *frag = nullptr;
    size_t frag_granted = 0;
    bool frag_valid = false;
    uint64_t frag_key[4] = {0, 0, 0, 0};
    hipStream_t frag_stream = nullptr;
    // the same for the prefiltered forward: per fragment the words its adjoint needs
    int *pkeep = nullptr;
    size_t pkeep_granted = 0;
    bool pkeep_valid = false;
    uint64_t pkeep_key[4] = {0, 0, 0, 0};
    hipStream_t pkeep_stream = nullptr;
    int keep_refs = 0;            // batches holding the kept-list buffers: no reallocation
    DevScene dev{};
};

// Batched renders (one kernel per stage over the jobs of many scenes).
struct BatchHost;
// flags: DVG_BATCH_ZERO_GRADS, DVG_BATCH_PERSISTENT
int batch_create(SceneHost *const *scenes, const dvg_render_args *args, int n, int flags, BatchHost **out);
int batch_render(BatchHost &b, hipStream_t st);
int batch_status(BatchHost &b);
void batch_destroy(BatchHost *b);
const std::vector<SceneHost *> &batch_scenes(const BatchHost &b);

int build_scene(const dvg_scene_desc &d, int gpu_index, SceneHost &sc);
int update_scene(SceneHost &sc, const dvg_scene_params &p, float filter_radius, hipStream_t st);
int scene_status(SceneHost &sc);
void free_scene(SceneHost &sc);
// `fixed`: a workspace the caller carved (the allocation-free entry point); the render
// then neither grows a workspace nor keeps fragment lists for its backward.
int render_tiled(SceneHost &sc, const dvg_render_args &a, Workspace *fixed = nullptr);
// dvg_render_noalloc: the TILED render with its workspace in a fixed device array.
int render_noalloc(SceneHost &sc, const dvg_render_args &a);
int render_plain(SceneHost &sc, const dvg_render_args &a);
// Rebuilds the PLAIN schedule's top-level BVH from the scene's current group boxes (after
// dvg_scene_update); synchronous.
int rebuild_top(SceneHost &sc);

}  // namespace dvg
