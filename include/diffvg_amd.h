/*
 * diffvg_amd.h -- C ABI of the MI355X render path of diffvg.
 *
 * The library replaces the native half of diffvg's render path: the pybind11
 * surface `diffvg.Scene(...)`, `diffvg.render(...)` and the per-shape gradient
 * getters (reference diffvg.cpp:1651-1792), backed by hand-written HIP kernels for
 * gfx950.  No torch types appear here: scenes are described by flat host arrays,
 * images by device pointers, and every call returns 0 on success or a non-zero
 * code with the message in dvg_last_error().  The library never aborts the process.
 */
#ifndef DIFFVG_AMD_H
#define DIFFVG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DVG_ABI_VERSION 1

/* ShapeType (reference shape.h:9-14), ColorType (color.h:7-11), FilterType
 * (filter.h:6-11) plus the Gaussian filter, which the reference does not have. */
enum { DVG_SHAPE_CIRCLE = 0, DVG_SHAPE_ELLIPSE = 1, DVG_SHAPE_PATH = 2, DVG_SHAPE_RECT = 3 };
enum { DVG_COLOR_NONE = -1, DVG_COLOR_CONSTANT = 0, DVG_COLOR_LINEAR = 1, DVG_COLOR_RADIAL = 2 };
enum { DVG_FILTER_BOX = 0, DVG_FILTER_TENT = 1, DVG_FILTER_PARABOLIC = 2, DVG_FILTER_HANN = 3,
       DVG_FILTER_GAUSSIAN = 4 };

/* Render schedules.  TILED is the MI355X path (tile-binned candidates, one wavefront
 * per pixel tile, gathered filter resolve, tile-binned edge samples).  PLAIN keeps the
 * reference's launch structure (one thread per sample in 64-thread blocks, per-sample
 * top-level BVH traversal, atomic splats, Morton-sorted edge samples, per-call
 * allocations, synchronous) and serves as the comparison baseline. */
enum { DVG_MODE_TILED = 0, DVG_MODE_PLAIN = 1 };

/* Flat scene description: the fields of Scene(canvas_width, canvas_height, shapes,
 * shape_groups, filter) (reference scene.h:19-25, shape.h:16-169, color.h:13-63)
 * laid out as structure-of-arrays.  All pointers are HOST pointers and are copied
 * during dvg_scene_create.
 *   circle params  : radius, cx, cy          ellipse params: rx, ry, cx, cy
 *   rect params    : x0, y0, x1, y1          path: num_control_points[], points[] (x,y)
 *   constant paint : r, g, b, a
 *   linear paint   : bx, by, ex, ey, offsets[n], stop colours[4n]
 *   radial paint   : cx, cy, rx, ry, offsets[n], stop colours[4n]
 *   shape_to_canvas: 9 floats per group, row-major. */
typedef struct dvg_scene_desc {
    int32_t canvas_width, canvas_height;
    int32_t num_shapes;
    const int32_t *shape_type;
    const float *stroke_width;
    const int32_t *shape_param_offset;   /* -1 for paths */
    const float *shape_params;
    int32_t num_shape_params;
    const int32_t *path_ctrl_offset;     /* per shape, offset into num_control_points */
    const int32_t *path_num_base_points; /* per shape, number of segments */
    const int32_t *path_point_offset;    /* per shape, offset (in points) into points */
    const int32_t *path_num_points;
    const uint8_t *path_has_thickness;   /* thickness[] indexed like points */
    const uint8_t *path_is_closed;
    const uint8_t *path_use_distance_approx;
    const int32_t *num_control_points;
    int32_t num_ctrl_total;
    const float *points;
    const float *thickness;              /* may be NULL when no path has thickness */
    int32_t num_points_total;
    int32_t num_groups;
    const int32_t *group_shape_offset;
    const int32_t *group_num_shapes;
    const int32_t *group_shape_ids;
    int32_t num_group_shape_ids;
    const int32_t *fill_type;            /* DVG_COLOR_* */
    const int32_t *fill_offset;
    const int32_t *fill_num_stops;
    const int32_t *stroke_type;
    const int32_t *stroke_offset;
    const int32_t *stroke_num_stops;
    const float *color_params;
    int32_t num_color_params;
    const uint8_t *use_even_odd_rule;
    const float *shape_to_canvas;
    int32_t filter_type;
    float filter_radius;
} dvg_scene_desc;

/* Arguments of render() (reference diffvg.cpp:1477-1492).  Image pointers are
 * DEVICE pointers on the scene's GPU (NULL = absent); layouts are the reference's:
 * images [H][W][4], sdf [H][W][1] or [N][1], d_translation [H][W][2], eval_positions
 * [N][2].  `stream` is a hipStream_t (NULL = the null stream). */
typedef struct dvg_render_args {
    const float *background_image;
    float *render_image;
    float *render_sdf;
    int32_t width, height, num_samples_x, num_samples_y;
    uint64_t seed;
    float *d_background_image;
    const float *d_render_image;
    const float *d_render_sdf;
    float *d_translation;
    int32_t use_prefiltering;
    const float *eval_positions;
    int32_t num_eval_positions;
    void *stream;
} dvg_render_args;

/* Gradient destinations, HOST pointers, in the layout of dvg_scene_desc (each may be
 * NULL to skip).  Replaces Scene.get_d_shape / get_d_shape_group /
 * get_d_filter_radius (reference scene.cpp:1025-1035, diffvg.cpp:1786-1788). */
typedef struct dvg_grads {
    float *d_shape_params;     /* [num_shape_params] */
    float *d_points;           /* [2 * num_points_total] */
    float *d_thickness;        /* [num_points_total] */
    float *d_stroke_width;     /* [num_shapes] */
    float *d_color_params;     /* [num_color_params] */
    float *d_shape_to_canvas;  /* [9 * num_groups] */
    float *d_filter_radius;    /* [1] */
} dvg_grads;

typedef struct dvg_scene dvg_scene;

/* Scene(...) (reference scene.cpp:919-998, bound at diffvg.cpp:1777-1789): copies the
 * description to the device of `gpu_index` (-1 = current device), derives boundary
 * lengths, sampling CDFs, boxes and the three BVH levels, and zeroes the gradients. */
int dvg_scene_create(const dvg_scene_desc *desc, int gpu_index, dvg_scene **out);
int dvg_scene_destroy(dvg_scene *scene);

/* render(...) (reference diffvg.cpp:1477-1649).  Forward when render_image/render_sdf
 * are set; backward when d_render_image/d_render_sdf are set, accumulating into the
 * scene's gradient buffers.  Asynchronous on args->stream in DVG_MODE_TILED (conditions
 * the reference asserts on -- fragment overflow, prefiltered circle strokes -- are then
 * reported by dvg_scene_status and dvg_scene_get_grads); synchronous in DVG_MODE_PLAIN,
 * like the reference. */
int dvg_render(dvg_scene *scene, const dvg_render_args *args);

/* dvg_render in DVG_MODE_TILED (whatever the scene's mode) that allocates no device
 * memory: no reference counterpart.  For callers that cannot allocate (code under HIP
 * graph capture, tight serving loops).  Its workspace is a fixed array in device memory,
 * one per GPU, held from the first call on that GPU until the process ends; renders of
 * at most 524,288 pixels and 524,288 samples (at most 256 per pixel), larger ones fail
 * with an error.  Every call on a GPU shares that GPU's array: calls that overlap in
 * time on one GPU (on different streams, or with dvg_render_noalloc calls of other
 * threads) race and must not be made.  The forward keeps no fragment lists for its
 * backward (the backward shades its samples again); results equal dvg_render's. */
int dvg_render_noalloc(dvg_scene *scene, const dvg_render_args *args);

/* Copies the accumulated gradients to host memory (synchronises the scene's last
 * stream first and fails with the conditions dvg_scene_status reports). */
int dvg_scene_get_grads(dvg_scene *scene, dvg_grads *out);
/* Layout of the scene's gradient block on the device: float offsets of the seven
 * arrays of dvg_grads (in that order) and the block size in floats. */
int dvg_scene_grad_layout(dvg_scene *scene, int64_t *offsets, int64_t *total_floats);
/* Copies the gradient block to device memory `dst` (total_floats floats), stream-ordered
 * on `stream`: the device-resident counterpart of dvg_scene_get_grads. */
int dvg_scene_copy_grads(dvg_scene *scene, float *dst, void *stream);
/* Clears the gradient buffers (stream-ordered on `stream`). */
int dvg_scene_zero_grads(dvg_scene *scene, void *stream);
/* Selects DVG_MODE_TILED (default) or DVG_MODE_PLAIN for later renders. */
int dvg_scene_set_mode(dvg_scene *scene, int mode);
/* Pixel tile width of the TILED schedule for later renders and batches of this scene
 * (a power of two; the tile height follows from 256 samples per tile; 0 restores the
 * square default).  Picked by the autotuner; results do not depend on it beyond float
 * summation order.  No reference counterpart (the reference has no tiles). */
int dvg_scene_set_tiling(dvg_scene *scene, int tile_width);
/* Boundary samples per wavefront of the TILED edge kernel for later renders and batches
 * of this scene: 32 (each lane shades one of the two colour queries of a sample; the
 * default, also chosen by 0) or 64 (each lane shades both).  Picked by the autotuner;
 * results do not depend on it beyond float summation order.  No reference counterpart. */
int dvg_scene_set_samples_per_wave(dvg_scene *scene, int edge_samples_per_wave);

/* Device-resident parameter values of a scene: DEVICE pointers on the scene's GPU, in the
 * layout of the float arrays of dvg_scene_desc (points as x,y pairs; thickness 0 where a
 * path has none; stroke width 0 for paths with thickness).  NULL keeps an array's
 * current values. */
typedef struct dvg_scene_params {
    const float *shape_params;    /* [num_shape_params] */
    const float *points;          /* [2 * num_points_total] */
    const float *thickness;       /* [num_points_total] */
    const float *stroke_width;    /* [num_shapes] */
    const float *color_params;    /* [num_color_params] */
    const float *shape_to_canvas; /* [9 * num_groups] */
} dvg_scene_params;

/* Rebuilds the scene on the device for new parameter values with the same topology
 * (shape kinds, segment counts, group members, paint kinds): the equivalent of
 * dvg_scene_create from a description that differs only in its float arrays, as the
 * reference constructs a new Scene for every forward pass (render_pytorch.py:366-368,
 * scene.cpp:919-998), without a host round trip.  Enqueued on `stream`: the values are
 * copied into the scene, then kernels recompute boxes, boundary lengths, sampling CDFs,
 * canvas_to_shape, and the segment and group BVH leaf orders with the reference's sort
 * (so results equal those of a scene created from the same values).  The gradient
 * buffers are not touched.  Conditions Scene construction rejects (non-finite points or
 * lengths, a total boundary length <= 0) are reported by dvg_scene_status.  The PLAIN
 * schedule's top-level BVH (scene.cpp:648-683) is rebuilt on the host from the updated
 * group boxes by the next PLAIN render of the scene, which is synchronous anyway.
 * Batches created earlier keep the filter radius they were created with. */
int dvg_scene_update(dvg_scene *scene, const dvg_scene_params *params, float filter_radius, void *stream);
/* Waits for the scene's last stream and reports the conditions its last update found and
 * those its renders found since the previous dvg_scene_status or dvg_scene_get_grads call
 * (the reference asserts or throws on them; the render conditions are cleared by the
 * call): 0 when there are none. */
int dvg_scene_status(dvg_scene *scene);
/* Copies the scene's device block (its derived tables; layout internal to this library
 * version) to host memory `dst` of `cap` bytes; `bytes` receives the block size.  The
 * PLAIN schedule's top-level BVH and the rebuild's sort scratch read as zeros, so the
 * blocks of a rebuilt scene and of a scene created from the same values are equal.  For
 * tests comparing a device rebuild with a host build; no reference counterpart. */
int dvg_scene_debug_block(dvg_scene *scene, void *dst, int64_t cap, int64_t *bytes);

/* Batched render: many (scene, render args) jobs -- scenes of any size, resolution,
 * sample count, output (colour or SDF) and prefiltering -- rendered with one kernel
 * per stage over the work items of all jobs.  It replaces the reference's loop of one
 * render(...) call per scene (diffvg.cpp:1477-1649 called per item, as
 * apps/generative_models/rendering.py:262-290 does); each job has the semantics of
 * dvg_render(scene, &args[i]) in DVG_MODE_TILED (at most 256 samples per pixel).
 *
 * dvg_batch_create validates the jobs, uploads the job table and reserves the
 * workspace; the scenes and every buffer named in `args` must stay alive and in place
 * until the batch is destroyed.  dvg_batch_render enqueues the kernels on `stream`: it
 * allocates nothing and waits for nothing, so it may be captured into a HIP graph and
 * replayed; with DVG_BATCH_ZERO_GRADS it first clears the gradients of the scenes with
 * backward jobs.  dvg_batch_status waits for the last render and reports the
 * conditions the reference asserts on (fragment overflow, prefiltered circle strokes).
 * Jobs of one batch that share a scene, or renders of the batch's scenes outside it,
 * must not overlap in time. */
#define DVG_BATCH_ZERO_GRADS 1
/* DVG_BATCH_PERSISTENT (one kernel launch per render whose workgroups take the work
 * items from a device queue) is disabled in this build: dvg_batch_create fails with it. */
#define DVG_BATCH_PERSISTENT 2
typedef struct dvg_batch dvg_batch;
int dvg_batch_create(dvg_scene *const *scenes, const dvg_render_args *args, int n, int flags, dvg_batch **out);
int dvg_batch_render(dvg_batch *batch, void *stream);
int dvg_batch_status(dvg_batch *batch);
int dvg_batch_destroy(dvg_batch *batch);

/* Kernel timing: when enabled, each kernel launch is bracketed by HIP events recorded
 * on its stream.  dvg_profile_read waits for the events and returns the number of
 * kernels, their names (newline separated) and summed milliseconds and launch counts,
 * then clears the records.  No reference counterpart (the reference only has
 * wall-clock print_timing, render_pytorch.py:8-12). */
int dvg_profile_enable(int on);
int dvg_profile_read(char *names, int cap, double *ms, int *counts, int max_kernels);

/* Last error message of the calling thread ("" when none). */
const char *dvg_last_error(void);
int dvg_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif
