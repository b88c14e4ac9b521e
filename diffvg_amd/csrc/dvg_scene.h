// dvg_scene.h -- scene layout shared by the host builder and the HIP kernels.
//
// The scene lives in HBM as structure-of-arrays, one device allocation per scene:
//   shapes    : type, stroke width, parameter offset, path ranges, local boxes
//   segments  : one record per path segment (point indices, kind, box, radius,
//               length CDF/PMF) plus, per path, the segment order in which the
//               reference's segment BVH visits its leaves (depth first, child1 first)
//   groups    : paint descriptors, even-odd flag, shape_to_canvas and its inverse,
//               the local root box, the canvas-space top-level leaf box and radius,
//               and the member shapes in group-BVH visit order
//   top BVH   : the reference's top-level BVH over groups (used by the PLAIN schedule)
//   sampling  : boundary lengths and the global shape CDF used by edge sampling
//   gradients : one zero-initialised float block with the parameter layout of the
//               scene description, so gradient read-back is a single copy.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DVG_HD __host__ __device__
#else
#define DVG_HD
#endif

namespace dvg {

struct BNode {  // BVHNode (reference scene.h:12-16), padded to 32 bytes
    int child0, child1;
    float minx, miny, maxx, maxy;
    float max_radius;
    int pad;
};

// float4 per segment of DevScene::seg_wf: wf::CubicPrep (3), wf::stroke_prep (5)
constexpr int kSegPrep = 8;

struct DevScene {
    int cw, ch;
    int ns, ng, nseg, npts, num_total_shapes;
    int filter_type;
    float filter_radius;
    // shapes [ns]
    const int *shape_type;
    const float *stroke_width;
    const int *param_off;
    const float *params;
    const int *seg_begin;      // first segment of a path (== its ctrl offset)
    const int *nsegs;          // number of segments of a path
    const int *pt_off;
    const int *npt;
    const uint8_t *has_thick;
    const uint8_t *closed;
    const uint8_t *approx;
    const float4 *shape_box;   // local-space box (scene.cpp:500-570)
    // segments [nseg]
    const int4 *seg_idx;       // absolute point indices, -1 when unused
    const int *seg_kind;       // number of control points: 0 line, 1 quadratic, 2 cubic
    const int *seg_pid;        // first point id relative to the path
    const float4 *seg_box;
    const float *seg_maxr;
    const float4 *seg_wf;      // [kSegPrep * nseg] fast-test prep of each cubic segment (from its points)
    const int *seg_dfs;        // per path (indexed from seg_begin): segment ids in visit order
    const float *seg_cdf;
    const float *seg_pmf;
    const float2 *points;
    const float *thickness;
    // groups [ng]
    const int *g_off;
    const int *g_n;
    const int *g_ids;          // member shapes in group-BVH visit order
    const int *fill_type, *fill_off, *fill_ns;
    const int *stroke_type, *stroke_off, *stroke_ns;
    const float *color;
    const uint8_t *even_odd;
    const float *s2c;          // shape_to_canvas, 9 per group
    const float *c2s;          // canvas_to_shape = inverse(shape_to_canvas)
    const float4 *g_lbox;      // local root box of the group BVH
    const float4 *g_box;       // canvas-space top-level leaf box
    const float *g_maxr;       // top-level leaf max_radius
    // SDF culling per group: (x, y) = a canvas point of the boundary whose distance
    // bounds the group's computed distance from above (NaN when no such bound holds:
    // non-similarity transform or approximate quadratic distances), z = 1 when the group
    // can report a distance at all (it has a path with segments or a rect)
    const float4 *g_sdf;
    // top-level BVH [2 ng - 1]
    const BNode *top;
    // edge sampling
    const float *shape_length;
    const float *shapes_cdf;
    const float *shapes_pmf;
    const int *sample_shape_id;
    const int *sample_group_id;
    // gradients (parameter layout of dvg_scene_desc)
    float *d_params;
    float *d_points;
    float *d_thickness;
    float *d_stroke_width;
    float *d_color;
    float *d_xform;
    float *d_filter_radius;
};

}  // namespace dvg
