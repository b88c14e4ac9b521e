// orc_scene.h -- oracle scene container (test infrastructure only).
//
// Holds the flat copy of the scene descriptor plus everything scene.cpp derives from
// it: canvas_to_shape = inverse(shape_to_canvas) (shape.h:123), boundary lengths and
// sampling CDFs (scene.cpp:113-333), shape boxes and the three BVH levels
// (scene.cpp:431-684), and the gradient buffers (scene.cpp:45-111, 335-429).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "orc_math.h"
#include "oracle.h"

namespace orc {

enum ShapeType { kCircle = 0, kEllipse = 1, kPath = 2, kRect = 3 };
enum ColorType { kNoColor = -1, kConstant = 0, kLinear = 1, kRadial = 2 };

struct Node {  // BVHNode (scene.h:12-16)
    int child0, child1;
    Box box;
    float max_radius;
};

struct Scene {
    int cw = 0, ch = 0;
    // shapes
    int ns = 0;
    std::vector<int> type;
    std::vector<float> stroke_width;
    std::vector<int> param_off;
    std::vector<float> params;
    std::vector<int> ctrl_off, nbase, pt_off, npts;
    std::vector<uint8_t> has_thick, closed, approx;
    std::vector<int> ncp;
    std::vector<float> points, thickness;
    // groups
    int ng = 0;
    std::vector<int> g_off, g_n, g_ids;
    std::vector<int> fill_type, fill_off, fill_ns, stroke_type, stroke_off, stroke_ns;
    std::vector<float> color;
    std::vector<uint8_t> even_odd;
    std::vector<M3> shape_to_canvas, canvas_to_shape;
    int filter_type = 0;
    float filter_radius = 0.5f;
    // derived (scene.cpp)
    int num_total_shapes = 0;
    std::vector<float> shape_length;
    std::vector<float> shapes_cdf, shapes_pmf;
    std::vector<int> sample_shape_id, sample_group_id;
    std::vector<float> path_cdf, path_pmf;  // indexed by ctrl_off
    std::vector<int> path_point_id_map;     // indexed by ctrl_off
    std::vector<Box> shape_box;
    std::vector<int> path_bvh_off;          // per shape, -1 for non-path
    std::vector<Node> path_bvh;
    std::vector<int> group_bvh_off;
    std::vector<Node> group_bvh;
    std::vector<Node> top_bvh;
    // gradients
    std::vector<float> d_params, d_points, d_thickness, d_stroke_width, d_color, d_xform;
    float d_filter_radius = 0.f;

    V2 pt(int s, int i) const {
        int k = pt_off[s] + i;
        return {points[2 * k], points[2 * k + 1]};
    }
    float thick(int s, int i) const { return thickness[pt_off[s] + i]; }
    const Node *pbvh(int s) const { return &path_bvh[path_bvh_off[s]]; }
    const Node *gbvh(int g) const { return &group_bvh[group_bvh_off[g]]; }
};

// scene.cpp
bool build_scene(const orc_scene_desc &d, Scene &S, std::string &err);
void build_bvh(const Scene &S, Node *nodes, int n, bool sort);

}  // namespace orc
