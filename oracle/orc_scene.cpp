// orc_scene.cpp -- oracle scene construction (test infrastructure only).
//
// Restates Scene::Scene (scene.cpp:919-998): copy parameters, compute boundary lengths
// (compute_shape_length, scene.cpp:113-205), the global shape CDF (build_shape_cdfs,
// scene.cpp:207-246), the per-path segment CDFs (build_path_cdfs, scene.cpp:248-333)
// and the boxes plus the three BVH levels (compute_bounding_boxes, scene.cpp:496-684,
// build_bvh, scene.cpp:439-494).  std::sort is called with the reference's comparators
// so that the (unstable) sort produces the same permutation on the same input.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "orc_scene.h"

namespace orc {

static V2 quad_eval(V2 p0, V2 p1, V2 p2, float t) {
    float tt = 1 - t;
    return (tt * tt) * p0 + (2 * tt * t) * p1 + (t * t) * p2;
}
static V2 cubic_eval(V2 p0, V2 p1, V2 p2, V2 p3, float t) {
    float tt = 1 - t;
    return (tt * tt * tt) * p0 + (3 * tt * tt * t) * p1 + (3 * tt * t * t) * p2 + (t * t * t) * p3;
}

// build_bvh<sort> (scene.cpp:439-494)
void build_bvh(const Scene &S, Node *nodes, int n, bool sort) {
    int bvh_size = 2 * n - 1;
    if (bvh_size <= 1) return;
    if (sort) {
        std::sort(nodes, nodes + n, [&](const Node &n0, const Node &n1) {
            V2 c0 = 0.5f * (V2{n0.box.minx, n0.box.miny} + V2{n0.box.maxx, n0.box.maxy});
            V2 c1 = 0.5f * (V2{n1.box.minx, n1.box.miny} + V2{n1.box.maxx, n1.box.maxy});
            return morton2d(c0, S.cw, S.ch) < morton2d(c1, S.cw, S.ch);
        });
    }
    for (int i = n; i < bvh_size; i++) nodes[i] = Node{-1, -1, box_empty(), 0.f};
    int prev_beg = 0, prev_end = n;
    int leftover = prev_end % 2 == 0 ? -1 : prev_end - 1;
    while (prev_end - prev_beg >= 1 || leftover != -1) {
        int length = (prev_end - prev_beg) / 2;
        if ((prev_end - prev_beg) % 2 == 1 && leftover != -1 && leftover != prev_end - 1) length += 1;
        for (int i = 0; i < length; i++) {
            Node node;
            node.child0 = prev_beg + 2 * i;
            node.child1 = prev_beg + 2 * i + 1;
            if (node.child1 >= prev_end) {
                node.child1 = leftover;
                leftover = -1;
            }
            node.box = merge(nodes[node.child0].box, nodes[node.child1].box);
            node.max_radius = std::max(nodes[node.child0].max_radius, nodes[node.child1].max_radius);
            nodes[prev_end + i] = node;
        }
        if (length == 1 && leftover == -1) break;
        prev_beg = prev_end;
        prev_end = prev_beg + length;
        if (length % 2 == 1 && leftover == -1) leftover = prev_end - 1;
    }
}

// Segment endpoints of path s, segment i starting at point `pid` (scene.cpp:147-184).
static void seg_points(const Scene &S, int s, int i, int pid, V2 *p, int *idx) {
    int n = S.npts[s];
    int k = S.ncp[S.ctrl_off[s] + i];
    idx[0] = pid;
    if (k == 0) {
        idx[1] = (pid + 1) % n;
    } else if (k == 1) {
        idx[1] = pid + 1;
        idx[2] = (pid + 2) % n;
    } else {
        idx[1] = pid + 1;
        idx[2] = pid + 2;
        idx[3] = (pid + 3) % n;
    }
    for (int j = 0; j <= k + 1; j++) p[j] = S.pt(s, idx[j]);
}

bool build_scene(const orc_scene_desc &d, Scene &S, std::string &err) {
    char buf[256];
    S.cw = d.canvas_width;
    S.ch = d.canvas_height;
    S.ns = d.num_shapes;
    S.ng = d.num_groups;
    if (S.ns < 0 || S.ng < 0) {
        err = "negative shape or group count";
        return false;
    }
    S.type.assign(d.shape_type, d.shape_type + S.ns);
    S.stroke_width.assign(d.stroke_width, d.stroke_width + S.ns);
    S.param_off.assign(d.shape_param_offset, d.shape_param_offset + S.ns);
    if (d.num_shape_params > 0) S.params.assign(d.shape_params, d.shape_params + d.num_shape_params);
    S.ctrl_off.assign(d.path_ctrl_offset, d.path_ctrl_offset + S.ns);
    S.nbase.assign(d.path_num_base_points, d.path_num_base_points + S.ns);
    S.pt_off.assign(d.path_point_offset, d.path_point_offset + S.ns);
    S.npts.assign(d.path_num_points, d.path_num_points + S.ns);
    S.has_thick.assign(d.path_has_thickness, d.path_has_thickness + S.ns);
    S.closed.assign(d.path_is_closed, d.path_is_closed + S.ns);
    S.approx.assign(d.path_use_distance_approx, d.path_use_distance_approx + S.ns);
    if (d.num_ctrl_total > 0) S.ncp.assign(d.num_control_points, d.num_control_points + d.num_ctrl_total);
    if (d.num_points_total > 0) {
        S.points.assign(d.points, d.points + 2 * d.num_points_total);
        if (d.thickness)
            S.thickness.assign(d.thickness, d.thickness + d.num_points_total);
        else
            S.thickness.assign(d.num_points_total, 0.f);
    }
    S.g_off.assign(d.group_shape_offset, d.group_shape_offset + S.ng);
    S.g_n.assign(d.group_num_shapes, d.group_num_shapes + S.ng);
    if (d.num_group_shape_ids > 0)
        S.g_ids.assign(d.group_shape_ids, d.group_shape_ids + d.num_group_shape_ids);
    S.fill_type.assign(d.fill_type, d.fill_type + S.ng);
    S.fill_off.assign(d.fill_offset, d.fill_offset + S.ng);
    S.fill_ns.assign(d.fill_num_stops, d.fill_num_stops + S.ng);
    S.stroke_type.assign(d.stroke_type, d.stroke_type + S.ng);
    S.stroke_off.assign(d.stroke_offset, d.stroke_offset + S.ng);
    S.stroke_ns.assign(d.stroke_num_stops, d.stroke_num_stops + S.ng);
    if (d.num_color_params > 0) S.color.assign(d.color_params, d.color_params + d.num_color_params);
    S.even_odd.assign(d.use_even_odd_rule, d.use_even_odd_rule + S.ng);
    S.shape_to_canvas.resize(S.ng);
    S.canvas_to_shape.resize(S.ng);
    for (int g = 0; g < S.ng; g++) {
        S.shape_to_canvas[g] = m3_from(d.shape_to_canvas + 9 * g);
        S.canvas_to_shape[g] = m3_inverse(S.shape_to_canvas[g]);  // shape.h:123
    }
    S.filter_type = d.filter_type;
    S.filter_radius = d.filter_radius;

    // gradient buffers zero-initialised (scene.cpp:45-111, 335-429, 993)
    S.d_params.assign(S.params.size(), 0.f);
    S.d_points.assign(S.points.size(), 0.f);
    S.d_thickness.assign(S.thickness.size(), 0.f);
    S.d_stroke_width.assign(S.ns, 0.f);
    S.d_color.assign(S.color.size(), 0.f);
    S.d_xform.assign(9 * (size_t)S.ng, 0.f);
    S.d_filter_radius = 0.f;

    if (S.ns == 0) return true;  // scene.cpp:931-933: nothing else is built
    for (int g = 0; g < S.ng; g++) {
        if (S.g_n[g] <= 0) {
            err = "shape group without shapes";
            return false;
        }
    }

    // compute_shape_length (scene.cpp:113-205)
    S.shape_length.assign(S.ns, 0.f);
    for (int s = 0; s < S.ns; s++) {
        float shape_length = 0.f;
        const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
        switch (S.type[s]) {
            case kCircle:
                shape_length += (float)(2.f * kPi) * P[0];
                break;
            case kEllipse: {
                float a = P[0], b = P[1];
                shape_length = (float)((double)shape_length +
                                       (double)(float)kPi * ((double)(3 * (a + b)) -
                                                             std::sqrt((double)((3 * a + b) * (a + 3 * b)))));
                break;
            }
            case kPath: {
                float length = 0.f;
                int point_id = 0;
                for (int i = 0; i < S.nbase[s]; i++) {
                    V2 p[4];
                    int idx[4];
                    seg_points(S, s, i, point_id, p, idx);
                    int k = S.ncp[S.ctrl_off[s] + i];
                    point_id += k + 1;
                    if (k == 0) {
                        length += dist(p[1], p[0]);
                    } else if (k == 1) {
                        V2 v1 = quad_eval(p[0], p[1], p[2], 0.5f);
                        length += dist(v1, p[0]) + dist(v1, p[2]);
                    } else {
                        V2 v1 = cubic_eval(p[0], p[1], p[2], p[3], 1.f / 3.f);
                        V2 v2 = cubic_eval(p[0], p[1], p[2], p[3], 2.f / 3.f);
                        length += dist(v1, p[0]) + dist(v1, v2) + dist(v2, p[3]);
                    }
                }
                shape_length += length;
                break;
            }
            case kRect:
                shape_length += 2 * (P[2] - P[0] + P[3] - P[1]);
                break;
            default:
                err = "unknown shape type";
                return false;
        }
        S.shape_length[s] = shape_length;
    }

    // build_shape_cdfs (scene.cpp:207-246)
    int total = 0;
    for (int g = 0; g < S.ng; g++) total += S.g_n[g];
    S.num_total_shapes = total;
    S.shapes_cdf.assign(total, 0.f);
    S.shapes_pmf.assign(total, 0.f);
    S.sample_shape_id.assign(total, 0);
    S.sample_group_id.assign(total, 0);
    int sid = 0;
    for (int g = 0; g < S.ng; g++) {
        for (int i = 0; i < S.g_n[g]; i++) {
            int shape_id = S.g_ids[S.g_off[g] + i];
            float length = S.shape_length[shape_id];
            S.sample_shape_id[sid] = shape_id;
            S.shapes_cdf[sid] = sid == 0 ? length : length + S.shapes_cdf[sid - 1];
            S.shapes_pmf[sid] = length;
            S.sample_group_id[sid] = g;
            sid++;
        }
    }
    if (total == 0) {
        err = "scene has shapes but no shape group";
        return false;
    }
    float normalization = S.shapes_cdf[total - 1];
    if (normalization <= 0) {
        snprintf(buf, sizeof(buf),
                 "The total length of the shape boundaries in the scene is equal or less than 0. Length = %f",
                 normalization);
        err = buf;
        return false;
    }
    if (!std::isfinite(normalization)) {
        snprintf(buf, sizeof(buf),
                 "The total length of the shape boundaries in the scene is not a number. Length = %f",
                 normalization);
        err = buf;
        return false;
    }
    for (int i = 0; i < total; i++) {
        S.shapes_cdf[i] /= normalization;
        S.shapes_pmf[i] /= normalization;
    }

    // build_path_cdfs (scene.cpp:248-333)
    S.path_cdf.assign(S.ncp.size(), 0.f);
    S.path_pmf.assign(S.ncp.size(), 0.f);
    S.path_point_id_map.assign(S.ncp.size(), 0);
    for (int s = 0; s < S.ns; s++) {
        if (S.type[s] != kPath) continue;
        float *pmf = &S.path_pmf[S.ctrl_off[s]];
        float *cdf = &S.path_cdf[S.ctrl_off[s]];
        int *pmap = &S.path_point_id_map[S.ctrl_off[s]];
        float inv_length = 1.f / S.shape_length[s];
        int point_id = 0;
        for (int i = 0; i < S.nbase[s]; i++) {
            pmap[i] = point_id;
            V2 p[4];
            int idx[4];
            seg_points(S, s, i, point_id, p, idx);
            int k = S.ncp[S.ctrl_off[s] + i];
            point_id += k + 1;
            float dd;
            if (k == 0) {
                dd = dist(p[0], p[1]) * inv_length;
            } else if (k == 1) {
                V2 v1 = quad_eval(p[0], p[1], p[2], 0.5f);
                dd = (dist(p[0], v1) + dist(v1, p[2])) * inv_length;
            } else {
                V2 v1 = cubic_eval(p[0], p[1], p[2], p[3], 1.f / 3.f);
                V2 v2 = cubic_eval(p[0], p[1], p[2], p[3], 2.f / 3.f);
                dd = (dist(v1, p[0]) + dist(v1, v2) + dist(v2, p[3])) * inv_length;
            }
            pmf[i] = dd;
            cdf[i] = i == 0 ? dd : dd + cdf[i - 1];
        }
    }

    // compute_bounding_boxes (scene.cpp:496-684)
    S.shape_box.assign(S.ns, box_empty());
    S.path_bvh_off.assign(S.ns, -1);
    size_t path_nodes = 0;
    for (int s = 0; s < S.ns; s++)
        if (S.type[s] == kPath) path_nodes += 2 * (size_t)S.nbase[s] - 1;
    S.path_bvh.assign(path_nodes, Node{-1, -1, box_empty(), 0.f});
    size_t cursor = 0;
    for (int s = 0; s < S.ns; s++) {
        const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
        switch (S.type[s]) {
            case kCircle:
                S.shape_box[s] = Box{P[1] - P[0], P[2] - P[0], P[1] + P[0], P[2] + P[0]};
                break;
            case kEllipse:
                S.shape_box[s] = Box{P[2] - P[0], P[3] - P[1], P[2] + P[0], P[3] + P[1]};
                break;
            case kRect:
                S.shape_box[s] = Box{P[0], P[1], P[2], P[3]};
                break;
            case kPath: {
                Box box = box_empty();
                if (S.npts[s] > 0) {
                    V2 p0 = S.pt(s, 0);
                    box = Box{p0.x, p0.y, p0.x, p0.y};
                }
                for (int i = 1; i < S.npts[s]; i++) box = merge(box, S.pt(s, i));
                S.shape_box[s] = box;
                int nb = S.nbase[s];
                std::vector<Box> boxes(nb);
                std::vector<float> thick(nb);
                std::vector<int> first_point_id(nb);
                float r = S.stroke_width[s];
                int point_id = 0;
                for (int i = 0; i < nb; i++) {
                    first_point_id[i] = point_id;
                    V2 p[4];
                    int idx[4];
                    seg_points(S, s, i, point_id, p, idx);
                    int k = S.ncp[S.ctrl_off[s] + i];
                    point_id += k + 1;
                    boxes[i] = box_empty();
                    float rr = -1;
                    for (int j = 0; j <= k + 1; j++) {
                        boxes[i] = merge(boxes[i], p[j]);
                        float rj = S.has_thick[s] ? S.thick(s, idx[j]) : r;
                        rr = j == 0 ? rj : fmax_r(rr, rj);
                    }
                    thick[i] = rr;
                }
                std::vector<int> order(nb);
                std::iota(order.begin(), order.end(), 0);
                std::sort(order.begin(), order.end(), [&](int i0, int i1) {
                    float b0y = 0.5f * (boxes[i0].miny + boxes[i0].maxy);
                    float b1y = 0.5f * (boxes[i1].miny + boxes[i1].maxy);
                    return b0y < b1y;
                });
                Node *nodes = &S.path_bvh[cursor];
                S.path_bvh_off[s] = (int)cursor;
                for (int i = 0; i < nb; i++)
                    nodes[i] = Node{order[i], -(first_point_id[order[i]] + 1), boxes[order[i]],
                                    thick[order[i]]};
                build_bvh(S, nodes, nb, false);
                cursor += 2 * (size_t)nb - 1;
                break;
            }
            default:
                err = "unknown shape type";
                return false;
        }
    }
    // group BVHs (scene.cpp:628-646)
    S.group_bvh_off.assign(S.ng, 0);
    size_t gnodes = 0;
    for (int g = 0; g < S.ng; g++) gnodes += 2 * (size_t)S.g_n[g] - 1;
    S.group_bvh.assign(gnodes, Node{-1, -1, box_empty(), 0.f});
    cursor = 0;
    for (int g = 0; g < S.ng; g++) {
        Node *nodes = &S.group_bvh[cursor];
        S.group_bvh_off[g] = (int)cursor;
        for (int i = 0; i < S.g_n[g]; i++) {
            int shape_id = S.g_ids[S.g_off[g] + i];
            float r = S.stroke_type[g] == kNoColor ? 0.f : S.stroke_width[shape_id];
            nodes[i] = Node{shape_id, -1, S.shape_box[shape_id], r};
        }
        build_bvh(S, nodes, S.g_n[g], true);
        cursor += 2 * (size_t)S.g_n[g] - 1;
    }
    // top level (scene.cpp:648-683); thickness paths read path_bvhs[id][0] (a leaf)
    S.top_bvh.assign(S.ng > 0 ? 2 * (size_t)S.ng - 1 : 0, Node{-1, -1, box_empty(), 0.f});
    for (int g = 0; g < S.ng; g++) {
        auto shape_radius = [&](int sid_) {
            float r = S.stroke_width[sid_];
            if (S.type[sid_] == kPath && S.has_thick[sid_]) r = S.pbvh(sid_)[0].max_radius;
            return r;
        };
        float max_radius = shape_radius(S.g_ids[S.g_off[g]]);
        for (int i = 1; i < S.g_n[g]; i++) max_radius = std::max(max_radius, shape_radius(S.g_ids[S.g_off[g] + i]));
        Box bbox = S.gbvh(g)[2 * S.g_n[g] - 2].box;
        S.top_bvh[g].child0 = g;
        S.top_bvh[g].child1 = -1;
        S.top_bvh[g].box = transform_box(S.shape_to_canvas[g], bbox);
        S.top_bvh[g].max_radius = S.stroke_type[g] == kNoColor ? 0.f : max_radius;
    }
    if (S.ng > 0) build_bvh(S, S.top_bvh.data(), S.ng, true);
    return true;
}

}  // namespace orc
