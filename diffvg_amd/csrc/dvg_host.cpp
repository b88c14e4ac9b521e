// dvg_host.cpp -- scene construction and the C ABI.
//
// dvg_scene_create derives, from the flat description, everything Scene::Scene
// (reference scene.cpp:919-998) derives: canvas_to_shape (shape.h:123), boundary
// lengths and sampling CDFs (scene.cpp:113-333), boxes and the three BVH levels
// (scene.cpp:431-684).  std::sort is called with the reference's comparators so the
// BVH topology, and with it the leaf visit orders the kernels follow, is the
// reference's.  The derived arrays are packed into one device allocation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "dvg_build.h"
#include "dvg_internal.h"

namespace dvg {

static thread_local std::string t_error;

// ---- device memory cache (size classes of powers of two, per device)
namespace {
std::mutex g_cache_mu;
std::map<std::pair<int, size_t>, std::vector<void *>> g_cache;
size_t size_class(size_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}
}  // namespace

void *cache_alloc(int device, size_t bytes, size_t *granted) {
    size_t c = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto &v = g_cache[{device, c}];
        if (!v.empty()) {
            void *p = v.back();
            v.pop_back();
            *granted = c;
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) return nullptr;
    *granted = c;
    return p;
}

void cache_free(int device, void *p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache[{device, size_class(bytes)}].push_back(p);
}

// A non-blocking stream per device for scene and batch uploads: a copy on it waits only
// for itself, never for kernels that other streams have in flight (the null stream
// would order it after all of them).
hipStream_t upload_stream(int device) {
    static std::mutex m;
    static std::map<int, hipStream_t> streams;
    std::lock_guard<std::mutex> lk(m);
    hipStream_t &s = streams[device];
    if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    return s;
}

int upload(int device, void *dst, const void *src, size_t bytes) {
    hipStream_t s = upload_stream(device);
    if (!s) return set_error("cannot create the upload stream");
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_error(std::string("upload: ") + hipGetErrorString(e));
    return 0;
}

int clear_device(int device, void *dst, size_t bytes) {
    hipStream_t s = upload_stream(device);
    if (!s) return set_error("cannot create the upload stream");
    hipError_t e = hipMemsetAsync(dst, 0, bytes, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_error(std::string("clear: ") + hipGetErrorString(e));
    return 0;
}

int set_error(const std::string &msg) {
    t_error = msg;
    return 1;
}

namespace {

using HV2 = bld::P2;
using HBox = bld::Box;
using bld::box_empty;
using bld::merge;
using bld::mx;

struct HNode {
    int child0, child1;
    HBox box;
    float max_radius;
};

// pairwise bottom-up build with the odd-node carry (scene.cpp:439-494)
void build_bvh(HNode *nodes, int n, bool sort, int cw, int ch) {
    int size = 2 * n - 1;
    if (size <= 1) return;
    if (sort) {
        std::sort(nodes, nodes + n, [&](const HNode &a, const HNode &b) {
            return bld::morton_of_box(a.box, cw, ch) < bld::morton_of_box(b.box, cw, ch);
        });
    }
    for (int i = n; i < size; i++) nodes[i] = HNode{-1, -1, box_empty(), 0.f};
    int beg = 0, end = n;
    int leftover = end % 2 == 0 ? -1 : end - 1;
    while (end - beg >= 1 || leftover != -1) {
        int length = (end - beg) / 2;
        if ((end - beg) % 2 == 1 && leftover != -1 && leftover != end - 1) length += 1;
        for (int i = 0; i < length; i++) {
            HNode node;
            node.child0 = beg + 2 * i;
            node.child1 = beg + 2 * i + 1;
            if (node.child1 >= end) {
                node.child1 = leftover;
                leftover = -1;
            }
            node.box = merge(nodes[node.child0].box, nodes[node.child1].box);
            node.max_radius = std::max(nodes[node.child0].max_radius, nodes[node.child1].max_radius);
            nodes[end + i] = node;
        }
        if (length == 1 && leftover == -1) break;
        beg = end;
        end = beg + length;
        if (length % 2 == 1 && leftover == -1) leftover = end - 1;
    }
}

// Leaves in the order the reference's stack-based traversal pops them (child1 first).
void dfs_leaves(const HNode *nodes, int n, std::vector<int> &out) {
    if (n <= 0) return;
    std::vector<int> stack;
    stack.push_back(2 * n - 2);
    while (!stack.empty()) {
        int k = stack.back();
        stack.pop_back();
        const HNode &nd = nodes[k];
        if (nd.child1 < 0) {
            out.push_back(nd.child0);
        } else {
            stack.push_back(nd.child0);
            stack.push_back(nd.child1);
        }
    }
}

// Leaf positions (0..n-1 in sorted order) in the order the traversal visits them: the
// pairwise build's shape depends on n only, so this is topology the device rebuild reuses.
const std::vector<int> &visit_positions(int n) {
    static std::mutex mu;
    static std::map<int, std::vector<int>> memo;
    std::lock_guard<std::mutex> lock(mu);
    auto it = memo.find(n);
    if (it != memo.end()) return it->second;
    std::vector<HNode> nodes(n > 0 ? 2 * n - 1 : 0, HNode{-1, -1, box_empty(), 0.f});
    for (int i = 0; i < n; i++) nodes[i] = HNode{i, -1, box_empty(), 0.f};
    build_bvh(nodes.data(), n, false, 1, 1);
    std::vector<int> out;
    dfs_leaves(nodes.data(), n, out);
    return memo.emplace(n, std::move(out)).first->second;
}

// Packs host arrays into one device allocation with 256-byte aligned sub-arrays.
struct Packer {
    std::vector<uint8_t> host;
    template <typename T>
    size_t add(const std::vector<T> &v) {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + std::max<size_t>(v.size() * sizeof(T), 16));
        if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

}  // namespace

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return set_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int build_scene(const dvg_scene_desc &d, int gpu_index, SceneHost &sc) {
    char buf[256];
    const int ns = d.num_shapes, ng = d.num_groups;
    if (ns < 0 || ng < 0 || d.canvas_width <= 0 || d.canvas_height <= 0)
        return set_error("invalid canvas size or shape/group count");
    if (ns > 0 && (!d.shape_type || !d.stroke_width || !d.shape_param_offset || !d.path_ctrl_offset ||
                   !d.path_num_base_points || !d.path_point_offset || !d.path_num_points || !d.path_has_thickness ||
                   !d.path_is_closed || !d.path_use_distance_approx))
        return set_error("null shape array in the scene description");
    if (ng > 0 && (!d.group_shape_offset || !d.group_num_shapes || !d.group_shape_ids || !d.fill_type ||
                   !d.fill_offset || !d.fill_num_stops || !d.stroke_type || !d.stroke_offset || !d.stroke_num_stops ||
                   !d.use_even_odd_rule || !d.shape_to_canvas))
        return set_error("null group array in the scene description");
    if (d.filter_type < 0 || d.filter_type > 4) return set_error("unknown filter type");
    if (!(d.filter_radius > 0) || !std::isfinite(d.filter_radius)) return set_error("filter radius must be positive");
    sc.cw = d.canvas_width;
    sc.ch = d.canvas_height;
    sc.num_shapes = ns;
    sc.num_groups = ng;
    sc.num_shape_params = d.num_shape_params;
    sc.num_points_total = d.num_points_total;
    sc.num_color_params = d.num_color_params;

    // ---- shapes and segments
    std::vector<int> shape_type(ns), param_off(ns), seg_begin(ns, 0), nsegs(ns, 0), pt_off(ns, 0), npt(ns, 0);
    std::vector<float> stroke_width(ns), params(d.shape_params ? d.shape_params : (const float *)nullptr,
                                                d.shape_params ? d.shape_params + d.num_shape_params : nullptr);
    std::vector<uint8_t> has_thick(ns, 0), closed(ns, 0), approx(ns, 0);
    std::vector<float4> shape_box(ns);
    std::vector<HBox> sbox(ns, box_empty());
    const int nseg = d.num_ctrl_total;
    std::vector<int4> seg_idx(nseg);
    std::vector<int> seg_kind(nseg, 0), seg_pid(nseg, 0), seg_dfs(nseg, 0), seg_perm(nseg, 0);
    std::vector<float4> seg_box(nseg);
    std::vector<float> seg_maxr(nseg, 0.f), seg_cdf(nseg, 0.f), seg_pmf(nseg, 0.f);
    std::vector<float2> points(d.num_points_total);
    std::vector<float> thickness(d.num_points_total, 0.f);
    for (int i = 0; i < d.num_points_total; i++) points[i] = make_float2(d.points[2 * i], d.points[2 * i + 1]);
    if (d.thickness)
        for (int i = 0; i < d.num_points_total; i++) thickness[i] = d.thickness[i];
    std::vector<float> shape_length(ns, 0.f);
    std::vector<float> thick_leaf0(ns, 0.f);  // max_radius of leaf 0 of a thickness path's BVH
    std::vector<HNode> pnodes;
    std::vector<int> long_ids;  // paths the device rebuild gives a workgroup each
    std::vector<int> big_gids;  // groups whose members the device rebuild sorts with a workgroup
    for (int s = 0; s < ns; s++) {
        shape_type[s] = d.shape_type[s];
        stroke_width[s] = d.stroke_width[s];
        param_off[s] = d.shape_param_offset[s];
        int t = shape_type[s];
        if (t < 0 || t > 3) return set_error("unknown shape type");
        if (t != 2) {
            int need = t == 0 ? 3 : 4;
            if (param_off[s] < 0 || param_off[s] + need > d.num_shape_params)
                return set_error("shape parameter offset out of range");
            shape_length[s] = bld::basic_shape(t, params.data() + param_off[s], sbox[s]);
        } else {
            param_off[s] = 0;
            int c0 = d.path_ctrl_offset[s], nb = d.path_num_base_points[s];
            int p0 = d.path_point_offset[s], np = d.path_num_points[s];
            if (nb <= 0 || np <= 0 || c0 < 0 || c0 + nb > nseg || p0 < 0 || p0 + np > d.num_points_total)
                return set_error("path ranges out of bounds");
            seg_begin[s] = c0;
            nsegs[s] = nb;
            if (nb > kRefitSegCap) long_ids.push_back(s);
            pt_off[s] = p0;
            npt[s] = np;
            has_thick[s] = d.path_has_thickness[s] ? 1 : 0;
            if (has_thick[s] && !d.thickness) return set_error("path with thickness but no thickness array");
            closed[s] = d.path_is_closed[s] ? 1 : 0;
            approx[s] = d.path_use_distance_approx[s] ? 1 : 0;
            HBox box{points[p0].x, points[p0].y, points[p0].x, points[p0].y};
            for (int i = 1; i < np; i++) box = merge(box, HV2{points[p0 + i].x, points[p0 + i].y});
            sbox[s] = box;
            float r = stroke_width[s];
            float length = 0.f;
            int pid = 0;
            std::vector<HBox> boxes(nb);
            std::vector<float> thick(nb);
            std::vector<float> seglen(nb);
            for (int i = 0; i < nb; i++) {
                int k = d.num_control_points[c0 + i];
                if (k < 0 || k > 2) return set_error("num_control_points must be 0, 1 or 2");
                int li[4] = {pid, -1, -1, -1};
                if (k == 0) {
                    li[1] = (pid + 1) % np;
                } else if (k == 1) {
                    li[1] = pid + 1;
                    li[2] = (pid + 2) % np;
                } else {
                    li[1] = pid + 1;
                    li[2] = pid + 2;
                    li[3] = (pid + 3) % np;
                }
                for (int j = 0; j <= k + 1; j++)
                    if (li[j] < 0 || li[j] >= np) return set_error("path has fewer points than its segments need");
                HV2 p[4];
                for (int j = 0; j <= k + 1; j++) p[j] = HV2{points[p0 + li[j]].x, points[p0 + li[j]].y};
                int seg = c0 + i;
                seg_idx[seg] = make_int4(p0 + li[0], p0 + li[1], k >= 1 ? p0 + li[2] : -1, k >= 2 ? p0 + li[3] : -1);
                seg_kind[seg] = k;
                seg_pid[seg] = pid;
                HBox b = box_empty();
                float rr = -1;
                for (int j = 0; j <= k + 1; j++) {
                    b = merge(b, p[j]);
                    float rj = has_thick[s] ? thickness[p0 + li[j]] : r;
                    rr = j == 0 ? rj : mx(rr, rj);
                }
                boxes[i] = b;
                thick[i] = rr;
                seg_box[seg] = make_float4(b.minx, b.miny, b.maxx, b.maxy);
                seg_maxr[seg] = rr;
                float l = bld::segment_length(k, p);
                seglen[i] = l;
                length += l;
                pid += k + 1;
            }
            shape_length[s] = 0.f + length;
            // build_path_cdfs (scene.cpp:248-333)
            float inv_length = 1.f / shape_length[s];
            for (int i = 0; i < nb; i++) {
                float dd = seglen[i] * inv_length;
                seg_pmf[c0 + i] = dd;
                seg_cdf[c0 + i] = i == 0 ? dd : dd + seg_cdf[c0 + i - 1];
            }
            // segment BVH: leaves sorted by box y-centre (scene.cpp:602-619)
            std::vector<int> order(nb);
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(), [&](int i0, int i1) {
                return bld::box_ycentre(boxes[i0]) < bld::box_ycentre(boxes[i1]);
            });
            pnodes.assign(2 * nb - 1, HNode{-1, -1, box_empty(), 0.f});
            for (int i = 0; i < nb; i++) pnodes[i] = HNode{order[i], -1 - i, boxes[order[i]], thick[order[i]]};
            build_bvh(pnodes.data(), nb, false, sc.cw, sc.ch);
            std::vector<int> leaves;
            dfs_leaves(pnodes.data(), nb, leaves);
            const std::vector<int> &vp = visit_positions(nb);
            for (int i = 0; i < nb; i++) {
                seg_dfs[c0 + i] = c0 + leaves[i];
                seg_perm[c0 + i] = vp[i];
            }
            // the top-level radius of a thickness path reads leaf 0 of this BVH (scene.cpp:651-656)
            thick_leaf0[s] = pnodes[0].max_radius;
        }
        if (!std::isfinite(shape_length[s])) return set_error("non-finite shape boundary length");
    }
    for (int s = 0; s < ns; s++) shape_box[s] = make_float4(sbox[s].minx, sbox[s].miny, sbox[s].maxx, sbox[s].maxy);

    // ---- groups
    std::vector<int> g_off(ng), g_n(ng), g_ids(d.num_group_shape_ids), g_ids_desc(d.num_group_shape_ids),
        g_perm(d.num_group_shape_ids);
    std::vector<int> fill_type(ng), fill_off(ng), fill_ns(ng), stroke_type(ng), stroke_off(ng), stroke_ns(ng);
    std::vector<uint8_t> even_odd(ng);
    std::vector<float> s2c(9 * (size_t)ng), c2s(9 * (size_t)ng);
    std::vector<float4> g_lbox(ng), g_box(ng);
    std::vector<float> g_maxr(ng);
    std::vector<float4> g_sdf(ng);
    std::vector<int> g_anchor(ng, -1), g_sdf_flags(ng, 0);
    std::vector<HNode> top(ng > 0 ? 2 * (size_t)ng - 1 : 1, HNode{-1, -1, box_empty(), 0.f});
    int total = 0;
    for (int g = 0; g < ng; g++) {
        g_off[g] = d.group_shape_offset[g];
        g_n[g] = d.group_num_shapes[g];
        if (g_n[g] > kRefitGroupLane) big_gids.push_back(g);
        if (g_n[g] <= 0 || g_off[g] < 0 || g_off[g] + g_n[g] > d.num_group_shape_ids)
            return set_error("shape group member range out of bounds");
        for (int i = 0; i < g_n[g]; i++) {
            int s = d.group_shape_ids[g_off[g] + i];
            if (s < 0 || s >= ns) return set_error("shape group refers to a missing shape");
        }
        total += g_n[g];
        auto check_paint = [&](int t, int off, int nstops) -> bool {
            if (t < -1 || t > 2) return false;
            if (t < 0) return true;
            int need = t == 0 ? 4 : 4 + 5 * nstops;
            if (t > 0 && nstops <= 0) return false;
            return off >= 0 && off + need <= d.num_color_params;
        };
        fill_type[g] = d.fill_type[g];
        fill_off[g] = d.fill_offset[g];
        fill_ns[g] = d.fill_num_stops[g];
        stroke_type[g] = d.stroke_type[g];
        stroke_off[g] = d.stroke_offset[g];
        stroke_ns[g] = d.stroke_num_stops[g];
        if (!check_paint(fill_type[g], fill_off[g], fill_ns[g]) ||
            !check_paint(stroke_type[g], stroke_off[g], stroke_ns[g]))
            return set_error("paint description out of bounds");
        even_odd[g] = d.use_even_odd_rule[g] ? 1 : 0;
        std::memcpy(&s2c[9 * (size_t)g], d.shape_to_canvas + 9 * (size_t)g, 9 * sizeof(float));
        bld::inverse3(&s2c[9 * (size_t)g], &c2s[9 * (size_t)g]);
    }
    if (ns > 0) {
        // build_shape_cdfs (scene.cpp:207-246)
        if (total == 0) return set_error("scene has shapes but no shape group");
        std::vector<float> cdf(total), pmf(total);
        std::vector<int> sid(total), gid(total);
        int k = 0;
        for (int g = 0; g < ng; g++)
            for (int i = 0; i < g_n[g]; i++) {
                int s = d.group_shape_ids[g_off[g] + i];
                float l = shape_length[s];
                sid[k] = s;
                cdf[k] = k == 0 ? l : l + cdf[k - 1];
                pmf[k] = l;
                gid[k] = g;
                k++;
            }
        float norm = cdf[total - 1];
        if (norm <= 0) {
            snprintf(buf, sizeof(buf),
                     "The total length of the shape boundaries in the scene is equal or less than 0. Length = %f", norm);
            return set_error(buf);
        }
        if (!std::isfinite(norm)) {
            snprintf(buf, sizeof(buf),
                     "The total length of the shape boundaries in the scene is not a number. Length = %f", norm);
            return set_error(buf);
        }
        for (int i = 0; i < total; i++) {
            cdf[i] /= norm;
            pmf[i] /= norm;
        }
        // group BVHs (Morton-sorted shape boxes) -> member visit order and root box
        for (int g = 0; g < ng; g++) {
            int n = g_n[g];
            std::vector<HNode> nodes(2 * n - 1, HNode{-1, -1, box_empty(), 0.f});
            for (int i = 0; i < n; i++) {
                int s = d.group_shape_ids[g_off[g] + i];
                float r = stroke_type[g] < 0 ? 0.f : stroke_width[s];
                nodes[i] = HNode{s, -1, sbox[s], r};
            }
            build_bvh(nodes.data(), n, true, sc.cw, sc.ch);
            std::vector<int> leaves;
            dfs_leaves(nodes.data(), n, leaves);
            const std::vector<int> &vp = visit_positions(n);
            for (int i = 0; i < n; i++) {
                g_ids[g_off[g] + i] = leaves[i];
                g_ids_desc[g_off[g] + i] = d.group_shape_ids[g_off[g] + i];
                g_perm[g_off[g] + i] = vp[i];
            }
            HBox rb = nodes[2 * n - 2].box;
            g_lbox[g] = make_float4(rb.minx, rb.miny, rb.maxx, rb.maxy);
            // top-level leaf (scene.cpp:648-681)
            auto radius_of = [&](int s) {
                float r = stroke_width[s];
                if (shape_type[s] == 2 && has_thick[s]) r = thick_leaf0[s];
                return r;
            };
            float maxr = radius_of(d.group_shape_ids[g_off[g]]);
            for (int i = 1; i < n; i++) maxr = std::max(maxr, radius_of(d.group_shape_ids[g_off[g] + i]));
            const float *m = &s2c[9 * (size_t)g];
            HBox tb = bld::canvas_box(m, rb);
            float r = stroke_type[g] < 0 ? 0.f : maxr;
            g_box[g] = make_float4(tb.minx, tb.miny, tb.maxx, tb.maxy);
            g_maxr[g] = r;
            // SDF bound (sample_distance, diffvg.cpp:709-775): a boundary point q of a
            // member whose computed distance is always <= |pt - q|: the first point of a
            // path (its first segment's t = 0 end is always a candidate of the segment's
            // closest-point search, and a pruned segment is farther than the running
            // minimum).  Valid when the transform preserves distance ratios and no member
            // uses the approximate quadratic distance.
            int flags = 0;
            for (int i = 0; i < n; i++) {
                int s = d.group_shape_ids[g_off[g] + i];
                if (shape_type[s] == 2 && nsegs[s] > 0) {
                    flags |= bld::kSdfFindable;
                    if (approx[s])
                        for (int q = 0; q < nsegs[s]; q++)
                            if (seg_kind[seg_begin[s] + q] == 1) flags |= bld::kSdfApproxQuad;
                    if (g_anchor[g] < 0) g_anchor[g] = s;
                } else if (shape_type[s] == 3) {
                    // a rect reports one of its corners, chosen by its nearest side, so
                    // no single point bounds its distance: findable, never an anchor
                    flags |= bld::kSdfFindable;
                }
            }
            g_sdf_flags[g] = flags;
            float qnan = std::numeric_limits<float>::quiet_NaN();
            float fz = (flags & bld::kSdfFindable) ? 1.f : 0.f;
            if ((flags & bld::kSdfFindable) && g_anchor[g] >= 0 && bld::similarity(m) &&
                !(flags & bld::kSdfApproxQuad)) {
                HV2 q = bld::xform(m, HV2{points[pt_off[g_anchor[g]]].x, points[pt_off[g_anchor[g]]].y});
                g_sdf[g] = make_float4(q.x, q.y, fz, 0.f);
            } else {
                g_sdf[g] = make_float4(qnan, qnan, fz, 0.f);
            }
            top[g] = HNode{g, -1, tb, r};
        }
        if (ng > 0) build_bvh(top.data(), ng, true, sc.cw, sc.ch);
        sc.dev.num_total_shapes = total;
        // pack
        Packer P;
        std::vector<BNode> topd(top.size());
        for (size_t i = 0; i < top.size(); i++)
            topd[i] = BNode{top[i].child0, top[i].child1, top[i].box.minx, top[i].box.miny, top[i].box.maxx,
                            top[i].box.maxy, top[i].max_radius, 0};
        size_t o_shape_type = P.add(shape_type), o_sw = P.add(stroke_width), o_poff = P.add(param_off),
               o_params = P.add(params), o_sb = P.add(seg_begin), o_nsegs = P.add(nsegs), o_ptoff = P.add(pt_off),
               o_npt = P.add(npt), o_thick_flag = P.add(has_thick), o_closed = P.add(closed), o_approx = P.add(approx),
               o_sbox = P.add(shape_box), o_sidx = P.add(seg_idx), o_skind = P.add(seg_kind), o_spid = P.add(seg_pid),
               o_sbx = P.add(seg_box), o_smr = P.add(seg_maxr), o_swf = P.add(std::vector<float4>(kSegPrep * (size_t)nseg)), o_sdfs = P.add(seg_dfs), o_scdf = P.add(seg_cdf),
               o_spmf = P.add(seg_pmf), o_pts = P.add(points), o_th = P.add(thickness), o_goff = P.add(g_off),
               o_gn = P.add(g_n), o_gids = P.add(g_ids), o_ft = P.add(fill_type), o_fo = P.add(fill_off),
               o_fn = P.add(fill_ns), o_st = P.add(stroke_type), o_so = P.add(stroke_off), o_sn = P.add(stroke_ns),
               o_color = P.add(std::vector<float>(d.color_params, d.color_params + d.num_color_params)),
               o_eo = P.add(even_odd), o_s2c = P.add(s2c), o_c2s = P.add(c2s), o_glb = P.add(g_lbox),
               o_gb = P.add(g_box), o_gmr = P.add(g_maxr), o_gsdf = P.add(g_sdf), o_top = P.add(topd),
               o_len = P.add(shape_length), o_cdf = P.add(cdf), o_pmf = P.add(pmf), o_sid = P.add(sid),
               o_gid = P.add(gid);
        // topology tables and scratch of the device rebuild (dvg_scene_update)
        size_t o_sperm = P.add(seg_perm), o_gdesc = P.add(g_ids_desc), o_gperm = P.add(g_perm),
               o_ganchor = P.add(g_anchor), o_gflags = P.add(g_sdf_flags), o_tl0 = P.add(thick_leaf0),
               o_order = P.add(std::vector<int>(nseg, 0)), o_gorder = P.add(std::vector<int>(total, 0)),
               o_long = P.add(long_ids), o_bigg = P.add(big_gids);
        sc.block_bytes = P.host.size();
        sc.top_offset = o_top;
        sc.top_bytes = topd.size() * sizeof(BNode);
        sc.scratch_offset[0] = o_order;
        sc.scratch_bytes[0] = (size_t)nseg * sizeof(int);
        sc.scratch_offset[1] = o_gorder;
        sc.scratch_bytes[1] = (size_t)total * sizeof(int);
        // validation and host-side construction are done: only now is a device needed
        if (gpu_index >= 0) HIP_OK(hipSetDevice(gpu_index));
        HIP_OK(hipGetDevice(&sc.device));
        sc.device_set = true;
        sc.block = cache_alloc(sc.device, sc.block_bytes, &sc.block_granted);
        if (!sc.block) return set_error("out of device memory for the scene");
        if (int e = upload(sc.device, sc.block, P.host.data(), sc.block_bytes)) return e;
        uint8_t *B = (uint8_t *)sc.block;
        DevScene &D = sc.dev;
        D.shape_type = (const int *)(B + o_shape_type);
        D.stroke_width = (const float *)(B + o_sw);
        D.param_off = (const int *)(B + o_poff);
        D.params = (const float *)(B + o_params);
        D.seg_begin = (const int *)(B + o_sb);
        D.nsegs = (const int *)(B + o_nsegs);
        D.pt_off = (const int *)(B + o_ptoff);
        D.npt = (const int *)(B + o_npt);
        D.has_thick = (const uint8_t *)(B + o_thick_flag);
        D.closed = (const uint8_t *)(B + o_closed);
        D.approx = (const uint8_t *)(B + o_approx);
        D.shape_box = (const float4 *)(B + o_sbox);
        D.seg_idx = (const int4 *)(B + o_sidx);
        D.seg_kind = (const int *)(B + o_skind);
        D.seg_pid = (const int *)(B + o_spid);
        D.seg_box = (const float4 *)(B + o_sbx);
        D.seg_maxr = (const float *)(B + o_smr);
        D.seg_wf = (const float4 *)(B + o_swf);
        D.seg_dfs = (const int *)(B + o_sdfs);
        D.seg_cdf = (const float *)(B + o_scdf);
        D.seg_pmf = (const float *)(B + o_spmf);
        D.points = (const float2 *)(B + o_pts);
        D.thickness = (const float *)(B + o_th);
        D.g_off = (const int *)(B + o_goff);
        D.g_n = (const int *)(B + o_gn);
        D.g_ids = (const int *)(B + o_gids);
        D.fill_type = (const int *)(B + o_ft);
        D.fill_off = (const int *)(B + o_fo);
        D.fill_ns = (const int *)(B + o_fn);
        D.stroke_type = (const int *)(B + o_st);
        D.stroke_off = (const int *)(B + o_so);
        D.stroke_ns = (const int *)(B + o_sn);
        D.color = (const float *)(B + o_color);
        D.even_odd = (const uint8_t *)(B + o_eo);
        D.s2c = (const float *)(B + o_s2c);
        D.c2s = (const float *)(B + o_c2s);
        D.g_lbox = (const float4 *)(B + o_glb);
        D.g_box = (const float4 *)(B + o_gb);
        D.g_maxr = (const float *)(B + o_gmr);
        D.g_sdf = (const float4 *)(B + o_gsdf);
        D.top = (const BNode *)(B + o_top);
        D.shape_length = (const float *)(B + o_len);
        D.shapes_cdf = (const float *)(B + o_cdf);
        D.shapes_pmf = (const float *)(B + o_pmf);
        D.sample_shape_id = (const int *)(B + o_sid);
        D.sample_group_id = (const int *)(B + o_gid);
        RefitTables &R = sc.refit;
        R.seg_perm = (const int *)(B + o_sperm);
        R.g_ids_desc = (const int *)(B + o_gdesc);
        R.g_perm = (const int *)(B + o_gperm);
        R.g_anchor = (const int *)(B + o_ganchor);
        R.g_sdf_flags = (const int *)(B + o_gflags);
        R.thick_leaf0 = (float *)(B + o_tl0);
        R.order = (int *)(B + o_order);
        R.gorder = (int *)(B + o_gorder);
        R.long_ids = (const int *)(B + o_long);
        R.n_long = (int)long_ids.size();
        R.big_gids = (const int *)(B + o_bigg);
        R.n_big = (int)big_gids.size();
    }
    DevScene &D = sc.dev;
    D.cw = sc.cw;
    D.ch = sc.ch;
    D.ns = ns;
    D.ng = ns > 0 ? ng : 0;
    D.nseg = nseg;
    D.npts = d.num_points_total;
    D.filter_type = d.filter_type;
    D.filter_radius = d.filter_radius;
    if (ns > 0 && nseg > 0) {
        if (int e = prep_segments(sc, nullptr)) return e;
    }
    // gradient block: params | points | thickness | stroke width | colours | xforms | filter radius
    auto al = [](size_t n) { return (n + 63) & ~size_t(63); };
    sc.g_params = 0;
    sc.g_points = sc.g_params + al(d.num_shape_params);
    sc.g_thick = sc.g_points + al(2 * (size_t)d.num_points_total);
    sc.g_sw = sc.g_thick + al(d.num_points_total);
    sc.g_color = sc.g_sw + al(ns);
    sc.g_xform = sc.g_color + al(d.num_color_params);
    sc.g_filter = sc.g_xform + al(9 * (size_t)ng);
    sc.grad_floats = sc.g_filter + 64;
    sc.grad = (float *)cache_alloc(sc.device, sc.grad_floats * sizeof(float), &sc.grad_granted);
    if (!sc.grad) return set_error("out of device memory for the gradients");
    if (int e = clear_device(sc.device, sc.grad, sc.grad_floats * sizeof(float))) return e;
    D.d_params = sc.grad + sc.g_params;
    D.d_points = sc.grad + sc.g_points;
    D.d_thickness = sc.grad + sc.g_thick;
    D.d_stroke_width = sc.grad + sc.g_sw;
    D.d_color = sc.grad + sc.g_color;
    D.d_xform = sc.grad + sc.g_xform;
    D.d_filter_radius = sc.grad + sc.g_filter;
    size_t eg = 0;
    sc.err = (int *)cache_alloc(sc.device, kErrInts * sizeof(int), &eg);
    if (!sc.err) return set_error("out of device memory");
    if (int e = clear_device(sc.device, sc.err, kErrInts * sizeof(int))) return e;
    return 0;
}

// The top-level BVH (scene.cpp:648-683) of a scene whose parameters dvg_scene_update
// rewrote: its leaves are the groups' canvas boxes and radii, which the device rebuild
// keeps current (k_refit_groups), so they are read back and the tree is built by the
// same host code as in build_scene (std::sort by Morton code, pairwise build with carry).
// Waits for the scene's last stream; the upload is complete when this returns.
int rebuild_top(SceneHost &sc) {
    const int ng = sc.num_groups;
    if (ng > 0) {
        if (sc.last_stream) HIP_OK(hipStreamSynchronize(sc.last_stream));
        std::vector<float4> gb(ng);
        std::vector<float> gr(ng);
        HIP_OK(hipMemcpy(gb.data(), sc.dev.g_box, ng * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(gr.data(), sc.dev.g_maxr, ng * sizeof(float), hipMemcpyDeviceToHost));
        std::vector<HNode> top(2 * (size_t)ng - 1, HNode{-1, -1, box_empty(), 0.f});
        for (int g = 0; g < ng; g++) top[g] = HNode{g, -1, HBox{gb[g].x, gb[g].y, gb[g].z, gb[g].w}, gr[g]};
        build_bvh(top.data(), ng, true, sc.cw, sc.ch);
        std::vector<BNode> topd(top.size());
        for (size_t i = 0; i < top.size(); i++)
            topd[i] = BNode{top[i].child0, top[i].child1, top[i].box.minx, top[i].box.miny, top[i].box.maxx,
                            top[i].box.maxy, top[i].max_radius, 0};
        if (topd.size() * sizeof(BNode) > sc.top_bytes) return set_error("top-level BVH larger than its slot");
        if (int e = upload(sc.device, (uint8_t *)sc.block + sc.top_offset, topd.data(), topd.size() * sizeof(BNode)))
            return e;
    }
    sc.top_stale = false;
    return 0;
}

void free_scene(SceneHost &sc) {
    if (!sc.device_set) return;
    int prev = -1;
    if (hipGetDevice(&prev) == hipSuccess) (void)hipSetDevice(sc.device);
    cache_free(sc.device, sc.block, sc.block_granted);
    cache_free(sc.device, sc.grad, sc.grad_granted);
    cache_free(sc.device, sc.err, kErrInts * sizeof(int));
    if (sc.frag) cache_free(sc.device, sc.frag, sc.frag_granted);
    if (sc.pkeep) cache_free(sc.device, sc.pkeep, sc.pkeep_granted);
    sc.pkeep = nullptr;
    sc.pkeep_valid = false;
    sc.block = nullptr;
    sc.grad = nullptr;
    sc.err = nullptr;
    sc.frag = nullptr;
    sc.frag_valid = false;
    if (prev >= 0) (void)hipSetDevice(prev);
}

}  // namespace dvg

using namespace dvg;

struct dvg_scene {
    SceneHost h;
};

extern "C" {

const char *dvg_last_error(void) { return t_error.c_str(); }
int dvg_abi_version(void) { return DVG_ABI_VERSION; }

int dvg_scene_create(const dvg_scene_desc *desc, int gpu_index, dvg_scene **out) {
    if (!desc || !out) return set_error("null argument");
    t_error.clear();
    auto *s = new (std::nothrow) dvg_scene();
    if (!s) return set_error("out of host memory");
    int prev = -1;
    (void)hipGetDevice(&prev);
    int rc = build_scene(*desc, gpu_index, s->h);
    if (prev >= 0) (void)hipSetDevice(prev);
    if (rc != 0) {
        free_scene(s->h);
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int dvg_scene_destroy(dvg_scene *scene) {
    if (!scene) return 0;
    // memory goes back to the cache, so every kernel that may still read it must be done
    (void)hipSetDevice(scene->h.device);
    (void)hipStreamSynchronize(scene->h.last_stream);
    free_scene(scene->h);
    delete scene;
    return 0;
}

int dvg_scene_set_tiling(dvg_scene *scene, int tile_width) {
    if (!scene) return set_error("null scene");
    if (tile_width < 0 || tile_width > 256 || (tile_width & (tile_width - 1)))
        return set_error("tile width must be 0 (default) or a power of two up to 256");
    scene->h.tile_w = tile_width;
    return 0;
}

int dvg_scene_set_samples_per_wave(dvg_scene *scene, int edge_samples_per_wave) {
    if (!scene) return set_error("null scene");
    if (edge_samples_per_wave != 0 && edge_samples_per_wave != 32 && edge_samples_per_wave != 64)
        return set_error("edge samples per wave must be 0 (default), 32 or 64");
    scene->h.edge_spw = edge_samples_per_wave == 64 ? 64 : 32;
    return 0;
}

int dvg_scene_set_mode(dvg_scene *scene, int mode) {
    if (!scene) return set_error("null scene");
    if (mode != DVG_MODE_TILED && mode != DVG_MODE_PLAIN) return set_error("unknown render mode");
    scene->h.mode = mode;
    return 0;
}

int dvg_render(dvg_scene *scene, const dvg_render_args *args) {
    if (!scene || !args) return set_error("null argument");
    const dvg_render_args &a = *args;
    if (a.width <= 0 || a.height <= 0 || a.num_samples_x <= 0 || a.num_samples_y <= 0)
        return set_error("width, height and sample counts must be positive");
    if ((int64_t)a.width * a.height * a.num_samples_x * a.num_samples_y > (int64_t)0x7fffffff)
        return set_error("too many samples for one render");
    if (a.num_eval_positions < 0) return set_error("negative eval position count");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != scene->h.device) (void)hipSetDevice(scene->h.device);
    int rc;
    bool plain = scene->h.mode == DVG_MODE_PLAIN || a.num_samples_x * a.num_samples_y > 256;
    rc = plain ? render_plain(scene->h, a) : render_tiled(scene->h, a);
    scene->h.last_stream = (hipStream_t)a.stream;
    if (prev >= 0 && prev != scene->h.device) (void)hipSetDevice(prev);
    return rc;
}

int dvg_render_noalloc(dvg_scene *scene, const dvg_render_args *args) {
    if (!scene || !args) return set_error("null argument");
    const dvg_render_args &a = *args;
    if (a.width <= 0 || a.height <= 0 || a.num_samples_x <= 0 || a.num_samples_y <= 0)
        return set_error("width, height and sample counts must be positive");
    if (a.num_eval_positions < 0) return set_error("negative eval position count");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != scene->h.device) (void)hipSetDevice(scene->h.device);
    int rc = render_noalloc(scene->h, a);
    scene->h.last_stream = (hipStream_t)a.stream;
    if (prev >= 0 && prev != scene->h.device) (void)hipSetDevice(prev);
    return rc;
}

struct dvg_batch {
    dvg::BatchHost *b;
};

int dvg_batch_create(dvg_scene *const *scenes, const dvg_render_args *args, int n, int flags, dvg_batch **out) {
    if (!scenes || !args || !out) return set_error("null argument");
    t_error.clear();
    std::vector<SceneHost *> hs(n > 0 ? n : 0);
    for (int i = 0; i < n; i++) {
        if (!scenes[i]) return set_error("null scene in batch");
        hs[i] = &scenes[i]->h;
    }
    dvg::BatchHost *b = nullptr;
    if (int rc = dvg::batch_create(hs.data(), args, n, flags, &b)) return rc;
    *out = new dvg_batch{b};
    return 0;
}

int dvg_batch_render(dvg_batch *batch, void *stream) {
    if (!batch) return set_error("null batch");
    int rc = dvg::batch_render(*batch->b, (hipStream_t)stream);
    for (auto *sc : dvg::batch_scenes(*batch->b)) sc->last_stream = (hipStream_t)stream;
    return rc;
}

int dvg_batch_status(dvg_batch *batch) {
    if (!batch) return set_error("null batch");
    return dvg::batch_status(*batch->b);
}

int dvg_batch_destroy(dvg_batch *batch) {
    if (!batch) return 0;
    dvg::batch_destroy(batch->b);
    delete batch;
    return 0;
}

int dvg_scene_update(dvg_scene *scene, const dvg_scene_params *params, float filter_radius, void *stream) {
    if (!scene || !params) return set_error("null argument");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (scene->h.device_set && prev != scene->h.device) (void)hipSetDevice(scene->h.device);
    int rc = update_scene(scene->h, *params, filter_radius, (hipStream_t)stream);
    if (prev >= 0 && scene->h.device_set && prev != scene->h.device) (void)hipSetDevice(prev);
    return rc;
}

int dvg_scene_status(dvg_scene *scene) {
    if (!scene) return set_error("null scene");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (scene->h.device_set && prev != scene->h.device) (void)hipSetDevice(scene->h.device);
    int rc = scene_status(scene->h);
    if (prev >= 0 && scene->h.device_set && prev != scene->h.device) (void)hipSetDevice(prev);
    return rc;
}

int dvg_scene_debug_block(dvg_scene *scene, void *dst, int64_t cap, int64_t *bytes) {
    if (!scene) return set_error("null scene");
    SceneHost &h = scene->h;
    if (bytes) *bytes = (int64_t)h.block_bytes;
    if (!dst || !h.block || h.block_bytes == 0) return 0;
    if (cap < (int64_t)h.block_bytes) return set_error("destination smaller than the scene block");
    (void)hipSetDevice(h.device);
    if (h.last_stream) HIP_OK(hipStreamSynchronize(h.last_stream));
    HIP_OK(hipMemcpy(dst, h.block, h.block_bytes, hipMemcpyDeviceToHost));
    // the PLAIN schedule's top-level BVH (not rebuilt by dvg_scene_update) and the
    // rebuild's sort scratch are zeroed, so blocks of equal scenes compare equal
    uint8_t *b = (uint8_t *)dst;
    std::memset(b + h.top_offset, 0, h.top_bytes);
    for (int i = 0; i < 2; i++) std::memset(b + h.scratch_offset[i], 0, h.scratch_bytes[i]);
    return 0;
}

int dvg_scene_zero_grads(dvg_scene *scene, void *stream) {
    if (!scene) return set_error("null scene");
    SceneHost &h = scene->h;
    hipError_t e = hipMemsetAsync(h.grad, 0, h.grad_floats * sizeof(float), (hipStream_t)stream);
    if (e != hipSuccess) return set_error(std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return 0;
}

int dvg_scene_grad_layout(dvg_scene *scene, int64_t *offsets, int64_t *total_floats) {
    if (!scene) return set_error("null scene");
    SceneHost &h = scene->h;
    if (offsets) {
        offsets[0] = (int64_t)h.g_params;
        offsets[1] = (int64_t)h.g_points;
        offsets[2] = (int64_t)h.g_thick;
        offsets[3] = (int64_t)h.g_sw;
        offsets[4] = (int64_t)h.g_color;
        offsets[5] = (int64_t)h.g_xform;
        offsets[6] = (int64_t)h.g_filter;
    }
    if (total_floats) *total_floats = (int64_t)h.grad_floats;
    return 0;
}

int dvg_scene_copy_grads(dvg_scene *scene, float *dst, void *stream) {
    if (!scene || !dst) return set_error("null argument");
    SceneHost &h = scene->h;
    HIP_OK(hipMemcpyAsync(dst, h.grad, h.grad_floats * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return 0;
}

int dvg_scene_get_grads(dvg_scene *scene, dvg_grads *o) {
    if (!scene || !o) return set_error("null argument");
    SceneHost &h = scene->h;
    if (h.last_stream) HIP_OK(hipStreamSynchronize(h.last_stream));
    HIP_OK(hipDeviceSynchronize());
    if (int e = scene_status(h)) return e;
    auto cp = [&](float *dst, size_t off, size_t n) -> int {
        if (dst && n) HIP_OK(hipMemcpy(dst, h.grad + off, n * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    };
    if (cp(o->d_shape_params, h.g_params, h.num_shape_params)) return 1;
    if (cp(o->d_points, h.g_points, 2 * (size_t)h.num_points_total)) return 1;
    if (cp(o->d_thickness, h.g_thick, h.num_points_total)) return 1;
    if (cp(o->d_stroke_width, h.g_sw, h.num_shapes)) return 1;
    if (cp(o->d_color_params, h.g_color, h.num_color_params)) return 1;
    if (cp(o->d_shape_to_canvas, h.g_xform, 9 * (size_t)h.num_groups)) return 1;
    if (cp(o->d_filter_radius, h.g_filter, 1)) return 1;
    return 0;
}

}  // extern "C"
