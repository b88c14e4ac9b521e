// dvg_kernels.hip -- render kernels for gfx950 and their launch sequences.
//
// TILED schedule (default):
//   k_weight      per pixel: gathers the filter weights of the samples in its footprint
//                 (the reference's weight_kernel, diffvg.cpp:1115-1158, as a gather)
//   k_color       one 256-lane workgroup (four wave64) per pixel tile: the workgroup
//                 bins the groups whose inflated top-level box meets the tile into LDS,
//                 in increasing group id, then every lane shades one sample against
//                 that list, compositing fragments as they are found
//   k_resolve     per pixel: filter-weighted gather of the sample colours
//   k_color_bwd   per tile: recompute fragments, reverse compositing sweep, paint
//                 adjoints, filter-radius adjoint reduced per workgroup
//   k_edge_sample / k_edge_bin / k_edge   boundary samples binned by pixel tile, both
//                 sides of each boundary shaded against the tile's candidate list
// PLAIN schedule: one thread per sample in 64-thread blocks with per-sample top-level
// BVH traversal, fragment arrays sorted by insertion, atomic splats and Morton-sorted
// edge samples -- the reference's kernel sequence (diffvg.cpp:1477-1649).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <mutex>
#include <vector>

#include "dvg_build.h"
#include "dvg_internal.h"
#include "dvg_shade.h"
#include "dvg_sort.h"

namespace dvg {

constexpr int kBlock = 256;
constexpr int kCandCap = 2048;
// boundary samples per edge workgroup: kEdgeChunk when each lane shades one side of a
// sample (32 samples per wave, the default), kBlock when each lane shades both sides (64
// per wave); Geom::echunk holds the one a render uses (a tuning knob)
constexpr int kEdgeChunk = kBlock / 2;
// fragment lists kept from a colour forward for its backward: up to kFragSlots fragments
// per sample (a sample with more is shaded again), at most kFragCacheBytes per scene
constexpr int kFragSlots = 16;
constexpr size_t kFragCacheBytes = size_t(512) << 20;
constexpr int kFragMaxGroups = 32768;  // group ids 2 g + is_stroke are kept in 16 bits
// prefiltered fragments keep what their adjoint needs besides the paint: group id, closest
// shape, signed distance, closest point, path position and whether a distance was found
constexpr int kPFragSlots = 8;
constexpr int kPFragWords = 9;
// Waves per SIMD each heavy shading kernel is compiled for, i.e. its register budget
// (512 / waves VGPRs per lane; beyond it the compiler spills to scratch).  Chosen by
// measurement on the 300-path and 2,048-stroke scenes: these kernels wait on memory and
// on long FP64 dependency chains, and 4-6 waves hide that better than 2 waves without
// spills (k_edge 5.03 -> 3.16 ms, k_color 1.45 -> 0.92 ms, k_color_bwd 1.90 -> 1.33 ms);
// the colour backward, which mostly replays kept fragment lists, is best at 4.
#ifndef DVG_WAVES_COLOR
#define DVG_WAVES_COLOR 5
#endif
#ifndef DVG_WAVES_COLOR_BWD
#define DVG_WAVES_COLOR_BWD 4
#endif
// the prefiltered instantiations keep more state per lane: on the 300-path scene
// k_color<1> takes 3.69 ms at 3 waves, 3.75 at 5, and k_color_bwd<1> 1.71 at 3, 1.69 at 4
// (the same within run-to-run spread); the batched forward stays at 5 (serving stream:
// 1.91 ms at 5, 2.16 at 4)
#ifndef DVG_WAVES_COLOR_PRE
#define DVG_WAVES_COLOR_PRE 3
#endif
#ifndef DVG_WAVES_BATCH_COLOR_PRE
#define DVG_WAVES_BATCH_COLOR_PRE 5
#endif
#ifndef DVG_WAVES_COLOR_BWD_PRE
#define DVG_WAVES_COLOR_BWD_PRE 3
#endif
#ifndef DVG_WAVES_EDGE
#define DVG_WAVES_EDGE 6
#endif
// the edge kernel with both sides of a sample per lane keeps two compositors live
#ifndef DVG_WAVES_EDGE_BOTH
#define DVG_WAVES_EDGE_BOTH 5
#endif
// the batched edge kernel (job lookup and per-job fields live across its loop) measured
// fastest at 5 (256 uneven scenes: 31.1 ms at 5, 33.1 at 4, 46.2 at 6, 46.6 at 7)
// the batched forward (job lookup live across its loop) measured fastest at 3 (256 uneven
// scenes: 13.2 ms at 3, 14.3 at 4, 20.7 at 5)
#ifndef DVG_WAVES_BATCH_COLOR
#define DVG_WAVES_BATCH_COLOR 3
#endif
#ifndef DVG_WAVES_BATCH_EDGE
#define DVG_WAVES_BATCH_EDGE 5
#endif
#ifndef DVG_WAVES_SDF
#define DVG_WAVES_SDF 5
#endif
// the batched backward (job lookup live across its loop; it mostly replays kept fragment
// lists) measured fastest at 4 (256 uneven scenes: 15.0 ms at 4, 16.3 at 3, 18.1 at 2)
#ifndef DVG_WAVES_BATCH_COLOR_BWD
#define DVG_WAVES_BATCH_COLOR_BWD 4
#endif
// the persistent batch kernel runs every stage in one launch, so one setting for all
#ifndef DVG_WAVES_PERSIST
#define DVG_WAVES_PERSIST 4
#endif

// Render geometry shared by the kernels.
struct Geom {
    int W, H, nsx, nsy, spp;
    uint64_t seed;
    int tw, th, tiles_x, tiles_y;
    float sx_scale, sy_scale;  // canvas / render
    float margin;
    int echunk;  // boundary samples per edge workgroup: kEdgeChunk or kBlock
};

// Workgroup-shared storage of the tiled stages.  Each stage function takes its own
// struct from the caller: the single-stage kernels declare one, the persistent batch
// kernel overlays all of them in one union.
struct CandLds {
    int list[kCandCap];
    int w[kBlock / 64];
};
struct ColorBwdLds {
    CandLds c;
    float red[kBlock / 64];
    float4 dbg[kBlock];
    int dbg_kind[kBlock];
};
struct EdgeLds {
    CandLds c;
    int perm[kBlock];
    int key[kBlock];
};
struct SdfLds {
    CandLds c;
    float red[kBlock / 64];
};
template <int NT>
struct ScanLds {
    int s[NT];
    int sc[NT];
    int carry, carry_c;
};

DVG_DEV V2 sample_pos(const Geom &G, int x, int y, int sx, int sy, int idx, bool centre) {
    Pcg rng = pcg_init(idx, G.seed);
    float rx = pcg_float(rng);
    float ry = pcg_float(rng);
    if (centre) rx = ry = 0.5f;
    return V2{x + ((float)sx + rx) / G.nsx, y + ((float)sy + ry) / G.nsy};
}
// render-pixel position -> canvas position, as render_kernel + sample_color do it
DVG_DEV V2 to_canvas(const DevScene &S, const Geom &G, V2 pt) {
    V2 npt{pt.x / G.W, pt.y / G.H};
    return V2{npt.x * S.cw, npt.y * S.ch};
}

// Ordered compaction of the groups in [g0, g1) whose inflated top-level box meets `tb`.
__device__ int build_candidates(const DevScene &S, float4 tb, int g0, int g1, int *list, int *wsum) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
    int total = 0;
    for (int base = g0; base < g1; base += blockDim.x) {
        int g = base + tid;
        bool hit = false;
        if (g < g1) {
            float4 b = S.g_box[g];
            float r = S.g_maxr[g];
            // a single group is the root of the top-level BVH and is visited untested
            hit = S.ng == 1 || !(b.x - r > tb.z || b.z + r < tb.x || b.y - r > tb.w || b.w + r < tb.y);
        }
        uint64_t m = __ballot(hit);
        int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wid] = __popcll(m);
        __syncthreads();
        int off = total, chunk = 0;
        for (int w = 0; w < nw; w++) {
            if (w < wid) off += wsum[w];
            chunk += wsum[w];
        }
        if (hit) list[off + pre] = g;
        total += chunk;
        __syncthreads();
    }
    return total;
}

DVG_DEV float4 tile_box(const DevScene &S, const Geom &G, int tx, int ty, float extra) {
    int x0 = tx * G.tw, y0 = ty * G.th;
    int x1 = min(G.W, x0 + G.tw), y1 = min(G.H, y0 + G.th);
    float m = G.margin + extra;
    return make_float4((float)x0 * G.sx_scale - m, (float)y0 * G.sy_scale - m, (float)x1 * G.sx_scale + m,
                       (float)y1 * G.sy_scale + m);
}

// Sample of local index `si` in tile (tx, ty); returns false past the image edge.
DVG_DEV bool tile_sample(const Geom &G, int tx, int ty, int si, int &x, int &y, int &sx, int &sy, int &idx) {
    int tile_px = G.tw * G.th;
    if (si >= tile_px * G.spp) return false;
    int p = si / G.spp, sub = si - p * G.spp;
    x = tx * G.tw + p % G.tw;
    y = ty * G.th + p / G.tw;
    if (x >= G.W || y >= G.H) return false;
    sx = sub % G.nsx;
    sy = sub / G.nsx;
    idx = ((y * G.W + x) * G.nsy + sy) * G.nsx + sx;
    return true;
}

// Shades one canvas point against a candidate list (sample_color, diffvg.cpp:525-653).
DVG_DEV void shade_list(const DevScene &S, V2 pt, const int *list, int n, Compositor &comp, EdgeQuery *eq,
                        bool &trav_hit) {
    int target = eq ? eq->group : -1;
    for (int i = 0; i < n; i++) {
        int g = list[i];
        if (S.ng > 1 && !in_box(S.g_box[g], pt, S.g_maxr[g])) continue;
        M3 c2s = m3_load(S.c2s + 9 * g);
        V2 local = xform_pt(c2s, pt);
        int st = S.stroke_type[g];
        if (st >= 0) {
            EdgeQuery q{target, eq ? eq->shape : -1, false};
            if (within_distance_local(S, g, local, eq ? &q : nullptr)) {
                comp.add(sample_paint(S, st, S.stroke_off[g], S.stroke_ns[g], pt), g, target);
            }
            trav_hit |= q.hit;
        }
        int ft = S.fill_type[g];
        if (ft >= 0) {
            EdgeQuery q{target, eq ? eq->shape : -1, false};
            if (is_inside_local(S, g, local, eq ? &q : nullptr)) {
                comp.add(sample_paint(S, ft, S.fill_off[g], S.fill_ns[g], pt), g, target);
            }
            trav_hit |= q.hit;
        }
    }
}

// Prefiltered shading of one candidate list (sample_color_prefiltered, diffvg.cpp:835-994).
template <bool RECORD>
DVG_DEV void shade_list_prefiltered(const DevScene &S, V2 pt, const int *list, int n, Compositor &comp,
                                    PFragRec *frags, int &overflow) {
    const float inf = __builtin_huge_valf();
    for (int i = 0; i < n; i++) {
        int g = list[i];
        if (S.ng > 1 && !in_box(S.g_box[g], pt, S.g_maxr[g])) continue;
        int st = S.stroke_type[g];
        if (st >= 0) {
            int ms = -1;
            V2 cp{0, 0};
            PathInfo info{-1, -1, 0.f};
            float d = inf;
            compute_distance(S, g, pt, inf, ms, cp, info, d);
            if (ms >= 0) {
                float sw = S.stroke_width[ms];
                float ad = (float)fabs((double)d);
                float w = smoothstep(ad + sw) - smoothstep(ad - sw);
                if (w > 0) {
                    V4 c = sample_paint(S, st, S.stroke_off[g], S.stroke_ns[g], pt);
                    c.w *= w;
                    if (RECORD) {
                        if (comp.nf < kMaxPFrags)
                            frags[comp.nf] = PFragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, 2 * g + 1, ms,
                                                      d, cp, info, 1};
                        else
                            overflow = 1;
                    }
                    comp.add(c, g, -1);
                }
            } else {
                overflow = 2;  // the reference asserts here (diffvg.cpp:868)
            }
        }
        int ft = S.fill_type[g];
        if (ft >= 0) {
            int ms = -1;
            V2 cp{0, 0};
            PathInfo info{-1, -1, 0.f};
            float d = inf;
            bool found = compute_distance(S, g, pt, 1.f, ms, cp, info, d);
            M3 c2s = m3_load(S.c2s + 9 * g);
            bool in = is_inside_local(S, g, xform_pt(c2s, pt), nullptr);
            if (found || in) {
                if (!in) d = -d;
                float w = smoothstep(d);
                if (w > 0) {
                    V4 c = sample_paint(S, ft, S.fill_off[g], S.fill_ns[g], pt);
                    c.w *= w;
                    if (RECORD) {
                        if (comp.nf < kMaxPFrags)
                            frags[comp.nf] = PFragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, 2 * g, ms, d,
                                                      cp, info, found ? 1 : 0};
                        else
                            overflow = 1;
                    }
                    comp.add(c, g, -1);
                }
            }
        }
    }
}

// Shading with the fragment list kept for the adjoint (non-prefiltered).
// shade_list_prefiltered<false> that also writes, per fragment it composites (at most
// kPFragSlots), the words replay_pfrags needs: keep[(slot * kPFragWords + w) * nsamp + idx].
DVG_DEV void shade_list_prefiltered_keep(const DevScene &S, V2 pt, const int *list, int n, Compositor &comp,
                                         int &overflow, int *keep, size_t nsamp, int idx) {
    const float inf = __builtin_huge_valf();
    auto put = [&](int gid, int ms, float d, V2 cp, const PathInfo &info, int within) {
        if (comp.nf >= kPFragSlots) return;
        int *k = keep + (size_t)comp.nf * kPFragWords * nsamp + idx;
        k[0] = gid;
        k[nsamp] = ms;
        k[2 * nsamp] = __float_as_int(d);
        k[3 * nsamp] = __float_as_int(cp.x);
        k[4 * nsamp] = __float_as_int(cp.y);
        k[5 * nsamp] = info.base_point_id;
        k[6 * nsamp] = info.point_id;
        k[7 * nsamp] = __float_as_int(info.t_root);
        k[8 * nsamp] = within;
    };
    for (int i = 0; i < n; i++) {
        int g = list[i];
        if (S.ng > 1 && !in_box(S.g_box[g], pt, S.g_maxr[g])) continue;
        int st = S.stroke_type[g];
        if (st >= 0) {
            int ms = -1;
            V2 cp{0, 0};
            PathInfo info{-1, -1, 0.f};
            float d = inf;
            compute_distance(S, g, pt, inf, ms, cp, info, d);
            if (ms >= 0) {
                float sw = S.stroke_width[ms];
                float ad = (float)fabs((double)d);
                float w = smoothstep(ad + sw) - smoothstep(ad - sw);
                if (w > 0) {
                    V4 c = sample_paint(S, st, S.stroke_off[g], S.stroke_ns[g], pt);
                    c.w *= w;
                    put(2 * g + 1, ms, d, cp, info, 1);
                    comp.add(c, g, -1);
                }
            } else {
                overflow = 2;  // the reference asserts here (diffvg.cpp:868)
            }
        }
        int ft = S.fill_type[g];
        if (ft >= 0) {
            int ms = -1;
            V2 cp{0, 0};
            PathInfo info{-1, -1, 0.f};
            float d = inf;
            bool found = compute_distance(S, g, pt, 1.f, ms, cp, info, d);
            M3 c2s = m3_load(S.c2s + 9 * g);
            bool in = is_inside_local(S, g, xform_pt(c2s, pt), nullptr);
            if (found || in) {
                if (!in) d = -d;
                float w = smoothstep(d);
                if (w > 0) {
                    V4 c = sample_paint(S, ft, S.fill_off[g], S.fill_ns[g], pt);
                    c.w *= w;
                    put(2 * g, ms, d, cp, info, found ? 1 : 0);
                    comp.add(c, g, -1);
                }
            }
        }
    }
}

// The prefiltered fragment records shade_list_prefiltered<true> makes, from the words a
// forward kept: paint and coverage weight recomputed from the kept distance.
DVG_DEV void replay_pfrags(const DevScene &S, V2 pt, const int *keep, size_t nsamp, int idx, int nf,
                           Compositor &comp, PFragRec *frags) {
    for (int j = 0; j < nf; j++) {
        const int *k = keep + (size_t)j * kPFragWords * nsamp + idx;
        const int gid = k[0], ms = k[nsamp];
        const float d = __int_as_float(k[2 * nsamp]);
        const V2 cp{__int_as_float(k[3 * nsamp]), __int_as_float(k[4 * nsamp])};
        const PathInfo info{k[5 * nsamp], k[6 * nsamp], __int_as_float(k[7 * nsamp])};
        const int within = k[8 * nsamp];
        const int g = gid >> 1;
        V4 c;
        if (gid & 1) {
            float sw = S.stroke_width[ms];
            float ad = (float)fabs((double)d);
            float w = smoothstep(ad + sw) - smoothstep(ad - sw);
            c = sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt);
            c.w *= w;
        } else {
            float w = smoothstep(d);
            c = sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt);
            c.w *= w;
        }
        frags[comp.nf] = PFragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, gid, ms, d, cp, info, within};
        comp.add(c, g, -1);
    }
}

DVG_DEV void shade_list_record(const DevScene &S, V2 pt, const int *list, int n, Compositor &comp, FragRec *frags,
                               int &overflow) {
    for (int i = 0; i < n; i++) {
        int g = list[i];
        if (S.ng > 1 && !in_box(S.g_box[g], pt, S.g_maxr[g])) continue;
        M3 c2s = m3_load(S.c2s + 9 * g);
        V2 local = xform_pt(c2s, pt);
        int st = S.stroke_type[g];
        if (st >= 0 && within_distance_local(S, g, local, nullptr)) {
            V4 c = sample_paint(S, st, S.stroke_off[g], S.stroke_ns[g], pt);
            if (comp.nf < kMaxFrags)
                frags[comp.nf] = FragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, 2 * g + 1};
            else
                overflow = 1;
            comp.add(c, g, -1);
        }
        int ft = S.fill_type[g];
        if (ft >= 0 && is_inside_local(S, g, local, nullptr)) {
            V4 c = sample_paint(S, ft, S.fill_off[g], S.fill_ns[g], pt);
            if (comp.nf < kMaxFrags)
                frags[comp.nf] = FragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, 2 * g};
            else
                overflow = 1;
            comp.add(c, g, -1);
        }
    }
}

// shade_list that also writes the group ids of the fragments it composites (2 g +
// is_stroke, at most kFragSlots of them) to ids[slot * nsamp + idx].
DVG_DEV void shade_list_ids(const DevScene &S, V2 pt, const int *list, int n, Compositor &comp, uint16_t *ids,
                            size_t nsamp, int idx) {
    for (int i = 0; i < n; i++) {
        int g = list[i];
        if (S.ng > 1 && !in_box(S.g_box[g], pt, S.g_maxr[g])) continue;
        M3 c2s = m3_load(S.c2s + 9 * g);
        V2 local = xform_pt(c2s, pt);
        int st = S.stroke_type[g];
        if (st >= 0 && within_distance_local(S, g, local, nullptr)) {
            if (comp.nf < kFragSlots) ids[(size_t)comp.nf * nsamp + idx] = (uint16_t)(2 * g + 1);
            comp.add(sample_paint(S, st, S.stroke_off[g], S.stroke_ns[g], pt), g, -1);
        }
        int ft = S.fill_type[g];
        if (ft >= 0 && is_inside_local(S, g, local, nullptr)) {
            if (comp.nf < kFragSlots) ids[(size_t)comp.nf * nsamp + idx] = (uint16_t)(2 * g);
            comp.add(sample_paint(S, ft, S.fill_off[g], S.fill_ns[g], pt), g, -1);
        }
    }
}

// The fragment records shade_list_record makes, from the group ids a forward kept: the
// same paints at the same point composited in the same order.
DVG_DEV void replay_frags(const DevScene &S, V2 pt, const uint16_t *ids, size_t nsamp, int idx, int nf,
                          Compositor &comp, FragRec *frags) {
    for (int j = 0; j < nf; j++) {
        const int gid = ids[(size_t)j * nsamp + idx];
        const int g = gid >> 1;
        const bool stroke = (gid & 1) != 0;
        V4 c = stroke ? sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt)
                      : sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt);
        frags[comp.nf] = FragRec{V3{c.x, c.y, c.z}, c.w, comp.color, comp.alpha, gid};
        comp.add(c, g, -1);
    }
}

DVG_DEV V4 load_bg(const float *bg, int pix) {
    float4 b = reinterpret_cast<const float4 *>(bg)[pix];
    return V4{b.x, b.y, b.z, b.w};
}

// ------------------------------------------------------------------ TILED kernels
DVG_DEV void weight_pixel(const DevScene &S, const Geom &G, int pix, float *weight) {
    if (pix >= G.W * G.H) return;
    int px = pix % G.W, py = pix / G.W;
    float radius = S.filter_radius;
    int ri = (int)ceil((double)radius);
    float xc = px + 0.5f, yc = py + 0.5f;
    float acc = 0.f;
    for (int y = py - ri; y <= py + ri; y++) {
        if (y < 0 || y >= G.H) continue;
        for (int x = px - ri; x <= px + ri; x++) {
            if (x < 0 || x >= G.W) continue;
            for (int sy = 0; sy < G.nsy; sy++)
                for (int sx = 0; sx < G.nsx; sx++) {
                    int idx = ((y * G.W + x) * G.nsy + sy) * G.nsx + sx;
                    // always jittered: the reference's weight_kernel never sets
                    // use_prefiltering (diffvg.cpp:1525-1533)
                    V2 pt = sample_pos(G, x, y, sx, sy, idx, false);
                    acc += filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y);
                }
        }
    }
    weight[pix] = acc;
}
__global__ __launch_bounds__(kBlock) void k_weight(DevScene S, Geom G, float *weight) {
    weight_pixel(S, G, blockIdx.x * blockDim.x + threadIdx.x, weight);
}

// Key of kept fragment lists: seed, sizes, sample counts and the scene's parameter
// version on the device.  The forward that writes the lists writes the key beside them;
// a backward replays them only when the key matches its own, so a backward that follows
// another render, or a device update, on the scene shades again -- also when it is
// replayed from a HIP graph.
constexpr int kKeyInts = 6;
DVG_DEV void key_of(const Geom &G, const int *ver, int *k) {
    k[0] = (int)(uint32_t)G.seed;
    k[1] = (int)(uint32_t)(G.seed >> 32);
    k[2] = G.W;
    k[3] = G.H;
    k[4] = G.nsx | (G.nsy << 16);
    k[5] = *ver;
}
DVG_DEV bool key_matches(const Geom &G, const int *ver, const int *hdr) {
    int k[kKeyInts];
    key_of(G, ver, k);
    bool ok = true;
#pragma unroll
    for (int i = 0; i < kKeyInts; i++) ok = ok && hdr[i] == k[i];
    return ok;
}

template <int PRE>
DVG_DEV void color_tile(CandLds &L, const DevScene &S, const Geom &G, int tile, const float *bg, float4 *colors, int *err,
                        uint16_t *frag = nullptr, int *pkeep = nullptr) {
    // the key of the lists this forward keeps (err: see SceneHost::err)
    if ((PRE ? pkeep != nullptr : frag != nullptr) && tile == 0 && threadIdx.x == 0)
        key_of(G, err + 2, err + (PRE ? 10 : 4));
    constexpr int prefilter = PRE;
    int *s_list = L.list;
    int *s_w = L.w;
    int tx = tile % G.tiles_x, ty = tile / G.tiles_x;
    int x, y, sx, sy, idx;
    bool valid = tile_sample(G, tx, ty, threadIdx.x, x, y, sx, sy, idx);
    V2 pt{0, 0};
    if (valid) pt = to_canvas(S, G, sample_pos(G, x, y, sx, sy, idx, prefilter != 0));
    bool has_bg = bg != nullptr;
    V4 bgc = (valid && has_bg) ? load_bg(bg, y * G.W + x) : V4{0, 0, 0, 0};
    Compositor comp;
    comp.init(has_bg, bgc);
    float4 tb = tile_box(S, G, tx, ty, 0.f);
    int overflow = 0;
    for (int c0 = 0; c0 < S.ng; c0 += kCandCap) {
        int n = build_candidates(S, tb, c0, min(S.ng, c0 + kCandCap), s_list, s_w);
        if (valid) {
            if constexpr (PRE) {
                if (pkeep)
                    shade_list_prefiltered_keep(S, pt, s_list, n, comp, overflow, pkeep, (size_t)G.W * G.H * G.spp,
                                                idx);
                else
                    shade_list_prefiltered<false>(S, pt, s_list, n, comp, nullptr, overflow);
            } else if (frag) {
                shade_list_ids(S, pt, s_list, n, comp, frag, (size_t)G.W * G.H * G.spp, idx);
            } else {
                bool th = false;
                shade_list(S, pt, s_list, n, comp, nullptr, th);
            }
        }
        // the list is rebuilt only for another range of groups: after the last one, a
        // wave that is done does not wait for the slower waves of its workgroup
        if (c0 + kCandCap < S.ng) __syncthreads();
    }
    if (valid) {
        V4 c = comp.result(has_bg, bgc);
        colors[idx] = make_float4(c.x, c.y, c.z, c.w);
        if (!PRE && frag) frag[(size_t)kFragSlots * G.W * G.H * G.spp + idx] = (uint16_t)min(comp.nf, 65535);
        if (PRE && pkeep) pkeep[(size_t)kPFragSlots * kPFragWords * G.W * G.H * G.spp + idx] = comp.nf;
    }
    if (overflow == 2) atomicOr(err, 2);
}
template <int PRE>
__global__ __launch_bounds__(kBlock, PRE ? DVG_WAVES_COLOR_PRE : DVG_WAVES_COLOR) void k_color(DevScene S, Geom G, const float *bg, float4 *colors, int *err,
                                                  uint16_t *frag, int *pkeep) {
    __shared__ CandLds L;
    color_tile<PRE>(L, S, G, blockIdx.x, bg, colors, err, frag, pkeep);
}

DVG_DEV void resolve_pixel(const DevScene &S, const Geom &G, int pix, int prefilter, const float *weight,
                           const float4 *colors, float *out) {
    if (pix >= G.W * G.H) return;
    float ws = weight[pix];
    if (!(ws > 0)) return;
    int px = pix % G.W, py = pix / G.W;
    float radius = S.filter_radius;
    int ri = (int)ceil((double)radius);
    float xc = px + 0.5f, yc = py + 0.5f;
    V4 acc{0, 0, 0, 0};
    for (int y = py - ri; y <= py + ri; y++) {
        if (y < 0 || y >= G.H) continue;
        for (int x = px - ri; x <= px + ri; x++) {
            if (x < 0 || x >= G.W) continue;
            for (int sy = 0; sy < G.nsy; sy++)
                for (int sx = 0; sx < G.nsx; sx++) {
                    int idx = ((y * G.W + x) * G.nsy + sy) * G.nsx + sx;
                    V2 pt = sample_pos(G, x, y, sx, sy, idx, prefilter != 0);
                    float fw = filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y);
                    float4 c = colors[idx];
                    acc += vdiv(fw * V4{c.x, c.y, c.z, c.w}, ws);
                }
        }
    }
    float4 *o = reinterpret_cast<float4 *>(out) + pix;
    float4 prev = *o;
    *o = make_float4(prev.x + acc.x, prev.y + acc.y, prev.z + acc.z, prev.w + acc.w);
}
__global__ __launch_bounds__(kBlock) void k_resolve(DevScene S, Geom G, int prefilter, const float *weight,
                                                    const float4 *colors, float *out) {
    resolve_pixel(S, G, blockIdx.x * blockDim.x + threadIdx.x, prefilter, weight, colors, out);
}

template <int PRE>
DVG_DEV void color_bwd_tile(ColorBwdLds &L, const DevScene &S, const Geom &G, int tile, const float *bg, const float *weight,
                            const float *d_img, float *d_bg, float *d_translation, int *err,
                            const uint16_t *frag = nullptr, const int *pkeep = nullptr) {
    constexpr int prefilter = PRE;
    int *s_list = L.c.list;
    int *s_w = L.c.w;
    float *s_red = L.red;
    float4 *s_dbg = L.dbg;
    int *s_dbg_kind = L.dbg_kind;
    int tx = tile % G.tiles_x, ty = tile / G.tiles_x;
    int x, y, sx, sy, idx;
    bool valid = tile_sample(G, tx, ty, threadIdx.x, x, y, sx, sy, idx);
    V2 rpt{0, 0}, pt{0, 0};
    if (valid) {
        rpt = sample_pos(G, x, y, sx, sy, idx, prefilter != 0);
        pt = to_canvas(S, G, rpt);
    }
    bool has_bg = bg != nullptr;
    V4 bgc = (valid && has_bg) ? load_bg(bg, y * G.W + x) : V4{0, 0, 0, 0};
    Compositor comp;
    comp.init(has_bg, bgc);
    FragRec frags[PRE ? 1 : kMaxFrags];
    PFragRec pfrags[PRE ? kMaxPFrags : 1];
    int overflow = 0;
    float4 tb = tile_box(S, G, tx, ty, 0.f);
    // a sample whose forward kept its fragment list replays it; the others shade
    const size_t nsamp = (size_t)G.W * G.H * G.spp;
    if (!key_matches(G, err + 2, err + (PRE ? 10 : 4))) {
        frag = nullptr;
        pkeep = nullptr;
    }
    int kept = -1;
    if (!PRE && frag && valid) kept = frag[(size_t)kFragSlots * nsamp + idx];
    if (PRE && pkeep && valid) kept = pkeep[(size_t)kPFragSlots * kPFragWords * nsamp + idx];
    const bool replay = kept >= 0 && kept <= (PRE ? kPFragSlots : kFragSlots);
    if (__syncthreads_or(valid && !replay)) {
        for (int c0 = 0; c0 < S.ng; c0 += kCandCap) {
            int n = build_candidates(S, tb, c0, min(S.ng, c0 + kCandCap), s_list, s_w);
            if (valid && !replay) {
                if constexpr (PRE)
                    shade_list_prefiltered<true>(S, pt, s_list, n, comp, pfrags, overflow);
                else
                    shade_list_record(S, pt, s_list, n, comp, frags, overflow);
            }
            if (c0 + kCandCap < S.ng) __syncthreads();
        }
    }
    if constexpr (!PRE) {
        if (valid && replay) replay_frags(S, pt, frag, nsamp, idx, kept, comp, frags);
    } else {
        if (valid && replay) replay_pfrags(S, pt, pkeep, nsamp, idx, kept, comp, pfrags);
    }
    float d_radius = 0.f;
    int kind = -1;
    V4 dbg_val{0, 0, 0, 0};
    V4 d_color{0, 0, 0, 0}, color{0, 0, 0, 0};
    float *d_tr = nullptr;
    if (valid) {
        d_color = gather_d_color(S.filter_type, S.filter_radius, d_img, weight, G.W, G.H, rpt);
        d_tr = d_translation ? d_translation + 2 * (y * G.W + x) : nullptr;
        color = comp.result(has_bg, bgc);
    }
    if constexpr (PRE) {
        // all lanes: paint, stroke-width and distance adjoints are combined across the wave
        int nf = valid ? min(comp.nf, kMaxPFrags) : 0;
        V4 d_back = composite_adjoint_prefiltered_wave(S, pfrags, nf, V3{color.x, color.y, color.z}, color.w,
                                                       d_color, pt, d_tr);
        if (valid) {
            kind = comp.nf == 0 ? 0 : 1;
            dbg_val = comp.nf == 0 ? d_color : d_back;
        }
    } else {
        // all lanes: the constant-paint adjoints are combined across the wave
        int nf = valid ? min(comp.nf, kMaxFrags) : 0;
        V4 d_back = composite_adjoint_wave(S, frags, nf, V3{color.x, color.y, color.z}, color.w, d_color, pt, d_tr);
        if (valid) {
            kind = comp.nf == 0 ? 0 : 1;
            dbg_val = comp.nf == 0 ? d_color : d_back;
        }
    }
    if (valid) {
        // filter-radius adjoint of the splat (diffvg.cpp:1250-1268)
        float radius = S.filter_radius;
        int ri = (int)ceil((double)radius);
        for (int dy = -ri; dy <= ri; dy++)
            for (int dx = -ri; dx <= ri; dx++) {
                int xx = x + dx, yy = y + dy;
                if (xx >= 0 && xx < G.W && yy >= 0 && yy < G.H) {
                    float ws = weight[yy * G.W + xx];
                    if (ws > 0) {
                        float xc = xx + 0.5f, yc = yy + 0.5f;
                        float fw = filter_weight(S.filter_type, radius, xc - rpt.x, yc - rpt.y);
                        float4 dp = reinterpret_cast<const float4 *>(d_img)[yy * G.W + xx];
                        float dpc = dot(V4{dp.x, dp.y, dp.z, dp.w}, color);
                        float d_weight = (dpc * ws - fw * dpc * (ws - fw)) / (ws * ws);
                        d_radius += d_filter_weight(S.filter_type, radius, xc - rpt.x, yc - rpt.y, d_weight);
                    }
                }
            }
    }
    // workgroup reduction of the filter-radius adjoint, one atomic per tile
    float v = d_radius;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    if (d_bg && has_bg) {
        s_dbg[threadIdx.x] = make_float4(dbg_val.x, dbg_val.y, dbg_val.z, dbg_val.w);
        s_dbg_kind[threadIdx.x] = kind;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < kBlock / 64; w++) t += s_red[w];
        if (t != 0.f) aadd(S.d_filter_radius, t);
    }
    // d_background: a sample without fragments assigns, one with fragments adds
    // (diffvg.cpp:650-656, 700-704); folded per pixel in sample order.
    if (d_bg && has_bg && threadIdx.x < G.tw * G.th && G.spp * G.tw * G.th <= kBlock) {
        int p = threadIdx.x;
        int px = tx * G.tw + p % G.tw, py = ty * G.th + p / G.tw;
        if (px < G.W && py < G.H) {
            float4 *dst = reinterpret_cast<float4 *>(d_bg) + (py * G.W + px);
            float4 acc = *dst;
            for (int k = 0; k < G.spp; k++) {
                int si = p * G.spp + k;
                float4 v4 = s_dbg[si];
                if (s_dbg_kind[si] == 0)
                    acc = v4;
                else if (s_dbg_kind[si] == 1)
                    acc = make_float4(acc.x + v4.x, acc.y + v4.y, acc.z + v4.z, acc.w + v4.w);
            }
            *dst = acc;
        }
    }
    if (overflow) atomicOr(err, overflow);
}
template <int PRE>
__global__ __launch_bounds__(kBlock, PRE ? DVG_WAVES_COLOR_BWD_PRE : DVG_WAVES_COLOR_BWD) void k_color_bwd(DevScene S, Geom G, const float *bg, const float *weight,
                                                      const float *d_img, float *d_bg, float *d_translation,
                                                      int *err, const uint16_t *frag, const int *pkeep) {
    __shared__ ColorBwdLds L;
    color_bwd_tile<PRE>(L, S, G, blockIdx.x, bg, weight, d_img, d_bg, d_translation, err, frag, pkeep);
}

// One boundary sample and its tile; the count the atomic returns is the sample's position
// in the tile's bin, so binning (edge_bin_one) needs no second round of atomics.
DVG_DEV void edge_sample_one(const DevScene &S, const Geom &G, int idx, int count, EdgeSample *edges,
                             int *edge_tile, int *edge_rank, int *tile_count) {
    if (idx >= count) return;
    EdgeSample es{};
    int tile = -1, rank = 0;
    if (S.num_total_shapes > 0 && make_edge_sample(S, idx, G.seed, es)) {
        int bx = (int)(es.px * G.W), by = (int)(es.py * G.H);
        if (bx >= 0 && bx < G.W && by >= 0 && by < G.H) {
            tile = (by / G.th) * G.tiles_x + (bx / G.tw);
            rank = atomicAdd(tile_count + tile, 1);
        }
    }
    edges[idx] = es;
    edge_tile[idx] = tile;
    edge_rank[idx] = rank;
}
__global__ __launch_bounds__(kBlock) void k_edge_sample(DevScene S, Geom G, int count, EdgeSample *edges,
                                                        int *edge_tile, int *edge_rank, int *tile_count) {
    edge_sample_one(S, G, blockIdx.x * blockDim.x + threadIdx.x, count, edges, edge_tile, edge_rank, tile_count);
}

// Exclusive scans, in one workgroup, of the per-tile boundary-sample counts (-> tile
// offsets) and of the per-tile chunk counts ceil(count / echunk) (-> chunk offsets; entry
// n holds the total), so that k_edge can run one workgroup per chunk.
template <int NT>
DVG_DEV void scan_tiles(ScanLds<NT> &L, const int *in, int *out, int *chunk_off, int n, int echunk) {
    int *s = L.s;
    int *sc = L.sc;
    if (threadIdx.x == 0) L.carry = L.carry_c = 0;
    __syncthreads();
    for (int base = 0; base < n; base += NT) {
        int i = base + threadIdx.x;
        int v = i < n ? in[i] : 0;
        int vc = (v + echunk - 1) / echunk;
        s[threadIdx.x] = v;
        sc[threadIdx.x] = vc;
        __syncthreads();
        for (int o = 1; o < NT; o <<= 1) {
            int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            int tc = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            sc[threadIdx.x] += tc;
            __syncthreads();
        }
        if (i < n) {
            out[i] = L.carry + s[threadIdx.x] - v;
            chunk_off[i] = L.carry_c + sc[threadIdx.x] - vc;
        }
        __syncthreads();
        if (threadIdx.x == NT - 1) {
            L.carry += s[NT - 1];
            L.carry_c += sc[NT - 1];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_off[n] = L.carry_c;
}
__global__ __launch_bounds__(1024) void k_scan(const int *in, int *out, int *chunk_off, int n, int echunk) {
    __shared__ ScanLds<1024> L;
    scan_tiles<1024>(L, in, out, chunk_off, n, echunk);
}

DVG_DEV void edge_bin_one(int idx, int count, const int *edge_tile, const int *edge_rank, const int *tile_off,
                          int *sorted) {
    if (idx >= count) return;
    int t = edge_tile[idx];
    if (t < 0) return;
    sorted[tile_off[t] + edge_rank[idx]] = idx;
}
__global__ __launch_bounds__(kBlock) void k_edge_bin(int count, const int *edge_tile, const int *edge_rank,
                                                     const int *tile_off, int *sorted) {
    edge_bin_one(blockIdx.x * blockDim.x + threadIdx.x, count, edge_tile, edge_rank, tile_off, sorted);
}

// One workgroup per chunk of up to G.echunk boundary samples of one tile (the grid is an
// upper bound on the chunk count; workgroups past the total exit at once), so a tile on a
// busy boundary is spread over many workgroups instead of looping inside one.  BOTH
// selects how lanes share the two colour queries of a sample (render_edge_kernel,
// diffvg.cpp:1388-1475, evaluates the inside and the outside point): false -- lanes 2i and
// 2i + 1 shade the inside and the outside point of sample i (kEdgeChunk samples, less
// state per lane); true -- one lane shades both points of its sample (kBlock samples).
template <bool BOTH>
DVG_DEV void edge_chunk(EdgeLds &L, const DevScene &S, const Geom &G, int chunk, const EdgeSample *edges,
                        const int *tile_off, const int *tile_count, const int *chunk_off, int ntiles,
                        const int *sorted, const float *bg, const float *weight, const float *d_img,
                        float *d_translation) {
    constexpr int kChunk = BOTH ? kBlock : kEdgeChunk;
    int *s_list = L.c.list;
    int *s_w = L.c.w;
    int *s_perm = L.perm;
    if (chunk >= chunk_off[ntiles]) return;
    // tile = the last t with chunk_off[t] <= chunk (chunk_off is non-decreasing)
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (chunk_off[mid] <= chunk)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int tile = lo;
    const int cnt = tile_count[tile];
    int tx = tile % G.tiles_x, ty = tile / G.tiles_x;
    float ext = 1e-4f * (float)max(S.cw, S.ch) * 1.01f;
    float4 tb = tile_box(S, G, tx, ty, ext);
    int off = tile_off[tile];
    bool has_bg = bg != nullptr;
    const int b0 = (chunk - chunk_off[tile]) * kChunk;
    {
        // The chunk's samples in order of their shape (stable): samples of one boundary
        // share a wave, so fewer waves run the deep part of the winding test for a curve
        // close to one of their points, and the gradient scatter combines more lanes.
        int *s_key = L.key;
        const int nvalid = min(kChunk, cnt - b0);
        if (threadIdx.x < kChunk)
            s_key[threadIdx.x] = (int)threadIdx.x < nvalid ? edges[sorted[off + b0 + threadIdx.x]].shape : INT_MAX;
        __syncthreads();
        int rank = 0;
        if (threadIdx.x < kChunk) {
            const int my = s_key[threadIdx.x];
            for (int j = 0; j < kChunk; j++) {
                const int o = s_key[j];
                rank += (o < my || (o == my && j < (int)threadIdx.x)) ? 1 : 0;
            }
        }
        __syncthreads();
        if (threadIdx.x < kChunk) s_perm[rank] = threadIdx.x;
        __syncthreads();
    }
    {
        const int side = BOTH ? 0 : (threadIdx.x & 1);
        int k = b0 + s_perm[BOTH ? threadIdx.x : (threadIdx.x >> 1)];
        bool valid = k < cnt;
        const int eidx = valid ? sorted[off + k] : 0;
        int tgroup = -1, tshape = -1;
        V4 cin{0, 0, 0, 0}, cout{0, 0, 0, 0};
        bool in_hit = false, out_hit = false;
        EdgeSample es{};
        if constexpr (!BOTH) {
            V2 p{0, 0};
            Compositor comp;
            {
                EdgeSample es0{};
                if (valid) es0 = edges[eidx];
                V2 bp{es0.px, es0.py}, n0{es0.nx, es0.ny};
                V2 q = side ? bp + 1e-4f * n0 : bp - 1e-4f * n0;
                p = V2{q.x * S.cw, q.y * S.ch};
                if (valid) {
                    tgroup = es0.group;
                    tshape = es0.shape;
                }
                int bx0 = (int)(es0.px * G.W), by0 = (int)(es0.py * G.H);
                V4 bg0 = (valid && has_bg) ? load_bg(bg, by0 * G.W + bx0) : V4{0, 0, 0, 0};
                comp.init(has_bg, bg0);
            }
            EdgeQuery q{tgroup, tshape, false};
            bool hit = false;
            for (int c0 = 0; c0 < S.ng; c0 += kCandCap) {
                int nc = build_candidates(S, tb, c0, min(S.ng, c0 + kCandCap), s_list, s_w);
                if (valid) shade_list(S, p, s_list, nc, comp, &q, hit);
                if (c0 + kCandCap < S.ng) __syncthreads();
            }
            asm volatile("" ::: "memory");  // read the sample again, do not keep it live
            if (valid) es = edges[eidx];
            int bx = (int)(es.px * G.W), by = (int)(es.py * G.H);
            V4 bgc = (valid && has_bg) ? load_bg(bg, by * G.W + bx) : V4{0, 0, 0, 0};
            // this side's colour and edge hit; the partner lane holds the other side
            const bool my_hit = valid && comp.edge_hit(hit);
            const V4 mine = valid ? comp.result(has_bg, bgc) : V4{0, 0, 0, 0};
            const bool other_hit = __shfl_xor((int)my_hit, 1) != 0;
            const V4 other{__shfl_xor(mine.x, 1), __shfl_xor(mine.y, 1), __shfl_xor(mine.z, 1),
                           __shfl_xor(mine.w, 1)};
            in_hit = my_hit;
            out_hit = other_hit;
            cin = mine;
            cout = other;
        } else {
            V2 pin{0, 0}, pout{0, 0};
            Compositor ci, co;
            {
                EdgeSample es0{};
                if (valid) es0 = edges[eidx];
                V2 bp{es0.px, es0.py}, n0{es0.nx, es0.ny};
                V2 qi = bp - 1e-4f * n0, qo = bp + 1e-4f * n0;
                pin = V2{qi.x * S.cw, qi.y * S.ch};
                pout = V2{qo.x * S.cw, qo.y * S.ch};
                if (valid) {
                    tgroup = es0.group;
                    tshape = es0.shape;
                }
                int bx0 = (int)(es0.px * G.W), by0 = (int)(es0.py * G.H);
                V4 bg0 = (valid && has_bg) ? load_bg(bg, by0 * G.W + bx0) : V4{0, 0, 0, 0};
                ci.init(has_bg, bg0);
                co.init(has_bg, bg0);
            }
            EdgeQuery qin{tgroup, tshape, false};
            EdgeQuery qout = qin;
            bool hin = false, hout = false;
            for (int c0 = 0; c0 < S.ng; c0 += kCandCap) {
                int nc = build_candidates(S, tb, c0, min(S.ng, c0 + kCandCap), s_list, s_w);
                if (valid) {
                    shade_list(S, pin, s_list, nc, ci, &qin, hin);
                    shade_list(S, pout, s_list, nc, co, &qout, hout);
                }
                if (c0 + kCandCap < S.ng) __syncthreads();
            }
            asm volatile("" ::: "memory");  // read the sample again, do not keep it live
            if (valid) es = edges[eidx];
            int bx = (int)(es.px * G.W), by = (int)(es.py * G.H);
            V4 bgc = (valid && has_bg) ? load_bg(bg, by * G.W + bx) : V4{0, 0, 0, 0};
            in_hit = valid && ci.edge_hit(hin);
            out_hit = valid && co.edge_hit(hout);
            if (valid) {
                cin = ci.result(has_bg, bgc);
                cout = co.result(has_bg, bgc);
            }
        }
        V2 n{es.nx, es.ny};
        int bx = (int)(es.px * G.W), by = (int)(es.py * G.H);
        // every lane stays to the end: the gradient scatter combines lanes of a wave
        bool active = valid && side == 0 && (in_hit || out_hit);
        float contrib = 0.f;
        if (active) {
            if (!in_hit) {
                n = -n;
                V4 tmp = cin;
                cin = cout;
                cout = tmp;
            }
            V2 sb{es.px * G.W, es.py * G.H};
            V4 d_color = gather_d_color(S.filter_type, S.filter_radius, d_img, weight, G.W, G.H, sb);
            d_color = vdiv(d_color, (float)(S.cw * S.ch));
            contrib = dot(cin - cout, d_color) / es.pdf;
        }
        BoundaryData bd = active ? es.data : BoundaryData{0, 0, 0.f, 0};
        accumulate_boundary_gradient_wave(S, active, active ? es.group : 0, active ? es.shape : 0, contrib,
                                          active ? es.t : 0.f, n, bd, active ? V2{es.lx, es.ly} : V2{0, 0});
        if (active && d_translation) {
            aadd(d_translation + 2 * (by * G.W + bx) + 0, n.x * contrib);
            aadd(d_translation + 2 * (by * G.W + bx) + 1, n.y * contrib);
        }
    }
}
template <bool BOTH>
__global__ __launch_bounds__(kBlock, BOTH ? DVG_WAVES_EDGE_BOTH : DVG_WAVES_EDGE) void k_edge(
    DevScene S, Geom G, const EdgeSample *edges, const int *tile_off, const int *tile_count, const int *chunk_off,
    int ntiles, const int *sorted, const float *bg, const float *weight, const float *d_img, float *d_translation) {
    __shared__ EdgeLds L;
    edge_chunk<BOTH>(L, S, G, blockIdx.x, edges, tile_off, tile_count, chunk_off, ntiles, sorted, bg, weight, d_img,
                     d_translation);
}

// SDF (sample_distance, diffvg.cpp:709-775) -- a linear scan over groups, as the
// reference does; one thread per sample.
DVG_DEV void sdf_sample(const DevScene &S, const Geom &G, int idx, const float *eval_positions, int num_eval,
                        float *sdf, const float *d_sdf, float *d_translation) {
    int count = eval_positions ? num_eval : G.W * G.H * G.spp;
    if (idx >= count) return;
    V2 rpt;
    int x, y;
    if (eval_positions) {
        rpt = V2{eval_positions[2 * idx], eval_positions[2 * idx + 1]};
        x = (int)rpt.x;
        y = (int)rpt.y;
    } else {
        int sx = idx % G.nsx;
        int sy = (idx / G.nsx) % G.nsy;
        x = (idx / (G.nsx * G.nsy)) % G.W;
        y = idx / (G.nsx * G.nsy * G.W);
        rpt = sample_pos(G, x, y, sx, sy, idx, false);
    }
    V2 pt = to_canvas(S, G, rpt);
    float weight = eval_positions ? 1.f : 1.f / G.spp;
    const float inf = __builtin_huge_valf();
    int min_group = -1, min_shape = -1;
    float min_distance = 0.f;
    V2 closest{0, 0};
    PathInfo min_info{-1, -1, 0.f};
    for (int g = S.ng - 1; g >= 0; g--) {
        int s = -1;
        V2 p{0, 0};
        PathInfo li{-1, -1, 0.f};
        float d = inf;
        if (compute_distance(S, g, pt, inf, s, p, li, d)) {
            if (min_group == -1 || d < min_distance) {
                min_distance = d;
                min_group = g;
                min_shape = s;
                closest = p;
                min_info = li;
            }
        }
    }
    float dist_out = min_distance;
    if (min_group >= 0) {
        dist_out = min_distance * weight;
        bool in = false;
        if (S.fill_type[min_group] >= 0) {
            M3 c2s = m3_load(S.c2s + 9 * min_group);
            in = is_inside_local(S, min_group, xform_pt(c2s, pt), nullptr);
            if (in) dist_out = -dist_out;
        }
        if (d_sdf) {
            float dd = eval_positions ? d_sdf[idx] : d_sdf[y * G.W + x];
            float d_abs = in ? -dd : dd;
            // the reference indexes d_translation by the sample pixel even in eval mode;
            // positions outside the image have no pixel and are skipped here
            bool in_img = x >= 0 && x < G.W && y >= 0 && y < G.H;
            float *d_tr = (d_translation && in_img) ? d_translation + 2 * (y * G.W + x) : nullptr;
            d_compute_distance(S, min_group, min_shape, pt, closest, min_info, d_abs, d_tr);
        }
    }
    if (sdf) aadd(eval_positions ? sdf + idx : sdf + (y * G.W + x), dist_out);
}
// The SDF of one pixel tile (sample_distance, diffvg.cpp:709-775), with the groups that
// cannot be nearest culled per tile.  For a group g, lb_g = distance from the tile box to
// g's canvas box bounds its computed distance from below at every sample of the tile;
// for an anchored group (dvg_scene.h, g_sdf), ub_g = the largest distance from a tile
// corner to its anchor bounds it from above.  With U = min_g ub_g, a group with lb_g > U
// is strictly farther than the group attaining U at every sample, so it can neither win
// nor tie; the rest are visited in the reference's order (descending group id, first
// strict minimum kept), which gives the reference's result.  Both bounds carry a margin
// of 1e-3 + 1e-5 * |coordinate| for rounding.
DVG_DEV float box_dist(float4 b, float4 t) {
    float dx = fmaxf(0.f, fmaxf(b.x - t.z, t.x - b.z));
    float dy = fmaxf(0.f, fmaxf(b.y - t.w, t.y - b.w));
    return sqrtf(dx * dx + dy * dy);
}

DVG_DEV void sdf_tile(SdfLds &L, const DevScene &S, const Geom &G, int tile, float *sdf, const float *d_sdf,
                      float *d_translation) {
    int *s_list = L.c.list;
    int *s_w = L.c.w;
    float *s_red = L.red;
    const int tx = tile % G.tiles_x, ty = tile / G.tiles_x;
    int x, y, sx, sy, idx;
    bool valid = tile_sample(G, tx, ty, threadIdx.x, x, y, sx, sy, idx);
    float4 tb = tile_box(S, G, tx, ty, 0.f);
    const float margin = 1e-3f + 1e-5f * fmaxf(fmaxf(fabsf(tb.x), fabsf(tb.z)), fmaxf(fabsf(tb.y), fabsf(tb.w)));
    // U = min over anchored groups of the farthest tile corner's distance to the anchor
    float u = __builtin_huge_valf();
    for (int g = threadIdx.x; g < S.ng; g += blockDim.x) {
        float4 a = S.g_sdf[g];
        if (a.z > 0.f && a.x == a.x) {
            float ex = fmaxf(fabsf(tb.x - a.x), fabsf(tb.z - a.x)), ey = fmaxf(fabsf(tb.y - a.y), fabsf(tb.w - a.y));
            u = fminf(u, sqrtf(ex * ex + ey * ey) + margin);
        }
    }
    for (int o = 32; o > 0; o >>= 1) u = fminf(u, __shfl_down(u, o, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = u;
    __syncthreads();
    float U = s_red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) U = fminf(U, s_red[w]);
    V2 pt{0, 0};
    if (valid) pt = to_canvas(S, G, sample_pos(G, x, y, sx, sy, idx, false));
    const float inf = __builtin_huge_valf();
    int min_group = -1, min_shape = -1;
    float min_distance = 0.f;
    V2 closest{0, 0};
    PathInfo min_info{-1, -1, 0.f};
    // candidate ranges from the top, so that the visit order stays descending
    for (int c1 = S.ng; c1 > 0; c1 -= kCandCap) {
        int c0 = max(0, c1 - kCandCap);
        int total = 0;
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int base = c0; base < c1; base += blockDim.x) {
            int g = base + threadIdx.x;
            bool hit = false;
            if (g < c1) {
                float4 a = S.g_sdf[g];
                hit = a.z > 0.f && box_dist(S.g_box[g], tb) - margin <= U;
            }
            uint64_t m = __ballot(hit);
            int pre = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) s_w[wid] = __popcll(m);
            __syncthreads();
            int off = total, chunk = 0;
            for (int w = 0; w < nw; w++) {
                if (w < wid) off += s_w[w];
                chunk += s_w[w];
            }
            if (hit) s_list[off + pre] = g;
            total += chunk;
            __syncthreads();
        }
        if (valid) {
            for (int i = total - 1; i >= 0; i--) {
                int g = s_list[i];
                int sh = -1;
                V2 p{0, 0};
                PathInfo li{-1, -1, 0.f};
                float d = inf;
                if (compute_distance(S, g, pt, inf, sh, p, li, d)) {
                    if (min_group == -1 || d < min_distance) {
                        min_distance = d;
                        min_group = g;
                        min_shape = sh;
                        closest = p;
                        min_info = li;
                    }
                }
            }
        }
        if (c1 - kCandCap > 0) __syncthreads();
    }
    float weight = 1.f / G.spp;
    float dist_out = min_distance;
    const bool found = valid && min_group >= 0;
    bool in = false;
    if (found) {
        dist_out = min_distance * weight;
        if (S.fill_type[min_group] >= 0) {
            M3 c2s = m3_load(S.c2s + 9 * min_group);
            in = is_inside_local(S, min_group, xform_pt(c2s, pt), nullptr);
            if (in) dist_out = -dist_out;
        }
    }
    if (d_sdf) {
        // every lane: the distance adjoints of the tile's samples are combined per wave
        float dd = found ? d_sdf[y * G.W + x] : 0.f;
        float d_abs = in ? -dd : dd;
        float *d_tr = (found && d_translation) ? d_translation + 2 * (y * G.W + x) : nullptr;
        d_compute_distance_wave(S, found, found ? min_group : 0, found ? min_shape : 0, pt, closest, min_info, d_abs,
                                d_tr);
    }
    if (valid && sdf) aadd(sdf + (y * G.W + x), dist_out);
}
__global__ __launch_bounds__(kBlock, DVG_WAVES_SDF) void k_sdf_tiled(DevScene S, Geom G, float *sdf, const float *d_sdf,
                                                      float *d_translation) {
    __shared__ SdfLds L;
    sdf_tile(L, S, G, blockIdx.x, sdf, d_sdf, d_translation);
}
__global__ __launch_bounds__(kBlock) void k_sdf(DevScene S, Geom G, const float *eval_positions, int num_eval,
                                                float *sdf, const float *d_sdf, float *d_translation) {
    sdf_sample(S, G, blockIdx.x * blockDim.x + threadIdx.x, eval_positions, num_eval, sdf, d_sdf, d_translation);
}

// ------------------------------------------------------------------ PLAIN kernels
// One thread per sample in 64-thread blocks.  Each sample walks the top-level BVH
// (child1 first, stack of 64, diffvg.cpp:548-595), collects the reached groups and sorts
// them by insertion, then shades them with the same per-group code as the tiled path.
constexpr int kPlainCand = 1024;

DVG_DEV int plain_candidates(const DevScene &S, V2 pt, int *list, int &overflow) {
    int n = 0;
    if (S.ng == 0) return 0;
    int stack[64];
    int sp = 0;
    stack[sp++] = 2 * S.ng - 2;
    while (sp > 0) {
        const BNode node = S.top[stack[--sp]];
        if (node.child1 < 0) {
            if (n < kPlainCand)
                list[n++] = node.child0;
            else
                overflow = 1;
        } else {
            const BNode c0 = S.top[node.child0];
            const BNode c1 = S.top[node.child1];
            if (in_box(make_float4(c0.minx, c0.miny, c0.maxx, c0.maxy), pt, c0.max_radius) && sp < 64)
                stack[sp++] = node.child0;
            if (in_box(make_float4(c1.minx, c1.miny, c1.maxx, c1.maxy), pt, c1.max_radius) && sp < 64)
                stack[sp++] = node.child1;
        }
    }
    for (int i = 1; i < n; i++) {
        int j = i, v = list[i];
        while (j > 0 && list[j - 1] > v) {
            list[j] = list[j - 1];
            j--;
        }
        list[j] = v;
    }
    return n;
}

__global__ __launch_bounds__(64) void kp_weight(DevScene S, Geom G, float *weight) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= G.W * G.H * G.spp) return;
    int sx = idx % G.nsx, sy = (idx / G.nsx) % G.nsy;
    int x = (idx / (G.nsx * G.nsy)) % G.W, y = idx / (G.nsx * G.nsy * G.W);
    V2 pt = sample_pos(G, x, y, sx, sy, idx, false);
    float radius = S.filter_radius;
    int ri = (int)ceil((double)radius);
    for (int dy = -ri; dy <= ri; dy++)
        for (int dx = -ri; dx <= ri; dx++) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && xx < G.W && yy >= 0 && yy < G.H)
                aadd(weight + yy * G.W + xx, filter_weight(S.filter_type, radius, xx + 0.5f - pt.x, yy + 0.5f - pt.y));
        }
}

template <int PRE>
__global__ __launch_bounds__(64) void kp_render(DevScene S, Geom G, const float *bg, float *img, const float *weight,
                                                const float *d_img, float4 *dbg_rec, int *dbg_kind,
                                                float *d_translation, int *err) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= G.W * G.H * G.spp) return;
    int sx = idx % G.nsx, sy = (idx / G.nsx) % G.nsy;
    int x = (idx / (G.nsx * G.nsy)) % G.W, y = idx / (G.nsx * G.nsy * G.W);
    V2 rpt = sample_pos(G, x, y, sx, sy, idx, PRE != 0);
    V2 pt = to_canvas(S, G, rpt);
    bool has_bg = bg != nullptr;
    V4 bgc = has_bg ? load_bg(bg, y * G.W + x) : V4{0, 0, 0, 0};
    float *d_tr = d_translation ? d_translation + 2 * (y * G.W + x) : nullptr;
    int overflow = 0;
    int list[kPlainCand];
    int n = plain_candidates(S, pt, list, overflow);
    Compositor comp;
    comp.init(has_bg, bgc);
    FragRec frags[PRE ? 1 : kMaxFrags];
    PFragRec pfrags[PRE ? kMaxPFrags : 1];
    if constexpr (PRE) {
        if (d_img)
            shade_list_prefiltered<true>(S, pt, list, n, comp, pfrags, overflow);
        else
            shade_list_prefiltered<false>(S, pt, list, n, comp, pfrags, overflow);
    } else {
        if (d_img) {
            shade_list_record(S, pt, list, n, comp, frags, overflow);
        } else {
            bool th = false;
            shade_list(S, pt, list, n, comp, nullptr, th);
        }
    }
    V4 color = comp.result(has_bg, bgc);
    if (d_img) {
        V4 d_color = gather_d_color(S.filter_type, S.filter_radius, d_img, weight, G.W, G.H, rpt);
        // d_background: recorded per sample and folded per pixel in sample order by
        // kp_fold_dbg (a sample without fragments assigns, one with fragments adds)
        if (comp.nf == 0) {
            if (dbg_rec) {
                dbg_rec[idx] = make_float4(d_color.x, d_color.y, d_color.z, d_color.w);
                dbg_kind[idx] = 0;
            }
        } else {
            V4 d_back;
            V3 fc{color.x, color.y, color.z};
            if constexpr (PRE)
                d_back = composite_adjoint_prefiltered(S, pfrags, min(comp.nf, kMaxPFrags), fc, color.w, d_color, pt,
                                                       d_tr);
            else
                d_back = composite_adjoint(S, frags, min(comp.nf, kMaxFrags), fc, color.w, d_color, pt, d_tr);
            if (dbg_rec) {
                dbg_rec[idx] = make_float4(d_back.x, d_back.y, d_back.z, d_back.w);
                dbg_kind[idx] = 1;
            }
        }
    }
    float radius = S.filter_radius;
    int ri = (int)ceil((double)radius);
    for (int dy = -ri; dy <= ri; dy++)
        for (int dx = -ri; dx <= ri; dx++) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && xx < G.W && yy >= 0 && yy < G.H && weight[yy * G.W + xx] > 0) {
                float ws = weight[yy * G.W + xx];
                float xc = xx + 0.5f, yc = yy + 0.5f;
                float fw = filter_weight(S.filter_type, radius, xc - rpt.x, yc - rpt.y);
                if (img) aadd4(img + 4 * (yy * G.W + xx), vdiv(fw * color, ws));
                if (d_img) {
                    float4 dp = reinterpret_cast<const float4 *>(d_img)[yy * G.W + xx];
                    float dpc = dot(V4{dp.x, dp.y, dp.z, dp.w}, color);
                    float d_weight = (dpc * ws - fw * dpc * (ws - fw)) / (ws * ws);
                    aadd(S.d_filter_radius, d_filter_weight(S.filter_type, radius, xc - rpt.x, yc - rpt.y, d_weight));
                }
            }
        }
    if (overflow) atomicOr(err, overflow);
}

// Per-pixel fold of the d_background records of its samples, in sample order
// (diffvg.cpp:650-656, 700-704: assignment without fragments, addition with them).
__global__ __launch_bounds__(64) void kp_fold_dbg(Geom G, const float4 *dbg_rec, const int *dbg_kind, float *d_bg) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.W * G.H) return;
    float4 *dst = reinterpret_cast<float4 *>(d_bg) + p;
    float4 acc = *dst;
    for (int k = 0; k < G.spp; k++) {
        size_t si = (size_t)p * G.spp + k;
        float4 v = dbg_rec[si];
        if (dbg_kind[si] == 0)
            acc = v;
        else
            acc = make_float4(acc.x + v.x, acc.y + v.y, acc.z + v.z, acc.w + v.w);
    }
    *dst = acc;
}

__global__ __launch_bounds__(64) void kp_edge_sample(DevScene S, Geom G, int count, EdgeSample *edges, int *ids,
                                                     uint32_t *keys) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    EdgeSample es;
    uint32_t key = 0;
    if (make_edge_sample(S, idx, G.seed, es)) {
        uint32_t xi = f2u32_x86(es.px * 1023), yi = f2u32_x86(es.py * 1023);
        key = (expand_bits(xi) << 1u) | expand_bits(yi);
    }
    edges[idx] = es;
    ids[idx] = idx;
    keys[idx] = key;
}

__global__ __launch_bounds__(64) void kp_edge(DevScene S, Geom G, const EdgeSample *edges, const int *ids,
                                              const float *bg, const float *weight, const float *d_img,
                                              float *d_translation, int *err) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G.W * G.H * G.spp) return;
    EdgeSample es = edges[ids[k]];
    if (es.shape < 0) return;
    int bx = (int)(es.px * G.W), by = (int)(es.py * G.H);
    if (bx < 0 || bx >= G.W || by < 0 || by >= G.H) return;
    bool has_bg = bg != nullptr;
    V4 bgc = has_bg ? load_bg(bg, by * G.W + bx) : V4{0, 0, 0, 0};
    V2 bp{es.px, es.py}, n{es.nx, es.ny};
    V2 qi = bp - 1e-4f * n, qo = bp + 1e-4f * n;
    V2 pin{qi.x * S.cw, qi.y * S.ch}, pout{qo.x * S.cw, qo.y * S.ch};
    int overflow = 0;
    int list[kPlainCand];
    EdgeQuery qin{es.group, es.shape, false}, qout{es.group, es.shape, false};
    bool hin = false, hout = false;
    Compositor ci, co;
    ci.init(has_bg, bgc);
    co.init(has_bg, bgc);
    int nc = plain_candidates(S, pin, list, overflow);
    shade_list(S, pin, list, nc, ci, &qin, hin);
    nc = plain_candidates(S, pout, list, overflow);
    shade_list(S, pout, list, nc, co, &qout, hout);
    if (overflow) atomicOr(err, overflow);
    bool in_hit = ci.edge_hit(hin), out_hit = co.edge_hit(hout);
    if (!in_hit && !out_hit) return;
    V4 cin = ci.result(has_bg, bgc), cout = co.result(has_bg, bgc);
    if (!in_hit) {
        n = -n;
        V4 tmp = cin;
        cin = cout;
        cout = tmp;
    }
    V4 d_color = gather_d_color(S.filter_type, S.filter_radius, d_img, weight, G.W, G.H, V2{es.px * G.W, es.py * G.H});
    d_color = vdiv(d_color, (float)(S.cw * S.ch));
    float contrib = dot(cin - cout, d_color) / es.pdf;
    accumulate_boundary_gradient(S, es.group, es.shape, contrib, es.t, n, es.data, V2{es.lx, es.ly});
    if (d_translation) {
        aadd(d_translation + 2 * (by * G.W + bx) + 0, n.x * contrib);
        aadd(d_translation + 2 * (by * G.W + bx) + 1, n.y * contrib);
    }
}


// ------------------------------------------------------------------ kernel profiler
// When enabled, every launch is bracketed by HIP events recorded on the stream it runs
// on; dvg_profile_read() waits for them and reports the summed durations per kernel.
namespace prof {
struct Rec {
    const char *name;
    hipEvent_t a, b;
};
static bool enabled = false;
static std::vector<Rec> recs;
static std::mutex mu;
struct Scope {
    hipStream_t st;
    Rec r{};
    bool on;
    Scope(hipStream_t s, const char *name) : st(s), on(enabled) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (on && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) on = false;
        if (!on) return;
        r.name = name;
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, st);
    }
    ~Scope() {
        if (!on) return;
        (void)hipEventRecord(r.b, st);
        std::lock_guard<std::mutex> lk(mu);
        recs.push_back(r);
    }
};
}  // namespace prof
#define DVG_PROF(st, name) prof::Scope prof_scope_##__LINE__(st, name)

// ------------------------------------------------------------------ launch sequences
int Workspace::reserve(size_t npix, size_t nsamp, size_t ntiles) {
    if (npix <= npix_cap && nsamp <= nsamp_cap && ntiles <= ntiles_cap) return 0;
    // growing: earlier renders may still use the old buffers on any stream
    if (weight) (void)hipDeviceSynchronize();
    release();
    auto al = [](void **p, size_t bytes) { return hipMalloc(p, bytes < 256 ? 256 : bytes); };
    if (al((void **)&weight, npix * sizeof(float)) != hipSuccess ||
        al((void **)&colors, nsamp * sizeof(float4)) != hipSuccess ||
        al((void **)&edges, nsamp * sizeof(EdgeSample)) != hipSuccess ||
        al((void **)&edge_tile, nsamp * sizeof(int)) != hipSuccess ||
        al((void **)&edge_sorted, nsamp * sizeof(int)) != hipSuccess ||
        al((void **)&tile_count, ntiles * sizeof(int)) != hipSuccess ||
        al((void **)&tile_off, ntiles * sizeof(int)) != hipSuccess ||
        al((void **)&edge_rank, nsamp * sizeof(int)) != hipSuccess ||
        al((void **)&chunk_off, (ntiles + 1) * sizeof(int)) != hipSuccess) {
        release();
        return set_error("out of device memory for the render workspace");
    }
    npix_cap = npix;
    nsamp_cap = nsamp;
    ntiles_cap = ntiles;
    return 0;
}

namespace {
struct WsSlot {
    std::mutex mu;
    Workspace ws;
    hipEvent_t last = nullptr;
    bool used = false;
};
WsSlot &ws_slot(int device) {
    static std::mutex m;
    static std::map<int, WsSlot *> slots;
    std::lock_guard<std::mutex> lk(m);
    WsSlot *&s = slots[device];
    if (!s) s = new WsSlot();
    return *s;
}
}  // namespace

Workspace &acquire_workspace(int device, hipStream_t st) {
    WsSlot &s = ws_slot(device);
    s.mu.lock();
    if (s.used && s.last) (void)hipStreamWaitEvent(st, s.last, 0);
    return s.ws;
}

void release_workspace(int device, hipStream_t st) {
    WsSlot &s = ws_slot(device);
    if (!s.last) (void)hipEventCreateWithFlags(&s.last, hipEventDisableTiming);
    (void)hipEventRecord(s.last, st);
    s.used = true;
    s.mu.unlock();
}

void Workspace::release() {
    void *ps[] = {weight, colors, edges, edge_tile, edge_sorted, edge_rank, tile_count, tile_off, chunk_off};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    weight = nullptr;
    colors = nullptr;
    edges = nullptr;
    edge_tile = edge_sorted = edge_rank = tile_count = tile_off = chunk_off = nullptr;
    npix_cap = nsamp_cap = ntiles_cap = 0;
}

static Geom make_geom(const SceneHost &sc, const dvg_render_args &a) {
    Geom G;
    G.W = a.width;
    G.H = a.height;
    G.nsx = a.num_samples_x;
    G.nsy = a.num_samples_y;
    G.spp = G.nsx * G.nsy;
    G.seed = a.seed;
    int px = kBlock / G.spp;
    if (px < 1) px = 1;
    int tw = 1;
    while ((tw * 2) * (tw * 2) <= px) tw *= 2;
    // a tile width chosen by the caller (the autotuner) overrides the square default
    if (sc.tile_w > 0 && sc.tile_w <= px) tw = sc.tile_w;
    int th = px / tw;
    if (th < 1) th = 1;
    G.tw = tw;
    G.th = th;
    G.echunk = sc.edge_spw == 64 ? kBlock : kEdgeChunk;
    G.tiles_x = (G.W + tw - 1) / tw;
    G.tiles_y = (G.H + th - 1) / th;
    G.sx_scale = (float)sc.cw / (float)G.W;
    G.sy_scale = (float)sc.ch / (float)G.H;
    float mx = (float)std::max(std::max(sc.cw, sc.ch), std::max(G.W, G.H));
    G.margin = 1e-3f + 4e-6f * mx;
    return G;
}

#define DVG_CHECK(x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return set_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// conditions found by dvg_scene_update's kernels (err[1])
enum { kUpdNonFinitePoints = 1, kUpdNonFiniteLength = 2, kUpdTotalLength = 4 };

static int report_flags(const int *h) {
    if (h[1] & kUpdNonFinitePoints) return set_error("non-finite path points or thickness");
    if (h[1] & kUpdNonFiniteLength) return set_error("non-finite shape boundary length");
    if (h[1] & kUpdTotalLength)
        return set_error("The total length of the shape boundaries in the scene is equal or less than 0, or not a "
                         "number");
    if (h[0] & 2) return set_error("prefiltered stroke of a group without a path or rect shape (the reference asserts)");
    if (h[0] & 1) return set_error("fragment list overflow: more overlapping fragments than the reference supports");
    return 0;
}

// Reads the conditions found by the last update and by the renders since the previous
// read, and clears the latter: a condition a forward finds is still there when its
// backward has rendered on the same scene.
static int check_err_flag(SceneHost &sc, hipStream_t st) {
    int h[2] = {0, 0};
    DVG_CHECK(hipMemcpyAsync(h, sc.err, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    DVG_CHECK(hipMemsetAsync(sc.err, 0, sizeof(int), st));
    DVG_CHECK(hipStreamSynchronize(st));
    return report_flags(h);
}

int scene_status(SceneHost &sc) {
    if (!sc.device_set || !sc.err) return 0;
    return check_err_flag(sc, sc.last_stream);
}

// A scene without shapes (and so without groups): every sample sees only the
// background, which the box-normalised resolve returns unchanged (transparent black when
// there is none); the SDF adds 0 (diffvg.cpp:741-743) and no gradient is produced.
static int render_empty(const dvg_render_args &a, size_t npix, hipStream_t st) {
    if (a.render_image && a.background_image)
        DVG_CHECK(hipMemcpyAsync(a.render_image, a.background_image, npix * 16, hipMemcpyDeviceToDevice, st));
    else if (a.render_image)
        DVG_CHECK(hipMemsetAsync(a.render_image, 0, npix * 16, st));
    return 0;
}

// Makes `ptr` a kept-list buffer of at least `bytes` for scene sc: the current one when
// it is large enough, else a new one -- unless batches hold the current one, and then
// false (the render shades its backward samples).  `owner` is the stream that last used
// the buffer.
static bool keep_buffer(SceneHost &sc, void *&ptr, size_t &granted, hipStream_t owner, size_t bytes) {
    if (ptr && granted >= bytes) return true;
    if (sc.keep_refs > 0) return false;
    if (ptr) {
        if (hipStreamSynchronize(owner) != hipSuccess) return false;
        cache_free(sc.device, ptr, granted);
        ptr = nullptr;
    }
    ptr = cache_alloc(sc.device, bytes, &granted);
    return ptr != nullptr;
}

int render_tiled(SceneHost &sc, const dvg_render_args &a, Workspace *fixed) {
    hipStream_t st = (hipStream_t)a.stream;
    Geom G = make_geom(sc, a);
    const bool eval_mode = a.eval_positions && a.num_eval_positions > 0;
    const int pre = a.use_prefiltering ? 1 : 0;
    const size_t npix = (size_t)G.W * G.H;
    const size_t nsamp = npix * G.spp;
    DevScene S = sc.dev;
    if (eval_mode && (a.render_image || a.d_render_image))
        return set_error("eval_positions are supported for the SDF output only (as in the reference)");
    if (sc.num_shapes == 0) return render_empty(a, npix, st);
    const bool color_out = a.render_image || a.d_render_image;
    const bool sdf_out = a.render_sdf || a.d_render_sdf;
    int ntiles = G.tiles_x * G.tiles_y;
    if (color_out) {
        Workspace &ws = fixed ? *fixed : acquire_workspace(sc.device, st);
        struct Release {
            int d;
            hipStream_t s;
            bool on;
            ~Release() {
                if (on) release_workspace(d, s);
            }
        } release_guard{sc.device, st, fixed == nullptr};
        if (fixed) {
            if (npix > ws.npix_cap || nsamp > ws.nsamp_cap || (size_t)ntiles > ws.ntiles_cap)
                return set_error("render larger than the fixed workspace");
        } else if (int e = ws.reserve(npix, nsamp, ntiles)) {
            return e;
        }
        {
            DVG_PROF(st, "k_weight");
            k_weight<<<(unsigned)((npix + kBlock - 1) / kBlock), kBlock, 0, st>>>(S, G, ws.weight);
        }
        // The forward keeps each sample's fragment list (group ids in compositing order)
        // in the scene when it fits the budget; a backward of the same render (same seed,
        // sizes and parameter version, enqueued on the same stream) replays it.
        const uint64_t key[4] = {a.seed, (uint64_t)(uint32_t)G.W | ((uint64_t)(uint32_t)G.H << 32),
                                 (uint64_t)(uint32_t)G.nsx | ((uint64_t)(uint32_t)G.nsy << 32), sc.version};
        if (a.render_image) {
            uint16_t *frag = nullptr;
            int *pkeep = nullptr;
            sc.frag_valid = false;
            sc.pkeep_valid = false;
            const size_t frag_bytes = (size_t)(kFragSlots + 1) * nsamp * sizeof(uint16_t);
            const size_t pkeep_bytes = (size_t)(kPFragSlots * kPFragWords + 1) * nsamp * sizeof(int);
            // a backward on another stream may still read the lists being replaced; with
            // a fixed workspace nothing is kept (keeping may allocate)
            if (!fixed && !pre && frag_bytes <= kFragCacheBytes && S.ng < kFragMaxGroups) {
                if (sc.frag && sc.frag_stream != st) DVG_CHECK(hipStreamSynchronize(sc.frag_stream));
                void *b = sc.frag;
                const bool ok = keep_buffer(sc, b, sc.frag_granted, sc.frag_stream, frag_bytes);
                sc.frag = (uint16_t *)b;
                frag = ok ? sc.frag : nullptr;
            }
            if (!fixed && pre && pkeep_bytes <= kFragCacheBytes) {
                if (sc.pkeep && sc.pkeep_stream != st) DVG_CHECK(hipStreamSynchronize(sc.pkeep_stream));
                void *b = sc.pkeep;
                const bool ok = keep_buffer(sc, b, sc.pkeep_granted, sc.pkeep_stream, pkeep_bytes);
                sc.pkeep = (int *)b;
                pkeep = ok ? sc.pkeep : nullptr;
            }
            {
                DVG_PROF(st, "k_color");
                if (pre)
                    k_color<1><<<ntiles, kBlock, 0, st>>>(S, G, a.background_image, ws.colors, sc.err, nullptr, pkeep);
                else
                    k_color<0><<<ntiles, kBlock, 0, st>>>(S, G, a.background_image, ws.colors, sc.err, frag, nullptr);
            }
            if (frag) {
                sc.frag_valid = true;
                for (int k = 0; k < 4; k++) sc.frag_key[k] = key[k];
                sc.frag_stream = st;
            }
            if (pkeep) {
                sc.pkeep_valid = true;
                for (int k = 0; k < 4; k++) sc.pkeep_key[k] = key[k];
                sc.pkeep_stream = st;
            }
            {
                DVG_PROF(st, "k_resolve");
                k_resolve<<<(unsigned)((npix + kBlock - 1) / kBlock), kBlock, 0, st>>>(S, G, pre, ws.weight,
                                                                                      ws.colors, a.render_image);
            }
        }
        if (a.d_render_image) {
            {
                DVG_PROF(st, "k_color_bwd");
                bool kept = !pre && sc.frag_valid && sc.frag_stream == st;
                bool pkept = pre && sc.pkeep_valid && sc.pkeep_stream == st;
                for (int k = 0; k < 4; k++) {
                    kept = kept && sc.frag_key[k] == key[k];
                    pkept = pkept && sc.pkeep_key[k] == key[k];
                }
                if (pre)
                    k_color_bwd<1><<<ntiles, kBlock, 0, st>>>(S, G, a.background_image, ws.weight,
                                                              a.d_render_image, a.d_background_image,
                                                              a.d_translation, sc.err, nullptr,
                                                              pkept ? sc.pkeep : nullptr);
                else
                    k_color_bwd<0><<<ntiles, kBlock, 0, st>>>(S, G, a.background_image, ws.weight,
                                                              a.d_render_image, a.d_background_image,
                                                              a.d_translation, sc.err, kept ? sc.frag : nullptr,
                                                              nullptr);
            }
            if (!pre) {
                DVG_CHECK(hipMemsetAsync(ws.tile_count, 0, sizeof(int) * ntiles, st));
                {
                    DVG_PROF(st, "k_edge_sample");
                    k_edge_sample<<<(unsigned)((nsamp + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                        S, G, (int)nsamp, ws.edges, ws.edge_tile, ws.edge_rank, ws.tile_count);
                }
                {
                    DVG_PROF(st, "k_edge_bin");
                    k_scan<<<1, 1024, 0, st>>>(ws.tile_count, ws.tile_off, ws.chunk_off, ntiles, G.echunk);
                    k_edge_bin<<<(unsigned)((nsamp + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                        (int)nsamp, ws.edge_tile, ws.edge_rank, ws.tile_off, ws.edge_sorted);
                }
                {
                    DVG_PROF(st, "k_edge");
                    // at most one partial chunk per tile plus nsamp / echunk full ones
                    unsigned grid = (unsigned)(ntiles + (nsamp + G.echunk - 1) / G.echunk);
                    if (G.echunk == kBlock)
                        k_edge<true><<<grid, kBlock, 0, st>>>(S, G, ws.edges, ws.tile_off, ws.tile_count,
                                                              ws.chunk_off, ntiles, ws.edge_sorted,
                                                              a.background_image, ws.weight, a.d_render_image,
                                                              a.d_translation);
                    else
                        k_edge<false><<<grid, kBlock, 0, st>>>(S, G, ws.edges, ws.tile_off, ws.tile_count,
                                                               ws.chunk_off, ntiles, ws.edge_sorted,
                                                               a.background_image, ws.weight, a.d_render_image,
                                                               a.d_translation);
                }
            }
        }
    }
    if (sdf_out) {
        DVG_PROF(st, "k_sdf");
        if (eval_mode) {
            int count = a.num_eval_positions;
            k_sdf<<<(count + kBlock - 1) / kBlock, kBlock, 0, st>>>(S, G, a.eval_positions, a.num_eval_positions,
                                                                     a.render_sdf, a.d_render_sdf, a.d_translation);
        } else {
            k_sdf_tiled<<<ntiles, kBlock, 0, st>>>(S, G, a.render_sdf, a.d_render_sdf, a.d_translation);
        }
    }
    DVG_CHECK(hipGetLastError());
    // nothing waits here: conditions the kernels flag are reported by scene_status
    return 0;
}

// ---------------------------------------------------------------- allocation-free render
// dvg_render_noalloc's workspace: one file-scope device array per GPU, carved like
// Workspace for at most kNaMaxPixels pixels and kNaMaxSamples samples (54 MiB), held by
// every GPU the library runs on until the process ends.  Host code never uses the
// array's name as an address: na_base() asks the runtime for the device address, once per
// device, with that device current.
namespace {
constexpr size_t kNaMaxPixels = size_t(1) << 19;
constexpr size_t kNaMaxSamples = size_t(1) << 19;
constexpr size_t kNaMaxTiles = kNaMaxPixels;  // every tile holds at least one pixel
constexpr size_t na_align(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kNaWeight = 0;
constexpr size_t kNaColors = kNaWeight + na_align(kNaMaxPixels * sizeof(float));
constexpr size_t kNaEdges = kNaColors + na_align(kNaMaxSamples * sizeof(float4));
constexpr size_t kNaEdgeTile = kNaEdges + na_align(kNaMaxSamples * sizeof(EdgeSample));
constexpr size_t kNaEdgeSorted = kNaEdgeTile + na_align(kNaMaxSamples * sizeof(int));
constexpr size_t kNaTileCount = kNaEdgeSorted + na_align(kNaMaxSamples * sizeof(int));
constexpr size_t kNaTileOff = kNaTileCount + na_align(kNaMaxTiles * sizeof(int));
constexpr size_t kNaEdgeRank = kNaTileOff + na_align(kNaMaxTiles * sizeof(int));
constexpr size_t kNaChunkOff = kNaEdgeRank + na_align(kNaMaxSamples * sizeof(int));
constexpr size_t kNaBytes = kNaChunkOff + na_align((kNaMaxTiles + 1) * sizeof(int));
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

static std::mutex na_mu;
static unsigned char *na_ptr[kNaMaxDevices];

// The device address of `scratch` on `device` (the current device).
static int na_base(int device, unsigned char **out) {
    if (device < 0 || device >= kNaMaxDevices) return set_error("device index out of range");
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        na_ptr[device] = (unsigned char *)p;
    }
    *out = na_ptr[device];
    return 0;
}

int render_noalloc(SceneHost &sc, const dvg_render_args &a) {
    if ((int64_t)a.num_samples_x * a.num_samples_y > 256)
        return set_error("dvg_render_noalloc takes at most 256 samples per pixel");
    const size_t npix = (size_t)a.width * a.height;
    const size_t nsamp = npix * a.num_samples_x * a.num_samples_y;
    if (npix > kNaMaxPixels || nsamp > kNaMaxSamples)
        return set_error("dvg_render_noalloc renders at most " + std::to_string(kNaMaxPixels) + " pixels and " +
                         std::to_string(kNaMaxSamples) + " samples");
    unsigned char *base = nullptr;
    if (int e = na_base(sc.device, &base)) return e;
    Workspace ws;
    ws.npix_cap = kNaMaxPixels;
    ws.nsamp_cap = kNaMaxSamples;
    ws.ntiles_cap = kNaMaxTiles;
    ws.weight = (float *)(base + kNaWeight);
    ws.colors = (float4 *)(base + kNaColors);
    ws.edges = (EdgeSample *)(base + kNaEdges);
    ws.edge_tile = (int *)(base + kNaEdgeTile);
    ws.edge_sorted = (int *)(base + kNaEdgeSorted);
    ws.tile_count = (int *)(base + kNaTileCount);
    ws.tile_off = (int *)(base + kNaTileOff);
    ws.edge_rank = (int *)(base + kNaEdgeRank);
    ws.chunk_off = (int *)(base + kNaChunkOff);
    return render_tiled(sc, a, &ws);
}

int render_plain(SceneHost &sc, const dvg_render_args &a) {
    hipStream_t st = (hipStream_t)a.stream;
    Geom G = make_geom(sc, a);
    const bool eval_mode = a.eval_positions && a.num_eval_positions > 0;
    const int pre = a.use_prefiltering ? 1 : 0;
    const size_t npix = (size_t)G.W * G.H;
    const size_t nsamp = npix * G.spp;
    DevScene S = sc.dev;
    if (eval_mode && (a.render_image || a.d_render_image))
        return set_error("eval_positions are supported for the SDF output only (as in the reference)");
    // the device rebuild keeps the group boxes current but not the tree over them
    if (sc.top_stale)
        if (int e = rebuild_top(sc)) return e;
    if (sc.num_shapes == 0) {
        if (int e = render_empty(a, npix, st)) return e;
        DVG_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    float *weight = nullptr;
    if (!eval_mode) {
        DVG_CHECK(hipMalloc(&weight, npix * sizeof(float)));
        DVG_CHECK(hipMemsetAsync(weight, 0, npix * sizeof(float), st));
    }
    const bool color_out = a.render_image || a.d_render_image;
    const bool sdf_out = a.render_sdf || a.d_render_sdf;
    if ((color_out || sdf_out) && weight) {
        DVG_PROF(st, "kp_weight");
        kp_weight<<<(unsigned)((nsamp + 63) / 64), 64, 0, st>>>(S, G, weight);
    }
    float4 *dbg_rec = nullptr;
    int *dbg_kind = nullptr;
    if (color_out && a.d_render_image && a.d_background_image && a.background_image) {
        DVG_CHECK(hipMalloc(&dbg_rec, nsamp * sizeof(float4)));
        DVG_CHECK(hipMalloc(&dbg_kind, nsamp * sizeof(int)));
    }
    if (color_out) {
        DVG_PROF(st, "kp_render");
        if (pre)
            kp_render<1><<<(unsigned)((nsamp + 63) / 64), 64, 0, st>>>(S, G, a.background_image, a.render_image,
                                                                       weight, a.d_render_image, dbg_rec, dbg_kind,
                                                                       a.d_translation, sc.err);
        else
            kp_render<0><<<(unsigned)((nsamp + 63) / 64), 64, 0, st>>>(S, G, a.background_image, a.render_image,
                                                                       weight, a.d_render_image, dbg_rec, dbg_kind,
                                                                       a.d_translation, sc.err);
    }
    if (dbg_rec) {
        kp_fold_dbg<<<(unsigned)((npix + 63) / 64), 64, 0, st>>>(G, dbg_rec, dbg_kind, a.d_background_image);
        DVG_CHECK(hipStreamSynchronize(st));
        DVG_CHECK(hipFree(dbg_rec));
        DVG_CHECK(hipFree(dbg_kind));
    }
    if (sdf_out) {
        int count = eval_mode ? a.num_eval_positions : (int)nsamp;
        k_sdf<<<(count + 63) / 64, 64, 0, st>>>(S, G, eval_mode ? a.eval_positions : nullptr, a.num_eval_positions,
                                                a.render_sdf, a.d_render_sdf, a.d_translation);
    }
    if (!pre && a.d_render_image) {
        EdgeSample *edges = nullptr;
        int *ids = nullptr, *ids2 = nullptr;
        uint32_t *keys = nullptr, *keys2 = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        DVG_CHECK(hipMalloc(&edges, nsamp * sizeof(EdgeSample)));
        DVG_CHECK(hipMalloc(&ids, nsamp * sizeof(int)));
        DVG_CHECK(hipMalloc(&ids2, nsamp * sizeof(int)));
        DVG_CHECK(hipMalloc(&keys, nsamp * sizeof(uint32_t)));
        DVG_CHECK(hipMalloc(&keys2, nsamp * sizeof(uint32_t)));
        kp_edge_sample<<<(unsigned)((nsamp + 63) / 64), 64, 0, st>>>(S, G, (int)nsamp, edges, ids, keys);
        DVG_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, ids, ids2, (int)nsamp, 0, 32, st));
        DVG_CHECK(hipMalloc(&tmp, tmp_bytes));
        DVG_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, ids, ids2, (int)nsamp, 0, 32, st));
        {
            DVG_PROF(st, "kp_edge");
            kp_edge<<<(unsigned)((nsamp + 63) / 64), 64, 0, st>>>(S, G, edges, ids2, a.background_image, weight,
                                                                  a.d_render_image, a.d_translation, sc.err);
        }
        DVG_CHECK(hipStreamSynchronize(st));
        DVG_CHECK(hipFree(tmp));
        DVG_CHECK(hipFree(edges));
        DVG_CHECK(hipFree(ids));
        DVG_CHECK(hipFree(ids2));
        DVG_CHECK(hipFree(keys));
        DVG_CHECK(hipFree(keys2));
    }
    DVG_CHECK(hipGetLastError());
    DVG_CHECK(hipStreamSynchronize(st));
    if (weight) DVG_CHECK(hipFree(weight));
    return check_err_flag(sc, st);
}


// ------------------------------------------------------------------ BATCHED launches
// Many (scene, render args) jobs rendered with one kernel per stage over the work items
// of all jobs.  The job table lives in HBM; each stage has a prefix array of its
// workgroup counts per job, and a workgroup finds its job by binary search (uniform per
// workgroup, so the job's fields are read with scalar loads).  The workspace is carved
// when the batch is created, so dvg_batch_render allocates nothing, waits for nothing
// and can be captured into a HIP graph.
struct Job {
    DevScene S;
    Geom G;
    const float *bg;
    float *img;
    const float *d_img;
    float *d_bg;
    float *d_tr;
    float *sdf;
    const float *d_sdf;
    const float *eval;
    int num_eval;
    int pre;
    float *weight;
    float4 *colors;
    EdgeSample *edges;
    int *edge_tile, *edge_sorted, *edge_rank, *tile_count, *tile_off, *chunk_off;
    int *err;
    float *grad;
    int grad_n;
    int ntiles, npix, nsamp;
    uint16_t *frag;  // the scene's kept fragment lists, or null
    int *pkeep;      // its kept prefiltered fragments, or null
};

// kStChunk / kStChunk2: edge chunks of the jobs with one side / both sides per lane
enum { kStPix = 0, kStTile, kStSamp, kStChunk, kStSdf, kStZero, kStChunk2, kNumStages };

// prefix arrays: pref[stage * (nj + 1) + k] = first workgroup of job k in that stage
DVG_DEV int find_job(const int *pref, int nj, int blk) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void kb_zero(const Job *jobs, const int *pref, int nj, int zero_grads) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    int b = blockIdx.x - pref[j], nb = pref[j + 1] - pref[j];
    if (J.d_img && J.tile_count)
        for (int i = b * kBlock + threadIdx.x; i < J.ntiles; i += nb * kBlock) J.tile_count[i] = 0;
    if (zero_grads && (J.d_img || J.d_sdf))
        for (int i = b * kBlock + threadIdx.x; i < J.grad_n; i += nb * kBlock) J.grad[i] = 0.f;
}

__global__ __launch_bounds__(kBlock) void kb_weight(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    weight_pixel(J.S, J.G, (blockIdx.x - pref[j]) * kBlock + threadIdx.x, J.weight);
}

template <int PRE>
__global__ __launch_bounds__(kBlock, PRE ? DVG_WAVES_BATCH_COLOR_PRE : DVG_WAVES_BATCH_COLOR) void kb_color(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.img) return;
    __shared__ CandLds L;
    color_tile<PRE>(L, J.S, J.G, blockIdx.x - pref[j], J.bg, J.colors, J.err, PRE ? nullptr : J.frag,
                    PRE ? J.pkeep : nullptr);
}

__global__ __launch_bounds__(kBlock) void kb_resolve(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.img) return;
    resolve_pixel(J.S, J.G, (blockIdx.x - pref[j]) * kBlock + threadIdx.x, J.pre, J.weight, J.colors, J.img);
}

template <int PRE>
__global__ __launch_bounds__(kBlock, DVG_WAVES_BATCH_COLOR_BWD) void kb_color_bwd(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.d_img) return;
    __shared__ ColorBwdLds L;
    color_bwd_tile<PRE>(L, J.S, J.G, blockIdx.x - pref[j], J.bg, J.weight, J.d_img, J.d_bg, J.d_tr, J.err,
                        PRE ? nullptr : J.frag, PRE ? J.pkeep : nullptr);
}

__global__ __launch_bounds__(kBlock) void kb_edge_sample(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.d_img) return;
    edge_sample_one(J.S, J.G, (blockIdx.x - pref[j]) * kBlock + threadIdx.x, J.nsamp, J.edges, J.edge_tile,
                    J.edge_rank, J.tile_count);
}

__global__ __launch_bounds__(1024) void kb_scan(const Job *jobs, int j0) {
    const Job &J = jobs[j0 + blockIdx.x];
    if (!J.d_img) return;
    __shared__ ScanLds<1024> L;
    scan_tiles<1024>(L, J.tile_count, J.tile_off, J.chunk_off, J.ntiles, J.G.echunk);
}

__global__ __launch_bounds__(kBlock) void kb_edge_bin(const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.d_img) return;
    edge_bin_one((blockIdx.x - pref[j]) * kBlock + threadIdx.x, J.nsamp, J.edge_tile, J.edge_rank, J.tile_off,
                 J.edge_sorted);
}

// one launch per variant of the edge kernel, over the jobs that use it (the others have no
// workgroups in its stage)
template <bool BOTH>
__global__ __launch_bounds__(kBlock, BOTH ? DVG_WAVES_EDGE_BOTH : DVG_WAVES_BATCH_EDGE) void kb_edge(
    const Job *jobs, const int *pref, int nj) {
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (!J.d_img) return;
    __shared__ EdgeLds L;
    edge_chunk<BOTH>(L, J.S, J.G, blockIdx.x - pref[j], J.edges, J.tile_off, J.tile_count, J.chunk_off, J.ntiles,
                     J.edge_sorted, J.bg, J.weight, J.d_img, J.d_tr);
}

__global__ __launch_bounds__(kBlock, DVG_WAVES_SDF) void kb_sdf(const Job *jobs, const int *pref, int nj) {
    __shared__ SdfLds L;
    int j = find_job(pref, nj, blockIdx.x);
    const Job &J = jobs[j];
    if (J.eval)
        sdf_sample(J.S, J.G, (blockIdx.x - pref[j]) * kBlock + threadIdx.x, J.eval, J.num_eval, J.sdf, J.d_sdf,
                   J.d_tr);
    else
        sdf_tile(L, J.S, J.G, blockIdx.x - pref[j], J.sdf, J.d_sdf, J.d_tr);
}

// ------------------------------------------------------------------ persistent batch kernel
// One launch per batch render: a fixed grid of workgroups (as many as the device holds at
// once) takes tasks -- (stage, job, workgroup of that stage) -- from a counter in HBM, in
// a fixed order: stage by stage over all jobs, in the order of kPStage.  A task first
// waits until the stages it depends on are done for its own job (per-job counters of
// finished tasks), so a scene's resolve can start while other scenes still shade, and a
// small scene's edge pass overlaps the colour passes of large ones.  Every dependency is
// on a task earlier in the order, and a task is taken only by a running workgroup, so
// every wait ends: the earliest unfinished task waits for nothing.  As a safeguard a wait
// gives up after a time budget (the host reports it as an error); the results of such a
// render are not defined.  The last workgroup to leave resets the counters for the next
// launch, so the launch can be replayed from a HIP graph.
//
// DISABLED (batch_create refuses DVG_BATCH_PERSISTENT): the first GPU run of this kernel
// (a batch of ten small mixed jobs) did not finish within three minutes, and neither its
// dependency waits nor their time budget explain that on reading; the kernel is kept
// for reference, not launched.
enum PStage {
    kPZero = 0,
    kPWeight,
    kPColor0,
    kPColor1,
    kPResolve,
    kPColorBwd0,
    kPColorBwd1,
    kPEdgeSample,
    kPScan,
    kPEdgeBin,
    kPEdge,
    kPSdf,
    kNumPStages
};

// stages whose tasks of the same job must be done first (bit masks of PStage)
DVG_DEV uint32_t pstage_deps(int s) {
    switch (s) {
        case kPResolve: return (1u << kPWeight) | (1u << kPColor0) | (1u << kPColor1);
        // gradients cleared; weights; the fragment lists the job's forward keeps
        case kPColorBwd0: return (1u << kPZero) | (1u << kPWeight) | (1u << kPColor0);
        case kPColorBwd1: return (1u << kPZero) | (1u << kPWeight) | (1u << kPColor1);
        case kPEdgeSample: return 1u << kPZero;  // tile counts cleared
        case kPScan: return 1u << kPEdgeSample;
        case kPEdgeBin: return 1u << kPScan;
        case kPEdge: return (1u << kPEdgeBin) | (1u << kPWeight) | (1u << kPZero);
        case kPSdf: return 1u << kPZero;
        default: return 0u;
    }
}

// words of the task state ahead of the per-job counters
enum { kPNext = 0, kPExited, kPAbort, kPStateInts = 4 };

union PLds {
    CandLds color;
    ColorBwdLds bwd;
    EdgeLds edge;
    SdfLds sdf;
    ScanLds<kBlock> scan;
};

// One lane waits for the task's prerequisites (`done` = the job's counters); false when
// the render was given up (this or another wait exceeded the budget).
DVG_DEV bool pstage_wait(const int *ppref, int nj, int j, int s, int *done, int *state, uint64_t budget) {
    uint32_t deps = pstage_deps(s);
    uint64_t t0 = 0;
    while (deps) {
        const int d = __builtin_ctz(deps);
        deps &= deps - 1;
        const int *pf = ppref + d * (nj + 1);
        const int need = pf[j + 1] - pf[j];
        while (need > 0 && __hip_atomic_load(done + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(state + kPAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
            const uint64_t now = wall_clock64();
            if (t0 == 0) {
                t0 = now;
            } else if (now - t0 > budget) {
                atomicOr(state + kPAbort, 1);
                return false;
            }
            __builtin_amdgcn_s_sleep(8);
        }
    }
    __threadfence();  // acquire: what the finished tasks wrote is visible from here on
    return __hip_atomic_load(state + kPAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

DVG_DEV void run_pstage(PLds &L, const Job &J, int s, int b, int nb, int zero_grads) {
    switch (s) {
        case kPZero:
            if (J.d_img && J.tile_count)
                for (int i = b * kBlock + threadIdx.x; i < J.ntiles; i += nb * kBlock) J.tile_count[i] = 0;
            if (zero_grads && (J.d_img || J.d_sdf))
                for (int i = b * kBlock + threadIdx.x; i < J.grad_n; i += nb * kBlock) J.grad[i] = 0.f;
            break;
        case kPWeight: weight_pixel(J.S, J.G, b * kBlock + threadIdx.x, J.weight); break;
        case kPColor0: color_tile<0>(L.color, J.S, J.G, b, J.bg, J.colors, J.err, J.frag, nullptr); break;
        case kPColor1: color_tile<1>(L.color, J.S, J.G, b, J.bg, J.colors, J.err, nullptr, J.pkeep); break;
        case kPResolve:
            resolve_pixel(J.S, J.G, b * kBlock + threadIdx.x, J.pre, J.weight, J.colors, J.img);
            break;
        case kPColorBwd0:
            color_bwd_tile<0>(L.bwd, J.S, J.G, b, J.bg, J.weight, J.d_img, J.d_bg, J.d_tr, J.err, J.frag, nullptr);
            break;
        case kPColorBwd1:
            color_bwd_tile<1>(L.bwd, J.S, J.G, b, J.bg, J.weight, J.d_img, J.d_bg, J.d_tr, J.err, nullptr, J.pkeep);
            break;
        case kPEdgeSample:
            edge_sample_one(J.S, J.G, b * kBlock + threadIdx.x, J.nsamp, J.edges, J.edge_tile, J.edge_rank,
                            J.tile_count);
            break;
        case kPScan:
            scan_tiles<kBlock>(L.scan, J.tile_count, J.tile_off, J.chunk_off, J.ntiles, J.G.echunk);
            break;
        case kPEdgeBin:
            edge_bin_one(b * kBlock + threadIdx.x, J.nsamp, J.edge_tile, J.edge_rank, J.tile_off, J.edge_sorted);
            break;
        case kPEdge:
            if (J.G.echunk == kBlock)
                edge_chunk<true>(L.edge, J.S, J.G, b, J.edges, J.tile_off, J.tile_count, J.chunk_off, J.ntiles,
                                 J.edge_sorted, J.bg, J.weight, J.d_img, J.d_tr);
            else
                edge_chunk<false>(L.edge, J.S, J.G, b, J.edges, J.tile_off, J.tile_count, J.chunk_off, J.ntiles,
                                  J.edge_sorted, J.bg, J.weight, J.d_img, J.d_tr);
            break;
        case kPSdf:
            if (J.eval)
                sdf_sample(J.S, J.G, b * kBlock + threadIdx.x, J.eval, J.num_eval, J.sdf, J.d_sdf, J.d_tr);
            else
                sdf_tile(L.sdf, J.S, J.G, b, J.sdf, J.d_sdf, J.d_tr);
            break;
        default: break;
    }
}

// ppref[s * (nj + 1) + k]: first task of job k within stage s; pbase[s]: first task of
// stage s (pbase[kNumPStages] = all tasks); state: kPStateInts words, then the
// per-job counters done[j * kNumPStages + s].
__global__ __launch_bounds__(kBlock, DVG_WAVES_PERSIST) void kb_persist(const Job *jobs, int nj, const int *ppref,
                                                                       const int *pbase, int *state, int zero_grads,
                                                                       uint64_t budget) {
    __shared__ PLds L;
    __shared__ int s_task;
    __shared__ int s_go;
    __shared__ int s_last;
    int *done = state + kPStateInts;
    const int total = pbase[kNumPStages];
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(state + kPNext, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_task);
        if (t >= total) break;
        int s = 0;
        while (pbase[s + 1] <= t) s++;
        const int *pf = ppref + s * (nj + 1);
        const int local = t - pbase[s];
        const int j = __builtin_amdgcn_readfirstlane(find_job(pf, nj, local));
        const int b = local - pf[j], nb = pf[j + 1] - pf[j];
        // a render given up skips its remaining tasks (their inputs may not be there)
        if (threadIdx.x == 0) s_go = pstage_wait(ppref, nj, j, s, done + j * kNumPStages, state, budget) ? 1 : 0;
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(s_go)) run_pstage(L, jobs[j], s, b, nb, zero_grads);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();  // release: this task's writes before its count
            atomicAdd(done + j * kNumPStages + s, 1);
        }
    }
    // the last workgroup out clears the counters for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(state + kPExited, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        for (int i = threadIdx.x; i < nj * kNumPStages; i += blockDim.x) done[i] = 0;
        if (threadIdx.x == 0) {
            state[kPNext] = 0;
            state[kPExited] = 0;
        }
        __threadfence();
    }
}

struct BatchHost {
    int device = 0;
    int n = 0;
    int zero_grads = 0;
    std::vector<SceneHost *> scenes;
    std::vector<Job> jobs;         // host copy, in class order
    std::vector<int> order;        // jobs[k] renders item order[k]
    // classes: [0] colour without prefiltering, [1] colour prefiltered, [2] SDF
    int cls_begin[4] = {0, 0, 0, 0};
    bool any_img[3] = {false, false, false}, any_dimg[3] = {false, false, false};
    std::vector<int> pref;         // per class, per stage: (njobs_in_class + 1) entries
    std::vector<size_t> pref_off;  // offset into pref of (class, stage)
    std::vector<int> total;        // workgroups of (class, stage)
    void *dev_block = nullptr;     // job table + prefix arrays
    size_t dev_bytes = 0, dev_granted = 0;
    void *ws = nullptr;            // per-job scratch
    size_t ws_bytes = 0, ws_granted = 0;
    hipStream_t last_stream = nullptr;
    std::vector<SceneHost *> keep_holds;  // one entry per job holding its scene's kept lists
    // persistent launch (DVG_BATCH_PERSISTENT): task prefix arrays and state in dev_block
    bool persistent = false;
    size_t ppref_off = 0, pbase_off = 0, pstate_off = 0;
    int ptotal = 0, pgrid = 0;
    uint64_t pbudget = 0;
};

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

int batch_create(SceneHost *const *scenes, const dvg_render_args *args, int n, int flags, BatchHost **out) {
    if (n <= 0) return set_error("a batch needs at least one job");
    auto *B = new (std::nothrow) BatchHost();
    if (!B) return set_error("out of host memory");
    std::unique_ptr<BatchHost> guard(B);
    B->n = n;
    B->zero_grads = (flags & DVG_BATCH_ZERO_GRADS) ? 1 : 0;
    B->persistent = (flags & DVG_BATCH_PERSISTENT) != 0;
    // Switched off: its first GPU run did not finish (see kb_persist); the cause is not
    // known, so the launch is refused rather than risked.
    if (B->persistent)
        return set_error("DVG_BATCH_PERSISTENT is disabled in this build: the persistent batch kernel did not "
                         "finish on the GPU and the cause is not known; use the one-launch-per-stage batch");
    B->device = scenes[0]->device;
    // classify
    std::vector<int> cls(n);
    for (int i = 0; i < n; i++) {
        const dvg_render_args &a = args[i];
        if (!scenes[i]) return set_error("null scene in batch");
        if (scenes[i]->device != B->device) return set_error("all scenes of a batch must live on one GPU");
        if (a.width <= 0 || a.height <= 0 || a.num_samples_x <= 0 || a.num_samples_y <= 0)
            return set_error("width, height and sample counts must be positive");
        if (a.num_samples_x * a.num_samples_y > kBlock)
            return set_error("batched renders support at most 256 samples per pixel");
        if ((int64_t)a.width * a.height * a.num_samples_x * a.num_samples_y > (int64_t)0x7fffffff)
            return set_error("too many samples for one render");
        bool color = a.render_image || a.d_render_image;
        bool sdf = a.render_sdf || a.d_render_sdf;
        if (color == sdf) return set_error("a batch job renders either colour or SDF");
        bool eval = a.eval_positions && a.num_eval_positions > 0;
        if (color && eval) return set_error("eval_positions are supported for the SDF output only");
        cls[i] = sdf ? 2 : (a.use_prefiltering ? 1 : 0);
    }
    // colour jobs per scene: a scene in more than one colour job of the batch keeps no
    // fragment lists (its jobs would write one buffer at once)
    std::map<const SceneHost *, int> color_jobs;
    for (int i = 0; i < n; i++)
        if (cls[i] != 2) color_jobs[scenes[i]]++;
    // within a class, jobs with the most samples first: their long tasks start first
    for (int c = 0; c < 3; c++) {
        const size_t k0 = B->order.size();
        for (int i = 0; i < n; i++)
            if (cls[i] == c) B->order.push_back(i);
        auto samples = [&](int i) {
            return (int64_t)args[i].width * args[i].height * args[i].num_samples_x * args[i].num_samples_y;
        };
        std::stable_sort(B->order.begin() + k0, B->order.end(),
                         [&](int a, int b) { return samples(a) > samples(b); });
    }
    for (int c = 0, k = 0; c < 3; c++) {
        B->cls_begin[c] = k;
        for (int i = 0; i < n; i++) k += cls[i] == c;
        B->cls_begin[c + 1] = k;
    }
    // workspace layout
    std::vector<size_t> o_weight(n), o_colors(n), o_edges(n), o_etile(n), o_esort(n), o_tcount(n), o_toff(n),
        o_erank(n), o_chunk(n);
    size_t ws = 0;
    B->jobs.resize(n);
    for (int k = 0; k < n; k++) {
        int i = B->order[k];
        SceneHost &sc = *scenes[i];
        const dvg_render_args &a = args[i];
        Job &J = B->jobs[k];
        std::memset(&J, 0, sizeof(Job));
        J.S = sc.dev;
        J.G = make_geom(sc, a);
        J.bg = a.background_image;
        J.img = a.render_image;
        J.d_img = a.d_render_image;
        J.d_bg = a.d_background_image;
        J.d_tr = a.d_translation;
        J.sdf = a.render_sdf;
        J.d_sdf = a.d_render_sdf;
        bool eval = a.eval_positions && a.num_eval_positions > 0;
        J.eval = eval ? a.eval_positions : nullptr;
        J.num_eval = eval ? a.num_eval_positions : 0;
        J.pre = a.use_prefiltering ? 1 : 0;
        J.err = sc.err;
        J.grad = sc.grad;
        J.grad_n = (int)sc.grad_floats;
        J.npix = J.G.W * J.G.H;
        J.nsamp = J.npix * J.G.spp;
        J.ntiles = J.G.tiles_x * J.G.tiles_y;
        if (cls[i] == 2) continue;
        // kept fragment lists: a forward job writes them, a backward job replays them when
        // the key the forward wrote matches its own (checked on the device)
        if (color_jobs[&sc] == 1) {
            const size_t ns = (size_t)J.nsamp;
            if (!J.pre) {
                const size_t bytes = (size_t)(kFragSlots + 1) * ns * sizeof(uint16_t);
                if (bytes <= kFragCacheBytes && sc.dev.ng < kFragMaxGroups) {
                    void *b = sc.frag;
                    const bool ok = J.img ? keep_buffer(sc, b, sc.frag_granted, sc.frag_stream, bytes)
                                          : (b && sc.frag_granted >= bytes);
                    sc.frag = (uint16_t *)b;
                    J.frag = ok ? sc.frag : nullptr;
                }
            } else {
                const size_t bytes = (size_t)(kPFragSlots * kPFragWords + 1) * ns * sizeof(int);
                if (bytes <= kFragCacheBytes) {
                    void *b = sc.pkeep;
                    const bool ok = J.img ? keep_buffer(sc, b, sc.pkeep_granted, sc.pkeep_stream, bytes)
                                          : (b && sc.pkeep_granted >= bytes);
                    sc.pkeep = (int *)b;
                    J.pkeep = ok ? sc.pkeep : nullptr;
                }
            }
            if (J.frag || J.pkeep) {
                sc.keep_refs++;
                B->keep_holds.push_back(&sc);
            }
        }
        o_weight[k] = ws, ws += al256((size_t)J.npix * 4);
        if (J.img) o_colors[k] = ws, ws += al256((size_t)J.nsamp * 16);
        if (J.d_img && !J.pre) {
            o_edges[k] = ws, ws += al256((size_t)J.nsamp * sizeof(EdgeSample));
            o_etile[k] = ws, ws += al256((size_t)J.nsamp * 4);
            o_esort[k] = ws, ws += al256((size_t)J.nsamp * 4);
            o_tcount[k] = ws, ws += al256((size_t)J.ntiles * 4);
            o_toff[k] = ws, ws += al256((size_t)J.ntiles * 4);
            o_erank[k] = ws, ws += al256((size_t)J.nsamp * 4);
            o_chunk[k] = ws, ws += al256((size_t)(J.ntiles + 1) * 4);
        }
    }
    // prefix arrays per class and stage
    // workgroups of a job in each stage (0 when the job has no work there)
    auto blocks = [](const Job &J, int stage) -> int {
        bool color = J.img || J.d_img, edges = J.d_img && !J.pre, sdf = J.sdf || J.d_sdf;
        switch (stage) {
            case kStPix: return color ? (J.npix + kBlock - 1) / kBlock : 0;
            case kStTile: return color ? J.ntiles : 0;
            case kStSamp: return edges ? (J.nsamp + kBlock - 1) / kBlock : 0;
            case kStChunk:
                return edges && J.G.echunk != kBlock ? J.ntiles + (J.nsamp + J.G.echunk - 1) / J.G.echunk : 0;
            case kStChunk2:
                return edges && J.G.echunk == kBlock ? J.ntiles + (J.nsamp + J.G.echunk - 1) / J.G.echunk : 0;
            case kStSdf: return sdf ? (J.eval ? (J.num_eval + kBlock - 1) / kBlock : J.ntiles) : 0;
            default: return std::max(1, std::min(64, (std::max(J.grad_n, J.ntiles) + kBlock - 1) / kBlock));
        }
    };
    for (int c = 0; c < 3; c++)
        for (int k = B->cls_begin[c]; k < B->cls_begin[c + 1]; k++) {
            B->any_img[c] |= B->jobs[k].img != nullptr;
            B->any_dimg[c] |= B->jobs[k].d_img != nullptr;
        }
    B->pref_off.assign(3 * kNumStages, 0);
    B->total.assign(3 * kNumStages, 0);
    for (int c = 0; c < 3; c++) {
        int j0 = B->cls_begin[c], nj = B->cls_begin[c + 1] - j0;
        for (int st = 0; st < kNumStages; st++) {
            B->pref_off[c * kNumStages + st] = B->pref.size();
            int acc = 0;
            for (int k = 0; k < nj; k++) {
                B->pref.push_back(acc);
                acc += blocks(B->jobs[j0 + k], st);
            }
            B->pref.push_back(acc);
            B->total[c * kNumStages + st] = acc;
        }
    }
    // persistent launch: task counts per (stage, job) over all jobs
    std::vector<int> ppref, pbase;
    if (B->persistent) {
        auto ptasks = [&](const Job &J, int stage) -> int {
            bool color = J.img || J.d_img, edges = J.d_img && !J.pre, sdf = J.sdf || J.d_sdf;
            switch (stage) {
                case kPZero: return blocks(J, kStZero);
                case kPWeight: return color ? (J.npix + kBlock - 1) / kBlock : 0;
                case kPColor0: return J.img && !J.pre ? J.ntiles : 0;
                case kPColor1: return J.img && J.pre ? J.ntiles : 0;
                case kPResolve: return J.img ? (J.npix + kBlock - 1) / kBlock : 0;
                case kPColorBwd0: return J.d_img && !J.pre ? J.ntiles : 0;
                case kPColorBwd1: return J.d_img && J.pre ? J.ntiles : 0;
                case kPEdgeSample:
                case kPEdgeBin: return edges ? (J.nsamp + kBlock - 1) / kBlock : 0;
                case kPScan: return edges ? 1 : 0;
                case kPEdge: return edges ? J.ntiles + (J.nsamp + J.G.echunk - 1) / J.G.echunk : 0;
                case kPSdf: return sdf ? (J.eval ? (J.num_eval + kBlock - 1) / kBlock : J.ntiles) : 0;
                default: return 0;
            }
        };
        int64_t acc_all = 0;
        for (int st = 0; st < kNumPStages; st++) {
            pbase.push_back((int)acc_all);
            int acc = 0;
            for (int k = 0; k < n; k++) {
                ppref.push_back(acc);
                acc += ptasks(B->jobs[k], st);
            }
            ppref.push_back(acc);
            acc_all += acc;
        }
        if (acc_all > (int64_t)0x7fffffff - 4 * 65536) return set_error("too many tasks for one persistent batch");
        pbase.push_back((int)acc_all);
        B->ptotal = (int)acc_all;
    }
    B->scenes.assign(scenes, scenes + n);
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != B->device) (void)hipSetDevice(B->device);
    auto restore = [&]() {
        if (prev >= 0 && prev != B->device) (void)hipSetDevice(prev);
    };
    // both blocks come from the per-device cache, so batches created and destroyed in a
    // serving loop neither allocate nor free device memory (hipFree would wait for the
    // whole device)
    size_t jbytes = al256(sizeof(Job) * n);
    B->dev_bytes = jbytes + al256(B->pref.size() * sizeof(int));
    if (B->persistent) {
        B->ppref_off = B->dev_bytes;
        B->dev_bytes += al256(ppref.size() * sizeof(int));
        B->pbase_off = B->dev_bytes;
        B->dev_bytes += al256(pbase.size() * sizeof(int));
        B->pstate_off = B->dev_bytes;  // zeroed below with the rest of the block
        B->dev_bytes += al256((size_t)(kPStateInts + n * kNumPStages) * sizeof(int));
        int per_cu = 0, cus = 0, rate_khz = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_persist, kBlock, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->device) != hipSuccess ||
            hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, B->device) != hipSuccess) {
            if (prev >= 0 && prev != B->device) (void)hipSetDevice(prev);
            return set_error("persistent batch: device query failed");
        }
        B->pgrid = std::max(1, std::min(B->ptotal, std::max(1, per_cu) * std::max(1, cus)));
        // a dependency wait gives up after 20 s of wall clock (a render takes milliseconds)
        B->pbudget = (uint64_t)std::max(rate_khz, 1) * 1000ull * 20ull;
    }
    B->dev_block = cache_alloc(B->device, B->dev_bytes, &B->dev_granted);
    if (!B->dev_block) {
        restore();
        return set_error("out of device memory for the batch table");
    }
    if (ws) {
        B->ws = cache_alloc(B->device, ws, &B->ws_granted);
        if (!B->ws) {
            cache_free(B->device, B->dev_block, B->dev_granted);
            restore();
            return set_error("out of device memory for the batch workspace");
        }
    }
    B->ws_bytes = ws;
    uint8_t *W = (uint8_t *)B->ws;
    for (int k = 0; k < n; k++) {
        Job &J = B->jobs[k];
        if (cls[B->order[k]] == 2) continue;
        J.weight = (float *)(W + o_weight[k]);
        if (J.img) J.colors = (float4 *)(W + o_colors[k]);
        if (J.d_img && !J.pre) {
            J.edges = (EdgeSample *)(W + o_edges[k]);
            J.edge_tile = (int *)(W + o_etile[k]);
            J.edge_sorted = (int *)(W + o_esort[k]);
            J.tile_count = (int *)(W + o_tcount[k]);
            J.tile_off = (int *)(W + o_toff[k]);
            J.edge_rank = (int *)(W + o_erank[k]);
            J.chunk_off = (int *)(W + o_chunk[k]);
        }
    }
    std::vector<uint8_t> host(B->dev_bytes, 0);
    std::memcpy(host.data(), B->jobs.data(), sizeof(Job) * n);
    std::memcpy(host.data() + jbytes, B->pref.data(), B->pref.size() * sizeof(int));
    if (B->persistent) {
        std::memcpy(host.data() + B->ppref_off, ppref.data(), ppref.size() * sizeof(int));
        std::memcpy(host.data() + B->pbase_off, pbase.data(), pbase.size() * sizeof(int));
    }
    int rc = upload(B->device, B->dev_block, host.data(), B->dev_bytes);
    restore();
    if (rc) {
        cache_free(B->device, B->dev_block, B->dev_granted);
        if (B->ws) cache_free(B->device, B->ws, B->ws_granted);
        return rc;
    }
    *out = guard.release();
    return 0;
}

int batch_render(BatchHost &B, hipStream_t st) {
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != B.device) (void)hipSetDevice(B.device);
    const Job *jobs = (const Job *)B.dev_block;
    const int *pref_base = (const int *)((uint8_t *)B.dev_block + al256(sizeof(Job) * B.n));
    if (B.persistent) {
        if (B.ptotal > 0) {
            DVG_PROF(st, "kb_persist");
            uint8_t *base = (uint8_t *)B.dev_block;
            kb_persist<<<(unsigned)B.pgrid, kBlock, 0, st>>>(jobs, B.n, (const int *)(base + B.ppref_off),
                                                             (const int *)(base + B.pbase_off),
                                                             (int *)(base + B.pstate_off), B.zero_grads, B.pbudget);
        }
    } else
    for (int c = 0; c < 3; c++) {
        int j0 = B.cls_begin[c], nj = B.cls_begin[c + 1] - j0;
        if (nj == 0) continue;
        const Job *J = jobs + j0;
        auto P = [&](int stage) { return pref_base + B.pref_off[c * kNumStages + stage]; };
        auto T = [&](int stage) { return (unsigned)B.total[c * kNumStages + stage]; };
        kb_zero<<<T(kStZero), kBlock, 0, st>>>(J, P(kStZero), nj, B.zero_grads);
        if (c == 2) {
            DVG_PROF(st, "kb_sdf");
            kb_sdf<<<T(kStSdf), kBlock, 0, st>>>(J, P(kStSdf), nj);
            continue;
        }
        {
            DVG_PROF(st, "kb_weight");
            kb_weight<<<T(kStPix), kBlock, 0, st>>>(J, P(kStPix), nj);
        }
        if (B.any_img[c]) {
            {
                DVG_PROF(st, "kb_color");
                if (c == 1)
                    kb_color<1><<<T(kStTile), kBlock, 0, st>>>(J, P(kStTile), nj);
                else
                    kb_color<0><<<T(kStTile), kBlock, 0, st>>>(J, P(kStTile), nj);
            }
            kb_resolve<<<T(kStPix), kBlock, 0, st>>>(J, P(kStPix), nj);
        }
        if (B.any_dimg[c]) {
            {
                DVG_PROF(st, "kb_color_bwd");
                if (c == 1)
                    kb_color_bwd<1><<<T(kStTile), kBlock, 0, st>>>(J, P(kStTile), nj);
                else
                    kb_color_bwd<0><<<T(kStTile), kBlock, 0, st>>>(J, P(kStTile), nj);
            }
            if (c == 0) {
                {
                    DVG_PROF(st, "kb_edge_sample");
                    kb_edge_sample<<<T(kStSamp), kBlock, 0, st>>>(J, P(kStSamp), nj);
                    kb_scan<<<(unsigned)nj, 1024, 0, st>>>(jobs, j0);
                    kb_edge_bin<<<T(kStSamp), kBlock, 0, st>>>(J, P(kStSamp), nj);
                }
                {
                    DVG_PROF(st, "kb_edge");
                    if (T(kStChunk)) kb_edge<false><<<T(kStChunk), kBlock, 0, st>>>(J, P(kStChunk), nj);
                    if (T(kStChunk2)) kb_edge<true><<<T(kStChunk2), kBlock, 0, st>>>(J, P(kStChunk2), nj);
                }
            }
        }
    }
    hipError_t e = hipGetLastError();
    B.last_stream = st;
    if (prev >= 0 && prev != B.device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return set_error(std::string("batch launch: ") + hipGetErrorString(e));
    return 0;
}

int batch_status(BatchHost &B) {
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != B.device) (void)hipSetDevice(B.device);
    DVG_CHECK(hipStreamSynchronize(B.last_stream));
    // render conditions accumulate in each scene's word until read: read and clear them,
    // as dvg_scene_status does (a forward batch's findings are still there when its
    // backward batch has run)
    if (B.persistent) {
        int abort_word = 0;
        int *state = (int *)((uint8_t *)B.dev_block + B.pstate_off);
        DVG_CHECK(hipMemcpyAsync(&abort_word, state + kPAbort, sizeof(int), hipMemcpyDeviceToHost, B.last_stream));
        DVG_CHECK(hipStreamSynchronize(B.last_stream));
        if (abort_word) {
            // the launch's last workgroup reset the task counters; clear the flag too
            DVG_CHECK(hipMemsetAsync(state + kPAbort, 0, sizeof(int), B.last_stream));
            DVG_CHECK(hipStreamSynchronize(B.last_stream));
            if (prev >= 0 && prev != B.device) (void)hipSetDevice(prev);
            return set_error("persistent batch: a task waited longer than its time limit for the tasks it depends "
                             "on; the batch's results are not defined");
        }
    }
    int flags[2] = {0, 0};
    for (auto *sc : B.scenes) {
        int h[2] = {0, 0};
        DVG_CHECK(hipMemcpyAsync(h, sc->err, 2 * sizeof(int), hipMemcpyDeviceToHost, B.last_stream));
        DVG_CHECK(hipMemsetAsync(sc->err, 0, sizeof(int), B.last_stream));
        DVG_CHECK(hipStreamSynchronize(B.last_stream));
        flags[0] |= h[0];
        flags[1] |= h[1];
    }
    if (prev >= 0 && prev != B.device) (void)hipSetDevice(prev);
    return report_flags(flags);
}

const std::vector<SceneHost *> &batch_scenes(const BatchHost &b) { return b.scenes; }

void batch_destroy(BatchHost *B) {
    if (!B) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(B->device);
    (void)hipStreamSynchronize(B->last_stream);
    if (B->dev_block) cache_free(B->device, B->dev_block, B->dev_granted);
    if (B->ws) cache_free(B->device, B->ws, B->ws_granted);
    for (SceneHost *sc : B->keep_holds) sc->keep_refs--;
    if (prev >= 0) (void)hipSetDevice(prev);
    delete B;
}

// ------------------------------------------------------------------ device scene rebuild
// dvg_scene_update: new parameter values for an unchanged topology.  The arithmetic is
// dvg_build.h's, which the host builder uses too, and the leaf orders come from the
// reference's std::sort restated in dvg_sort.h, so every table equals the one
// dvg_scene_create derives from the same values.
struct RefitArgs {
    DevScene S;
    RefitTables R;
    int *uerr;
};

__global__ void k_refit_clear(int *uerr) {
    if (threadIdx.x == 0) {
        uerr[0] = 0;
        uerr[1] += 1;  // the parameter version (err[2]): kept fragment lists go stale
    }
}

template <class T>
DVG_DEV T *writable(const T *p) {
    return const_cast<T *>(p);
}
DVG_DEV bld::Box to_box(float4 b) { return bld::Box{b.x, b.y, b.z, b.w}; }
DVG_DEV float4 to_f4(const bld::Box &b) { return make_float4(b.minx, b.miny, b.maxx, b.maxy); }

// Per shape (scene.cpp:113-333, 496-619): local box, boundary length; for paths the
// segment boxes, radii, length CDF/PMF and the segment BVH's leaf visit order.
//
// One wave per shape.  Every value is the one the host build computes: boxes are folds
// whose result does not depend on the order of the points (the first point attaining
// each extreme, as the sequential fold keeps it), the length sum and the CDF are
// sequential (one lane, from LDS), and the leaf order is the sorted order of the
// segments' y-centres -- by rank when the keys are distinct, else by the same introsort
// on one lane.  Paths with more than kRefitSegCap segments get a workgroup each
// (k_refit_long).
constexpr int kRefitWaves = kBlock / 64;
// a leaf's sort key beside its index: the introsort then reads keys without an
// indirection, and yields the permutation it yields for indices with a key comparator
template <class K>
struct KeyOf {
    K k;
    int i;
};
using KeyIdx = KeyOf<float>;
struct KeyLess {
    template <class T>
    DVG_DEV bool operator()(const T &a, const T &b) const { return a.k < b.k; }
};
constexpr int kSortStack = 96;  // >= 3 (2 floor_log2(n) + 2) for n <= 4096

// where the introsort's final insertion pass puts element i of a[0..n): the stable
// rank, for keys without NaN
template <class K>
DVG_DEV int stable_rank(const KeyOf<K> *a, int n, int i) {
    const K ki = a[i].k;
    int pos = 0;
#pragma unroll 8
    for (int j = 0; j < n; j++) {
        const K kj = a[j].k;
        pos += (kj < ki || (j < i && !(ki < kj))) ? 1 : 0;
    }
    return pos;
}

// (v, i) <- the first (lowest index) of the extreme values in the wave, as a sequential
// fold `cur = p < cur ? p : cur` (MIN) or `cur = p > cur ? p : cur` keeps it
template <bool MIN>
DVG_DEV void wave_first_extreme(float &v, int &i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(v, off);
        int oi = __shfl_xor(i, off);
        bool better = MIN ? ov < v : ov > v;
        bool worse = MIN ? v < ov : v > ov;
        if (better || (!worse && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

DVG_DEV void refit_path_serial(const RefitArgs &A, int s) {
    const DevScene &S = A.S;
    const int c0 = S.seg_begin[s], nb = S.nsegs[s], p0 = S.pt_off[s], np = S.npt[s];
    const bool thick = S.has_thick[s] != 0;
    const float r = S.stroke_width[s];
    bool finite = true;
    bld::Box box{S.points[p0].x, S.points[p0].y, S.points[p0].x, S.points[p0].y};
    for (int i = 0; i < np; i++) {
        float2 q = S.points[p0 + i];
        finite = finite && isfinite(q.x) && isfinite(q.y) && (!thick || isfinite(S.thickness[p0 + i]));
        if (i > 0) box = bld::merge(box, bld::P2{q.x, q.y});
    }
    float4 *seg_box = writable(S.seg_box);
    float *seg_maxr = writable(S.seg_maxr), *seg_pmf = writable(S.seg_pmf), *seg_cdf = writable(S.seg_cdf);
    float length = 0.f;
    for (int i = 0; i < nb; i++) {
        const int seg = c0 + i, k = S.seg_kind[seg];
        const int4 ix = S.seg_idx[seg];
        const int id[4] = {ix.x, ix.y, ix.z, ix.w};
        bld::P2 p[4];
        bld::Box b = bld::box_empty();
        float rr = -1;
        for (int j = 0; j <= k + 1; j++) {
            float2 q = S.points[id[j]];
            p[j] = bld::P2{q.x, q.y};
            b = bld::merge(b, p[j]);
            float rj = thick ? S.thickness[id[j]] : r;
            rr = j == 0 ? rj : bld::mx(rr, rj);
        }
        seg_box[seg] = to_f4(b);
        seg_maxr[seg] = rr;
        float l = bld::segment_length(k, p);
        seg_pmf[seg] = l;
        length += l;
    }
    const float len = 0.f + length;
    const float inv_length = 1.f / len;
#pragma unroll 8
    for (int i = 0; i < nb; i++) {
        float dd = seg_pmf[c0 + i] * inv_length;
        seg_pmf[c0 + i] = dd;
        seg_cdf[c0 + i] = i == 0 ? dd : dd + seg_cdf[c0 + i - 1];
    }
    int *order = A.R.order + c0;
    for (int i = 0; i < nb; i++) order[i] = i;
    int stack[192];
    stl::sort(
        order, nb,
        [&](int a, int b) {
            return bld::box_ycentre(to_box(seg_box[c0 + a])) < bld::box_ycentre(to_box(seg_box[c0 + b]));
        },
        stack);
    int *seg_dfs = writable(S.seg_dfs);
    for (int i = 0; i < nb; i++) seg_dfs[c0 + i] = c0 + order[A.R.seg_perm[c0 + i]];
    A.R.thick_leaf0[s] = seg_maxr[c0 + order[0]];
    if (!finite) atomicOr(A.uerr, kUpdNonFinitePoints);
    writable(S.shape_box)[s] = to_f4(box);
    writable(S.shape_length)[s] = len;
    if (!isfinite(len)) atomicOr(A.uerr, kUpdNonFiniteLength);
}

__global__ __launch_bounds__(kBlock) void k_refit_shapes(RefitArgs A) {
    const DevScene &S = A.S;
    __shared__ KeyIdx s_kv[kRefitWaves][kRefitSegCap];
    __shared__ int s_stack[kRefitWaves][kSortStack];
    __shared__ float s_len[kRefitWaves][kRefitSegCap];
    __shared__ int s_ord[kRefitWaves][kRefitSegCap];
    __shared__ float s_maxr[kRefitWaves][kRefitSegCap];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s = blockIdx.x * kRefitWaves + wv;
    if (s >= S.ns) return;  // whole waves; nothing below synchronises the workgroup
    const int type = S.shape_type[s];
    if (type != 2) {
        if (lane == 0) {
            bld::Box box;
            float len = bld::basic_shape(type, S.params + S.param_off[s], box);
            writable(S.shape_box)[s] = to_f4(box);
            writable(S.shape_length)[s] = len;
            if (!isfinite(len)) atomicOr(A.uerr, kUpdNonFiniteLength);
        }
        return;
    }
    const int c0 = S.seg_begin[s], nb = S.nsegs[s], p0 = S.pt_off[s], np = S.npt[s];
    if (nb > kRefitSegCap) return;  // k_refit_long
    KeyIdx *kv = s_kv[wv];
    float *lens = s_len[wv], *maxr = s_maxr[wv];
    int *ord = s_ord[wv];
    const bool thick = S.has_thick[s] != 0;
    const float r = S.stroke_width[s];
    // shape box and finiteness over the points
    float x0 = __builtin_inff(), y0 = x0, x1 = -x0, y1 = -x0;
    int ix0 = INT_MAX, iy0 = INT_MAX, ix1 = INT_MAX, iy1 = INT_MAX;
    bool finite = true;
    for (int i = lane; i < np; i += 64) {
        float2 q = S.points[p0 + i];
        finite = finite && isfinite(q.x) && isfinite(q.y) && (!thick || isfinite(S.thickness[p0 + i]));
        if (ix0 == INT_MAX) {
            x0 = x1 = q.x;
            y0 = y1 = q.y;
            ix0 = iy0 = ix1 = iy1 = i;
        } else {
            if (q.x < x0) { x0 = q.x; ix0 = i; }
            if (q.y < y0) { y0 = q.y; iy0 = i; }
            if (q.x > x1) { x1 = q.x; ix1 = i; }
            if (q.y > y1) { y1 = q.y; iy1 = i; }
        }
    }
    wave_first_extreme<true>(x0, ix0);
    wave_first_extreme<true>(y0, iy0);
    wave_first_extreme<false>(x1, ix1);
    wave_first_extreme<false>(y1, iy1);
    const bool all_finite = __all(finite);
    // per segment: box, radius, length; y-centre keys for the leaf sort
    float4 *seg_box = writable(S.seg_box);
    float *seg_maxr = writable(S.seg_maxr), *seg_pmf = writable(S.seg_pmf), *seg_cdf = writable(S.seg_cdf);
    for (int i = lane; i < nb; i += 64) {
        const int seg = c0 + i, k = S.seg_kind[seg];
        const int4 ix = S.seg_idx[seg];
        const int id[4] = {ix.x, ix.y, ix.z, ix.w};
        bld::P2 p[4];
        bld::Box b = bld::box_empty();
        float rr = -1;
        for (int j = 0; j <= k + 1; j++) {
            float2 q = S.points[id[j]];
            p[j] = bld::P2{q.x, q.y};
            b = bld::merge(b, p[j]);
            float rj = thick ? S.thickness[id[j]] : r;
            rr = j == 0 ? rj : bld::mx(rr, rj);
        }
        seg_box[seg] = to_f4(b);
        seg_maxr[seg] = rr;
        maxr[i] = rr;
        lens[i] = bld::segment_length(k, p);
        kv[i] = KeyIdx{bld::box_ycentre(b), i};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // leaf order by rank when the keys are distinct numbers
    bool plain = true, number = true;
    int rank[kRefitSegCap / 64];
#pragma unroll
    for (int t = 0; t < kRefitSegCap / 64; t++) {
        const int i = lane + 64 * t;
        rank[t] = 0;
        if (i >= nb) continue;
        const float ki = kv[i].k;
        number = number && ki == ki;
        plain = plain && ki == ki;
#pragma unroll 8
        for (int j = 0; j < nb; j++) {
            const float kj = kv[j].k;
            rank[t] += kj < ki ? 1 : 0;
            plain = plain && (j == i || kj < ki || ki < kj);
        }
    }
    const bool distinct = __all(plain), numbers = __all(number);
    __builtin_amdgcn_wave_barrier();
    if (distinct) {
#pragma unroll
        for (int t = 0; t < kRefitSegCap / 64; t++) {
            const int i = lane + 64 * t;
            if (i < nb) ord[rank[t]] = i;
        }
    } else if (numbers) {
        // ties: the introsort's partition phase on one lane, its final insertion pass
        // as the stable rank on all lanes
        if (lane == 0) {
            KeyLess less;
            stl::partition_phase(kv, 0, nb, 2 * stl::floor_log2(nb), less, s_stack[wv]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < nb; i += 64) ord[stable_rank(kv, nb, i)] = kv[i].i;
    } else if (lane == 0) {
        // a NaN key (the update reports non-finite points): any permutation serves
        stl::sort(kv, nb, KeyLess{}, s_stack[wv]);
        for (int i = 0; i < nb; i++) ord[i] = kv[i].i;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int *seg_dfs = writable(S.seg_dfs);
    for (int i = lane; i < nb; i += 64) seg_dfs[c0 + i] = c0 + ord[A.R.seg_perm[c0 + i]];
    if (lane == 0) {
        // the sequential length sum, PMF and CDF of the host build
        float length = 0.f;
#pragma unroll 8
        for (int i = 0; i < nb; i++) length += lens[i];
        const float len = 0.f + length;
        const float inv_length = 1.f / len;
        float prev = 0.f;
#pragma unroll 8
        for (int i = 0; i < nb; i++) {
            float dd = lens[i] * inv_length;
            seg_pmf[c0 + i] = dd;
            prev = i == 0 ? dd : dd + prev;
            seg_cdf[c0 + i] = prev;
        }
        A.R.thick_leaf0[s] = maxr[ord[0]];
        if (!all_finite) atomicOr(A.uerr, kUpdNonFinitePoints);
        writable(S.shape_box)[s] = make_float4(x0, y0, x1, y1);
        writable(S.shape_length)[s] = len;
        if (!isfinite(len)) atomicOr(A.uerr, kUpdNonFiniteLength);
    }
}

// Paths of more than kRefitSegCap segments, one workgroup each, with the per-wave
// rebuild's arithmetic: first-extreme folds for the shape box (per wave, then across
// the waves in order), leaf order by rank when the y-centre keys are distinct (else the
// introsort on one thread), the sequential length sum and CDF on one thread from LDS.
// Paths of more than kRefitLongCap segments run on one thread from global memory.
constexpr int kRefitLongThreads = 1024;
constexpr int kRefitLongCap = 4096;

constexpr int kRefitBigRange = 192;  // ranges the workgroup partitions together
constexpr int kRefitMaxBig = kRefitLongCap / (kRefitBigRange + 1) + 1;
constexpr int kRefitMaxSmall = kRefitLongCap / (stl::kInsertionLimit + 1) + 1;

// exclusive prefix sum of one int per thread over a kRefitLongThreads workgroup;
// `total` gets the sum.  Every thread of the workgroup must call it.
DVG_DEV int block_excl_scan(int v, int &total, int *s_scan) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_scan[wv] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < kRefitLongThreads / 64; w++) {
        const int t = s_scan[w];
        base += w < wv ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

// stl::partition_by_stoppers(a, b, e, p) by the workgroup: each thread takes up to
// four consecutive positions, numbers its left stoppers by a prefix count and its right
// stoppers by a suffix count, and the first M pairs with L_k < R_k swap.  Returns the
// cut in every thread; a[] is final when it returns.  Lpos holds e - b + 1 ints, Rpos
// e - b + 2.
template <class K>
DVG_DEV int block_partition(KeyOf<K> *a, int b, int e, int p, int *Lpos, int *Rpos, int *s_scan) {
    const int tid = threadIdx.x;
    const int per = (e - b + kRefitLongThreads - 1) / kRefitLongThreads;
    const int i0 = min(e, b + tid * per), i1 = min(e, i0 + per);
    const K P = a[p].k;
    int nl = 0, nr = 0;
    for (int i = i0; i < i1; i++) {
        nl += !(a[i].k < P) ? 1 : 0;
        nr += !(P < a[i].k) ? 1 : 0;
    }
    int NL, NR;
    int kl = block_excl_scan(nl, NL, s_scan);
    const int pr = block_excl_scan(nr, NR, s_scan);
    int kr = NR - pr;  // right stoppers at positions >= i0: the index of the first one here
    for (int i = i0; i < i1; i++) {
        if (!(a[i].k < P)) Lpos[++kl] = i;
        if (!(P < a[i].k)) Rpos[kr--] = i;
    }
    if (tid == 0) Rpos[NR + 1] = p;
    __syncthreads();
    int cnt = 0;
    for (int k = tid + 1; k <= min(NL, NR + 1); k += kRefitLongThreads) cnt += Lpos[k] < Rpos[k] ? 1 : 0;
    int M;
    (void)block_excl_scan(cnt, M, s_scan);
    // the swapped pairs are disjoint: every swapped L lies below every swapped R
    for (int k = tid + 1; k <= M; k += kRefitLongThreads) {
        const KeyOf<K> x = a[Lpos[k]], y = a[Rpos[k]];
        a[Lpos[k]] = y;
        a[Rpos[k]] = x;
    }
    int cut = M + 1 <= NL ? Lpos[M + 1] : e;
    if (M >= 1 && Rpos[M] < cut) cut = Rpos[M];
    __syncthreads();
    return cut;
}

// Sorts kv[0..nb) (nb <= kRefitLongCap, in LDS) as dvg_sort.h's restated introsort does
// and writes the sorted element indices to ord (nb ints; rpos holds nb + 2).  By rank
// when the keys are distinct numbers; with ties, the introsort's own schedule: ranges
// longer than kRefitBigRange are partitioned by the workgroup one level at a time
// (block_partition), the others by the sequential partition phase on one thread each,
// and the final insertion pass -- a stable sort -- as the stable rank; with a NaN key
// (a float key of non-finite points, which the update reports) the sequential sort on
// one thread.  Every thread of the kRefitLongThreads workgroup calls it; the keys must
// be written before the call (it begins with a barrier).
template <class K>
DVG_DEV void block_sort(KeyOf<K> *kv, int nb, int *ord, int *rpos) {
    __shared__ int s_stack[kSortStack];
    __shared__ int s_scan[kRefitLongThreads / 64];
    __shared__ int4 s_big[2][kRefitMaxBig], s_small[kRefitMaxSmall];
    __shared__ int s_nbig[2], s_nsmall, s_tie;
    const int tid = threadIdx.x;
    if (tid == 0) s_tie = 0;
    __syncthreads();
    // leaf order by rank; a tie or a NaN key hands the order to the introsort
    for (int i = tid; i < nb; i += kRefitLongThreads) {
        const K ki = kv[i].k;
        bool plain = ki == ki;
        int rank = 0;
#pragma unroll 8
        for (int j = 0; j < nb; j++) {
            const K kj = kv[j].k;
            rank += kj < ki ? 1 : 0;
            plain = plain && (j == i || kj < ki || ki < kj);
        }
        if (plain) {
            ord[rank] = i;
        } else {
            atomicOr(&s_tie, ki == ki ? 1 : 2);
        }
    }
    __syncthreads();
    const int tie = s_tie;  // read after the barrier: uniform over the workgroup
    if (tie & 2) {
        // a NaN key (the update reports non-finite points): any permutation serves
        if (tid == 0) {
            stl::sort(kv, nb, KeyLess{}, s_stack);
            for (int i = 0; i < nb; i++) ord[i] = kv[i].i;
        }
        __syncthreads();
    } else if (tie) {
        // ties: the introsort's schedule, with the ranges longer than kRefitBigRange
        // partitioned by the workgroup one level at a time, the others by one thread
        // each, and the final insertion pass as the stable rank
        KeyLess less;
        auto classify = [&](int lo, int hi, int depth, int into) {
            if (hi - lo <= stl::kInsertionLimit) return;
            if (hi - lo <= kRefitBigRange || depth == 0)
                s_small[s_nsmall++] = make_int4(lo, hi, depth, 0);
            else
                s_big[into][s_nbig[into]++] = make_int4(lo, hi, depth, 0);
        };
        if (tid == 0) {
            s_nsmall = 0;
            s_nbig[0] = s_nbig[1] = 0;
            classify(0, nb, 2 * stl::floor_log2(nb), 0);
        }
        __syncthreads();
        for (int cur = 0; s_nbig[cur] > 0; cur ^= 1) {
            const int nrng = s_nbig[cur];
            if (tid == 0) s_nbig[cur ^ 1] = 0;
            for (int r = 0; r < nrng; r++) {
                const int4 R = s_big[cur][r];
                if (tid == 0) stl::median_to(kv, R.x, R.x + 1, R.x + (R.y - R.x) / 2, R.y - 1, less);
                __syncthreads();
                const int cut = block_partition(kv, R.x + 1, R.y, R.x, ord, rpos, s_scan);
                if (tid == 0) {
                    classify(R.x, cut, R.z - 1, cur ^ 1);
                    classify(cut, R.y, R.z - 1, cur ^ 1);
                }
            }
            __syncthreads();
        }
        for (int r = tid; r < s_nsmall; r += kRefitLongThreads) {
            const int4 R = s_small[r];
            int stack[kSortStack];
            stl::partition_phase(kv, R.x, R.y, R.z, less, stack);
        }
        __syncthreads();
        for (int i = tid; i < nb; i += kRefitLongThreads) ord[stable_rank(kv, nb, i)] = kv[i].i;
        __syncthreads();
    }
}

template <bool MIN>
DVG_DEV void first_extreme_combine(float &v, int &i, float ov, int oi) {
    bool better = MIN ? ov < v : ov > v;
    bool worse = MIN ? v < ov : v > ov;
    if (better || (!worse && oi < i)) {
        v = ov;
        i = oi;
    }
}

__global__ __launch_bounds__(kRefitLongThreads) void k_refit_long(RefitArgs A) {
    const DevScene &S = A.S;
    const int s = A.R.long_ids[blockIdx.x];
    const int c0 = S.seg_begin[s], nb = S.nsegs[s], p0 = S.pt_off[s], np = S.npt[s];
    const int tid = threadIdx.x, wv = tid >> 6;
    if (nb > kRefitLongCap) {  // uniform over the workgroup: no barrier is reached
        if (tid == 0) refit_path_serial(A, s);
        return;
    }
    constexpr int kW = kRefitLongThreads / 64;
    __shared__ KeyIdx kv[kRefitLongCap];
    __shared__ float lens[kRefitLongCap];
    __shared__ int ord[kRefitLongCap];  // also the left stoppers of the tie sort
    __shared__ int rpos[kRefitLongCap + 2];
    __shared__ float s_ext[4][kW];
    __shared__ int s_exti[4][kW];
    __shared__ int s_nonfinite;
    if (tid == 0) s_nonfinite = 0;
    const bool thick = S.has_thick[s] != 0;
    const float r = S.stroke_width[s];
    // shape box and finiteness over the points
    float x0 = __builtin_inff(), y0 = x0, x1 = -x0, y1 = -x0;
    int ix0 = INT_MAX, iy0 = INT_MAX, ix1 = INT_MAX, iy1 = INT_MAX;
    bool finite = true;
    for (int i = tid; i < np; i += kRefitLongThreads) {
        float2 q = S.points[p0 + i];
        finite = finite && isfinite(q.x) && isfinite(q.y) && (!thick || isfinite(S.thickness[p0 + i]));
        if (ix0 == INT_MAX) {
            x0 = x1 = q.x;
            y0 = y1 = q.y;
            ix0 = iy0 = ix1 = iy1 = i;
        } else {
            if (q.x < x0) { x0 = q.x; ix0 = i; }
            if (q.y < y0) { y0 = q.y; iy0 = i; }
            if (q.x > x1) { x1 = q.x; ix1 = i; }
            if (q.y > y1) { y1 = q.y; iy1 = i; }
        }
    }
    wave_first_extreme<true>(x0, ix0);
    wave_first_extreme<true>(y0, iy0);
    wave_first_extreme<false>(x1, ix1);
    wave_first_extreme<false>(y1, iy1);
    __syncthreads();  // s_nonfinite is set
    if ((tid & 63) == 0) {
        s_ext[0][wv] = x0, s_exti[0][wv] = ix0;
        s_ext[1][wv] = y0, s_exti[1][wv] = iy0;
        s_ext[2][wv] = x1, s_exti[2][wv] = ix1;
        s_ext[3][wv] = y1, s_exti[3][wv] = iy1;
    }
    if (!finite) s_nonfinite = 1;
    // per segment: box, radius, length; y-centre keys for the leaf sort
    float4 *seg_box = writable(S.seg_box);
    float *seg_maxr = writable(S.seg_maxr), *seg_pmf = writable(S.seg_pmf), *seg_cdf = writable(S.seg_cdf);
    for (int i = tid; i < nb; i += kRefitLongThreads) {
        const int seg = c0 + i, k = S.seg_kind[seg];
        const int4 ix = S.seg_idx[seg];
        const int id[4] = {ix.x, ix.y, ix.z, ix.w};
        bld::P2 p[4];
        bld::Box b = bld::box_empty();
        float rr = -1;
        for (int j = 0; j <= k + 1; j++) {
            float2 q = S.points[id[j]];
            p[j] = bld::P2{q.x, q.y};
            b = bld::merge(b, p[j]);
            float rj = thick ? S.thickness[id[j]] : r;
            rr = j == 0 ? rj : bld::mx(rr, rj);
        }
        seg_box[seg] = to_f4(b);
        seg_maxr[seg] = rr;
        lens[i] = bld::segment_length(k, p);
        kv[i] = KeyIdx{bld::box_ycentre(b), i};
    }
    __syncthreads();
    block_sort(kv, nb, ord, rpos);
    int *seg_dfs = writable(S.seg_dfs);
    for (int i = tid; i < nb; i += kRefitLongThreads) seg_dfs[c0 + i] = c0 + ord[A.R.seg_perm[c0 + i]];
    if (tid == 0) {
        // the sequential length sum, PMF and CDF of the host build
        float length = 0.f;
#pragma unroll 8
        for (int i = 0; i < nb; i++) length += lens[i];
        const float len = 0.f + length;
        const float inv_length = 1.f / len;
        float prev = 0.f;
#pragma unroll 8
        for (int i = 0; i < nb; i++) {
            float dd = lens[i] * inv_length;
            seg_pmf[c0 + i] = dd;
            prev = i == 0 ? dd : dd + prev;
            seg_cdf[c0 + i] = prev;
        }
        A.R.thick_leaf0[s] = seg_maxr[c0 + ord[0]];
        float e[4];
        int ei[4];
        for (int c = 0; c < 4; c++) {
            e[c] = s_ext[c][0];
            ei[c] = s_exti[c][0];
        }
        for (int w = 1; w < kW; w++) {
            first_extreme_combine<true>(e[0], ei[0], s_ext[0][w], s_exti[0][w]);
            first_extreme_combine<true>(e[1], ei[1], s_ext[1][w], s_exti[1][w]);
            first_extreme_combine<false>(e[2], ei[2], s_ext[2][w], s_exti[2][w]);
            first_extreme_combine<false>(e[3], ei[3], s_ext[3][w], s_exti[3][w]);
        }
        if (s_nonfinite) atomicOr(A.uerr, kUpdNonFinitePoints);
        writable(S.shape_box)[s] = make_float4(e[0], e[1], e[2], e[3]);
        writable(S.shape_length)[s] = len;
        if (!isfinite(len)) atomicOr(A.uerr, kUpdNonFiniteLength);
    }
}

// The member visit order of group g's BVH (scene.cpp:632-645): the members sorted by
// the Morton code of their box centre on one thread, from global memory.
DVG_DEV void group_order_serial(const RefitArgs &A, int g) {
    const DevScene &S = A.S;
    const int off = S.g_off[g], n = S.g_n[g];
    const int *desc = A.R.g_ids_desc + off;
    int *ord = A.R.gorder + off;
    for (int i = 0; i < n; i++) ord[i] = i;
    int stack[192];
    stl::sort(
        ord, n,
        [&](int a, int b) {
            return bld::morton_of_box(to_box(S.shape_box[desc[a]]), S.cw, S.ch) <
                   bld::morton_of_box(to_box(S.shape_box[desc[b]]), S.cw, S.ch);
        },
        stack);
    int *g_ids = writable(S.g_ids);
    for (int i = 0; i < n; i++) g_ids[off + i] = desc[ord[A.R.g_perm[off + i]]];
}

// The member order of a group of more than kRefitGroupLane shapes, one workgroup per
// group: the Morton codes in LDS, sorted by block_sort (groups of more than
// kRefitLongCap members on one thread).
__global__ __launch_bounds__(kRefitLongThreads) void k_refit_big_groups(RefitArgs A) {
    const DevScene &S = A.S;
    const int g = A.R.big_gids[blockIdx.x];
    const int off = S.g_off[g], n = S.g_n[g];
    if (n > kRefitLongCap) {  // uniform over the workgroup: no barrier is reached
        if (threadIdx.x == 0) group_order_serial(A, g);
        return;
    }
    __shared__ KeyOf<uint32_t> kv[kRefitLongCap];
    __shared__ int ord[kRefitLongCap];
    __shared__ int rpos[kRefitLongCap + 2];
    const int *desc = A.R.g_ids_desc + off;
    for (int i = threadIdx.x; i < n; i += kRefitLongThreads)
        kv[i] = KeyOf<uint32_t>{bld::morton_of_box(to_box(S.shape_box[desc[i]]), S.cw, S.ch), i};
    block_sort(kv, n, ord, rpos);
    int *g_ids = writable(S.g_ids);
    for (int i = threadIdx.x; i < n; i += kRefitLongThreads) g_ids[off + i] = desc[ord[A.R.g_perm[off + i]]];
}

// Per group (scene.cpp:621-684, shape.h:123): canvas_to_shape, member visit order of the
// Morton-sorted group BVH, local root box, canvas-space leaf box and radius, SDF anchor.
__global__ __launch_bounds__(kBlock) void k_refit_groups(RefitArgs A) {
    const DevScene &S = A.S;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.ng) return;
    const float *m = S.s2c + 9 * (size_t)g;
    bld::inverse3(m, writable(S.c2s) + 9 * (size_t)g);
    const int off = S.g_off[g], n = S.g_n[g];
    const int *desc = A.R.g_ids_desc + off;
    if (n <= kRefitGroupLane) group_order_serial(A, g);  // larger groups: k_refit_big_groups
    bld::Box lb = bld::box_empty();
    float maxr = 0.f;
    for (int i = 0; i < n; i++) {
        const int sh = desc[i];
        lb = bld::merge(lb, to_box(S.shape_box[sh]));
        float r = (S.shape_type[sh] == 2 && S.has_thick[sh]) ? A.R.thick_leaf0[sh] : S.stroke_width[sh];
        maxr = i == 0 ? r : (maxr < r ? r : maxr);
    }
    writable(S.g_lbox)[g] = to_f4(lb);
    writable(S.g_box)[g] = to_f4(bld::canvas_box(m, lb));
    writable(S.g_maxr)[g] = S.stroke_type[g] < 0 ? 0.f : maxr;
    const int flags = A.R.g_sdf_flags[g], anchor = A.R.g_anchor[g];
    const float fz = (flags & bld::kSdfFindable) ? 1.f : 0.f;
    float4 sd = make_float4(__builtin_nanf(""), __builtin_nanf(""), fz, 0.f);
    if ((flags & bld::kSdfFindable) && anchor >= 0 && bld::similarity(m) && !(flags & bld::kSdfApproxQuad)) {
        float2 q0 = S.points[S.pt_off[anchor]];
        bld::P2 q = bld::xform(m, bld::P2{q0.x, q0.y});
        sd = make_float4(q.x, q.y, fz, 0.f);
    }
    writable(S.g_sdf)[g] = sd;
}

// build_shape_cdfs (scene.cpp:207-246): the running sum is sequential, as in the
// reference, so the CDF has its rounding.  The workgroup gathers the lengths into LDS
// in chunks; one lane adds them in order; then every entry is normalised in parallel.
__global__ __launch_bounds__(kBlock) void k_refit_cdf(RefitArgs A) {
    const DevScene &S = A.S;
    const int total = S.num_total_shapes;
    float *cdf = writable(S.shapes_cdf), *pmf = writable(S.shapes_pmf);
    constexpr int kChunk = 8192;
    __shared__ float s_len[kChunk];
    __shared__ float s_acc;
    for (int base = 0; base < total; base += kChunk) {
        const int n = min(kChunk, total - base);
        for (int k = threadIdx.x; k < n; k += blockDim.x) s_len[k] = S.shape_length[S.sample_shape_id[base + k]];
        __syncthreads();
        if (threadIdx.x == 0) {
            float acc = base == 0 ? 0.f : s_acc;
            for (int k = 0; k < n; k++) {
                float l = s_len[k];
                acc = base + k == 0 ? l : l + acc;
                s_len[k] = acc;
            }
            s_acc = acc;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) cdf[base + k] = s_len[k];
        __syncthreads();
    }
    const float norm = total > 0 ? s_acc : 0.f;
    if (threadIdx.x == 0 && (!(norm > 0.f) || !isfinite(norm))) atomicOr(A.uerr, kUpdTotalLength);
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        cdf[k] = cdf[k] / norm;
        pmf[k] = S.shape_length[S.sample_shape_id[k]] / norm;
    }
}

// The segment-only parts of the fast cubic tests, one lane per segment, from the scene's
// current points: wf::cubic_prep in float4 0-2, wf::stroke_prep in float4 3-7.
__global__ __launch_bounds__(kBlock) void k_seg_prep(DevScene S) {
    const int seg = blockIdx.x * blockDim.x + threadIdx.x;
    if (seg >= S.nseg) return;
    float4 *dst = const_cast<float4 *>(S.seg_wf) + kSegPrep * (size_t)seg;
    if (S.seg_kind[seg] != 2) {
        for (int i = 0; i < kSegPrep; i++) dst[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const int4 idx = S.seg_idx[seg];
    const float2 a = S.points[idx.x], b = S.points[idx.y], c = S.points[idx.z], d = S.points[idx.w];
    const wf::CubicPrep q = wf::cubic_prep(a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y);
    dst[0] = make_float4(q.my, q.mx, q.E, q.ddy);
    dst[1] = make_float4(q.ddx, q.ddmax, q.xdmax, q.br1);
    dst[2] = make_float4(q.br2, 0.f, 0.f, 0.f);
    float poly[16], mc, dev;
    wf::stroke_prep(a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y, poly, mc, dev);
    for (int i = 0; i < 4; i++) dst[3 + i] = make_float4(poly[4 * i], poly[4 * i + 1], poly[4 * i + 2], poly[4 * i + 3]);
    dst[7] = make_float4(mc, dev, 0.f, 0.f);
}

int prep_segments(SceneHost &sc, hipStream_t st) {
    const DevScene &D = sc.dev;
    if (D.nseg <= 0) return 0;
    k_seg_prep<<<(D.nseg + kBlock - 1) / kBlock, kBlock, 0, st>>>(D);
    DVG_CHECK(hipGetLastError());
    if (!st) DVG_CHECK(hipStreamSynchronize(nullptr));
    return 0;
}

int update_scene(SceneHost &sc, const dvg_scene_params &p, float filter_radius, hipStream_t st) {
    sc.version++;
    sc.frag_valid = false;
    sc.pkeep_valid = false;
    if (!(filter_radius > 0) || !std::isfinite(filter_radius)) return set_error("filter radius must be positive");
    sc.dev.filter_radius = filter_radius;
    if (!sc.device_set || sc.num_shapes == 0) return 0;
    DevScene &D = sc.dev;
    auto copy = [&](const void *dst, const float *src, size_t floats) -> int {
        if (src && floats)
            DVG_CHECK(hipMemcpyAsync(const_cast<void *>(dst), src, floats * sizeof(float), hipMemcpyDeviceToDevice, st));
        return 0;
    };
    if (copy(D.params, p.shape_params, sc.num_shape_params) || copy(D.points, p.points, 2 * (size_t)D.npts) ||
        copy(D.thickness, p.thickness, D.npts) || copy(D.stroke_width, p.stroke_width, D.ns) ||
        copy(D.color, p.color_params, sc.num_color_params) || copy(D.s2c, p.shape_to_canvas, 9 * (size_t)D.ng))
        return 1;
    RefitArgs A{D, sc.refit, sc.err + 1};
    {
        DVG_PROF(st, "k_refit");
        k_refit_clear<<<1, 64, 0, st>>>(A.uerr);
        k_refit_shapes<<<(D.ns + kRefitWaves - 1) / kRefitWaves, kBlock, 0, st>>>(A);
        if (A.R.n_long > 0) k_refit_long<<<A.R.n_long, kRefitLongThreads, 0, st>>>(A);
        if (D.ng > 0) k_refit_groups<<<(D.ng + kBlock - 1) / kBlock, kBlock, 0, st>>>(A);
        if (A.R.n_big > 0) k_refit_big_groups<<<A.R.n_big, kRefitLongThreads, 0, st>>>(A);
        k_refit_cdf<<<1, kBlock, 0, st>>>(A);
        if (D.nseg > 0) k_seg_prep<<<(D.nseg + kBlock - 1) / kBlock, kBlock, 0, st>>>(D);
    }
    DVG_CHECK(hipGetLastError());
    sc.top_stale = true;
    sc.last_stream = st;
    return 0;
}

}  // namespace dvg

extern "C" {

int dvg_profile_enable(int on) {
    dvg::prof::enabled = on != 0;
    return 0;
}

// Waits for the recorded events and writes per-kernel totals: names as a
// newline-separated list into `names` (capacity `cap` bytes), total milliseconds and
// launch counts into ms[] / counts[] (capacity `max_kernels`).  Clears the records.
int dvg_profile_read(char *names, int cap, double *ms, int *counts, int max_kernels) {
    std::lock_guard<std::mutex> lk(dvg::prof::mu);
    std::map<std::string, std::pair<double, int>> acc;
    for (auto &r : dvg::prof::recs) {
        float t = 0.f;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&t, r.a, r.b);
        auto &e = acc[r.name];
        e.first += t;
        e.second += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    dvg::prof::recs.clear();
    std::string all;
    int n = 0;
    for (auto &kv : acc) {
        if (n >= max_kernels) break;
        all += kv.first + "\n";
        ms[n] = kv.second.first;
        counts[n] = kv.second.second;
        n++;
    }
    if (names && cap > 0) {
        size_t k = std::min((size_t)cap - 1, all.size());
        std::memcpy(names, all.data(), k);
        names[k] = 0;
    }
    return n;
}

}  // extern "C"
