/*
 * oracle.h -- CPU restatement of diffvg's render path (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke() check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and only
 * as the checker / CPU baseline.  The product (diffvg_amd) never links or calls it.
 *
 * The oracle restates, function by function, the reference algorithm in
 * diffvg.cpp, scene.cpp, winding_number.h, within_distance.h, compute_distance.h,
 * sample_boundary.h, solve.h, filter.h, pcg.h, cdf.h, matrix.h and vector.h,
 * including the reference's float/double promotion (its unqualified libm calls on
 * float arguments resolve to the double overloads) and its quirks.  Each function in
 * oracle.cpp cites the reference file:line it follows.
 *
 * Scene input is a flat structure-of-arrays descriptor with the same field layout
 * as dvg_scene_desc in include/diffvg_amd.h; all pointers are host pointers.
 */
#ifndef DIFFVG_ORACLE_H
#define DIFFVG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene_desc {
    int32_t canvas_width, canvas_height;
    int32_t num_shapes;
    const int32_t *shape_type;          /* 0 circle, 1 ellipse, 2 path, 3 rect */
    const float   *stroke_width;
    const int32_t *shape_param_offset;  /* into shape_params, -1 for paths */
    const float   *shape_params;        /* circle r,cx,cy | ellipse rx,ry,cx,cy | rect x0,y0,x1,y1 */
    int32_t num_shape_params;
    const int32_t *path_ctrl_offset;
    const int32_t *path_num_base_points;
    const int32_t *path_point_offset;
    const int32_t *path_num_points;
    const uint8_t *path_has_thickness;
    const uint8_t *path_is_closed;
    const uint8_t *path_use_distance_approx;
    const int32_t *num_control_points;
    int32_t num_ctrl_total;
    const float   *points;
    const float   *thickness;
    int32_t num_points_total;
    int32_t num_groups;
    const int32_t *group_shape_offset;
    const int32_t *group_num_shapes;
    const int32_t *group_shape_ids;
    int32_t num_group_shape_ids;
    const int32_t *fill_type;           /* -1 none, 0 constant, 1 linear, 2 radial */
    const int32_t *fill_offset;
    const int32_t *fill_num_stops;
    const int32_t *stroke_type;
    const int32_t *stroke_offset;
    const int32_t *stroke_num_stops;
    const float   *color_params;
    int32_t num_color_params;
    const uint8_t *use_even_odd_rule;
    const float   *shape_to_canvas;     /* num_groups * 9, row-major */
    int32_t filter_type;                /* 0 box, 1 tent, 2 parabolic, 3 hann, 4 gaussian */
    float filter_radius;
} orc_scene_desc;

typedef struct orc_render_args {
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
    void *reserved;
} orc_render_args;

typedef struct orc_grads {
    float *d_shape_params;
    float *d_points;
    float *d_thickness;
    float *d_stroke_width;
    float *d_color_params;
    float *d_shape_to_canvas;
    float *d_filter_radius;
} orc_grads;

/* Query counters of the last render (for the algorithmic-work cost model). */
typedef struct orc_counters {
    int64_t samples;
    int64_t top_box_tests;
    int64_t group_box_tests;
    int64_t seg_box_tests;
    int64_t line_tests;
    int64_t quad_tests;
    int64_t cubic_winding_tests;
    int64_t cubic_distance_tests;
    int64_t newton_iterations;
    int64_t closest_point_tests;
    int64_t fragments;
    int64_t boundary_samples;
} orc_counters;

typedef struct orc_scene orc_scene;

int orc_scene_create(const orc_scene_desc *desc, orc_scene **out);
int orc_scene_destroy(orc_scene *scene);
int orc_render(orc_scene *scene, const orc_render_args *args);
int orc_scene_get_grads(const orc_scene *scene, orc_grads *out);
int orc_scene_zero_grads(orc_scene *scene);
const char *orc_last_error(void);
void orc_set_num_threads(int n);
int orc_get_num_threads(void);
void orc_set_counting(int on);
/* Counters of the last render; phase 0 = weight and render kernels, 1 = boundary
 * sampling and edge kernel. */
void orc_get_counters(int phase, orc_counters *out);
/* While counting: samples of the last render whose distance queries reached the
 * isolating-interval slot that the reference reads uninitialised when the isolator
 * root -B/5 lies outside [0, 1] (within_distance.h:195-202, compute_distance.h:200-207).
 * phase 0 = interior samples, 1 = edge samples that contributed a gradient.  The
 * reference's result for such a sample depends on whatever that stack slot holds; the
 * oracle treats the slot as absent. */
int64_t orc_slot_samples(int phase);
/* The phase-1 tally split by the shape each edge sample was drawn on; returns the
 * number of entries written (min(n, number of shapes)). */
int orc_slot_edge_samples_by_shape(int64_t *out, int n);
/* Evaluate only every stride-th sample in each pass (for cost counting on a sample;
 * the images and gradients are then incomplete). */
void orc_set_sample_stride(int stride);

/* BVH inspection: level 0 = top level (groups), 1 = group `index`, 2 = path `index`.
 * Writes up to max_nodes nodes as (child0, child1, minx, miny, maxx, maxy, max_radius);
 * returns the node count or -1. */
int orc_scene_bvh(const orc_scene *scene, int level, int index,
                  int32_t *children, float *boxes, float *max_radius, int max_nodes);
/* Edge-sampling tables: per path shape the length CDF/PMF; global shape CDF/PMF. */
int orc_scene_sampling_tables(const orc_scene *scene, float *shape_length,
                              float *shapes_cdf, float *shapes_pmf,
                              int32_t *sample_shape_id, int32_t *sample_group_id);

/* Unit-level entry points used to pin individual functions. */
uint32_t orc_pcg32_sample(int idx, uint64_t seed, float *out2);
int orc_solve_quadratic_f(float a, float b, float c, float *t0, float *t1);
int orc_solve_cubic_f(float a, float b, float c, float d, float *t);
int orc_solve_cubic_d(double a, double b, double c, double d, double *t);
float orc_filter_weight(int type, float radius, float dx, float dy);
float orc_d_filter_weight(int type, float radius, float dx, float dy, float d_return);
int orc_winding_number(const orc_scene *scene, int group_id, float x, float y);
int orc_is_inside(const orc_scene *scene, int group_id, float x, float y);
int orc_within_distance(const orc_scene *scene, int group_id, float x, float y);
int orc_compute_distance(const orc_scene *scene, int group_id, float x, float y,
                         float max_radius, float *dist, float *closest_xy,
                         int32_t *shape_id, int32_t *path_info, float *t_root);
void orc_sample_color(const orc_scene *scene, float sx, float sy, float *rgba);
int orc_sample_boundary(const orc_scene *scene, int idx, uint64_t seed, float *out);
uint32_t orc_morton2d(float x, float y, int w, int h);

#ifdef __cplusplus
}
#endif

#endif
