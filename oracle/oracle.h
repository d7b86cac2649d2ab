/*
 * oracle.h — CPU oracle for the particle-sphere ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke check and
 * the CPU-baseline leg of the benchmark call it. The product (libmmhrt) never links,
 * loads or calls it.
 *
 * Every function restates a piece of the reference (MegaMol optix_hpg + mmospray path)
 * and cites the file:line it follows. Arithmetic is plain FP32 in the order the
 * reference writes it, compiled with -ffp-contract=off so that no multiply-add is fused.
 */
#ifndef MMHRT_ORACLE_H
#define MMHRT_ORACLE_H
#include <stdint.h>

typedef struct orc_camera {
    float position[3], direction[3], up[3];
    float fovy, aspect, near_plane, far_plane;
    float tile_start[2], tile_end[2];
} orc_camera;

typedef struct orc_frame_state {
    float up[3], right[3], front[3], center[3];
    float th, rw, nearp;
    float depth_a, depth_b, depth_c; /* proj[2][2], proj[2][3], proj[3][2] */
} orc_frame_state;

typedef struct orc_params {
    uint32_t width, height, spp;
    int32_t mode; /* 0 lambert, 1 ao */
    float ao_distance;
    uint32_t frame_idx;
    float background[4];
    int32_t want_aux;
    uint32_t row_begin, row_end; /* LBVH renderer only: rows [begin, end) (end 0 = all rows) */
    int32_t max_bounces;         /* mode 2: diffuse bounces of raygen.cu:144-154 */
} orc_params;

typedef struct orc_list {
    int32_t vert_type, col_type;
    const void* vert;
    const void* col;
    uint64_t vert_stride, col_stride, count;
    float global_radius;
    uint8_t global_rgba[4];
    float col_min, col_max;
} orc_list;

typedef struct orc_bvh {
    uint64_t n;          /* spheres */
    float* node_box;     /* (n-1) * 12: child0 lo xyz hi xyz, child1 lo xyz hi xyz */
    int32_t* node_child; /* (n-1) * 2: >=0 internal index, <0 leaf ~k */
    uint32_t* leaf_id;   /* n: particle index of leaf k */
    uint64_t* keys;      /* n: Morton key of leaf k */
    float* leaf_sph;     /* n * 4: spheres in leaf order */
    const float* xyzr_in; /* the spheres the build read, in particle order (not owned: the caller keeps them) */
    float amax;           /* max over spheres of |coordinate| + radius (the slab padding's scale) */
} orc_bvh;

/* Treelets of P-k-d trees (plugins/optix_owl/src/TreeletsCreate.h:117-166, PKDCreate.h:86-160):
 * the Morton-sorted spheres are cut into treelets of M consecutive spheres (the last one shorter),
 * each stored as a left-balanced k-d tree over its spheres in heap order (node i of treelet t at
 * position t*M + i, children 2i+1 and 2i+2), with a complete binary tree of boxes over the
 * treelets on top (top nodes 0 .. top_leaves-2 in heap order; child c >= top_leaves-1 is treelet
 * c - (top_leaves-1), absent when >= T). */
typedef struct orc_pkd {
    uint64_t n;          /* spheres */
    uint32_t M, T;       /* treelet size (power of two), treelet count */
    uint32_t top_leaves; /* leaves of the top tree: the smallest power of two >= max(T, 2) */
    float* sph;          /* n * 4: spheres in treelet heap order */
    uint32_t* id;        /* n: particle index */
    uint8_t* dim;        /* n: split dimension of the node */
    float* tl_box;       /* T * 6: sphere bounds of each treelet (lo xyz, hi xyz) */
    float* tl_rmax;      /* T: largest radius in each treelet */
    float* top_box;      /* (top_leaves-1) * 12: child0 lo xyz hi xyz, child1 lo xyz hi xyz */
    int32_t* top_child;  /* (top_leaves-1) * 2: >= 0 top node, < 0 treelet ~t, INT32_MIN absent */
    const float* xyzr_in; /* as orc_bvh */
    float amax;
} orc_pkd;

#ifdef __cplusplus
extern "C" {
#endif
void orc_tea_encrypt(uint32_t v[2], const uint32_t k[4], int cycles);
uint32_t orc_tea16(uint32_t v0, uint32_t v1);
void orc_ao_direction(const float n[3], uint32_t seed, uint32_t smp, uint32_t spp, float dir[3]);
uint32_t orc_lcg(uint32_t* state);
float orc_rnd(uint32_t* state);
void orc_frame_state_from_camera(const orc_camera* cam, orc_frame_state* fs);
void orc_primary_ray(const orc_frame_state* fs, uint32_t w, uint32_t h, uint32_t x, uint32_t y, float rx, float ry,
                     float o[3], float d[3]);
int orc_intersect_f32(const float o[3], const float d[3], float tmin, float tmax, const float sph[4], float* t_out);
int orc_intersect_f64(const double o[3], const double d[3], double tmin, const double sph[4], double* t_out,
                      double* disc_out);

int orc_convert_lists(const orc_list* lists, uint32_t n_lists, const float* tf, uint32_t tf_n, float tf_min,
                      float tf_max, int tf_auto_range, float* xyzr_out, float* rgba_out, uint64_t* n_out);
void orc_tf_sample(const float* tf, uint32_t tf_n, float vmin, float vmax, float value, float out[4]);

void orc_blend_frames(const float* old_col, const float* new_col, const int32_t* hit_id, uint64_t n_px,
                      uint32_t frame_idx, const float bg[4], float* out_col, uint32_t* rgba8);
void orc_pack_rgba8(const float* col, uint64_t n_px, uint32_t* out);
int orc_render_brute(const float* xyzr, const float* rgba, uint64_t n, const orc_camera* cam, const orc_params* p,
                     uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux);
int orc_primary_f64(const float* xyzr, uint64_t n, const orc_camera* cam, const orc_params* p, uint32_t x0,
                    uint32_t y0, uint32_t x1, uint32_t y1, int32_t* id_out, double* t_out, double* disc_out,
                    double* t2_out);
int orc_primary_f64_rects(const float* xyzr, uint64_t n, const orc_camera* cam, const orc_params* p,
                          const uint32_t* rects, uint32_t k, int threads, int32_t* id_out, double* t_out,
                          double* disc_out, double* t2_out);

int orc_bvh_build(const float* xyzr, uint64_t n, int bits, orc_bvh* out);
int orc_bvh_build_mt(const float* xyzr, uint64_t n, int bits, int threads, orc_bvh* out);
void orc_bvh_free(orc_bvh* b);
int orc_render_bvh(const orc_bvh* b, const float* rgba, const orc_camera* cam, const orc_params* p, int threads,
                   uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux,
                   uint64_t* counters);
int orc_bvh_copy_out(const orc_bvh* b, float* node_box, int32_t* node_child, uint32_t* leaf_id, uint64_t* keys);

/* Treelet P-k-d trees over the Morton order of an LBVH build (its leaf_sph / leaf_id). */
int orc_pkd_build(const orc_bvh* b, uint32_t M, int threads, orc_pkd* out);
void orc_pkd_free(orc_pkd* k);
int orc_render_pkd(const orc_pkd* k, const float* rgba, const orc_camera* cam, const orc_params* p, int threads,
                   uint32_t* rgba8, float* depth, float* color, float* t_out, int32_t* id_out, uint16_t* aux,
                   uint64_t* counters);
uint32_t orc_pkd_subtree_size(uint32_t i, uint32_t m);

double orc_now(void);
#ifdef __cplusplus
}
#endif
#endif
