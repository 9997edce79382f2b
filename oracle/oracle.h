/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's triangle SBVH build and traversal, used by the
 * tests to check the HIP path and by bench.py as the "port" CPU baseline.
 */
#ifndef SPROUT_ORACLE_H
#define SPROUT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_ERR_ARGS        = -1,
    ORACLE_ERR_UNSPLITTABLE = -3,
    ORACLE_ERR_STACK       = -4,
};

/* scene::bvh::Node, core/scene/bvh/scene_bvh_node.hpp:44-59 */
typedef struct {
    float    lo[3];
    uint32_t data; /* first child (inner) or first slot (leaf) */
    float    hi[3];
    uint8_t  axis;
    uint8_t  num; /* 0 = inner node, else number of slots */
    uint8_t  pad[2];
} onode;

typedef struct {
    onode*    nodes;
    uint32_t* slots; /* input triangle id per tree slot */
    uint32_t  num_nodes;
    uint32_t  num_slots;
    uint32_t  spatial_split_threshold;
} oracle_bvh;

/* reciprocal modes for the traversal */
enum {
    ORACLE_RCP_PINNED = 0, /* rcp(x) := correctly rounded 1/x, then the reference's Newton step */
    ORACLE_RCP_NATIVE = 1, /* rcp(x) := this host's rcpss/rcpps, then the reference's Newton step */
};

int32_t oracle_sbvh_build(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                          uint32_t num_triangles, uint32_t num_slices, uint32_t sweep_threshold,
                          uint32_t max_primitives, oracle_bvh* out);

void oracle_bvh_free(oracle_bvh* bvh);

float oracle_surface_area(float const lo[3], float const hi[3]);

/*
 * rays: 8 floats per ray {o.xyz, min_t, d.xyz, max_t}.
 * any == 0: hits = 4 words per ray {t, u, v, input triangle id or 0xFFFFFFFF}
 * any != 0: hits = 1 word per ray (1 occluded / 0 free)
 * counts (optional): 2 words per ray {nodes fetched, triangle tests}, counted
 * as the reference's loop performs them.
 * Returns 0, or ORACLE_ERR_STACK if a ray needed more than 127 stack entries.
 */
int32_t oracle_trace(onode const* nodes, uint32_t const* slots, float const* positions,
                     uint32_t const* indices, float const* rays, uint64_t num_rays, int32_t any,
                     int32_t rcp_mode, uint32_t* hits, uint32_t* counts, uint32_t num_threads);

/* Closest hit by testing every triangle in input order with the same
 * primitive test and tie rule (a later equal-t triangle wins). */
void oracle_brute_force(float const* positions, uint32_t const* indices, uint32_t num_triangles,
                        float const* rays, uint64_t num_rays, int32_t rcp_mode, uint32_t* hits);

/* The reciprocals as defined above, exposed for tests. */
float oracle_reciprocal(float x, int32_t rcp_mode);
void  oracle_reciprocal3(float const d[3], int32_t rcp_mode, float out[3]);

/* offset_ray, core/scene/ray_offset.inl:18-29 */
void oracle_offset_ray(float const p[3], float const n[3], float out[3]);

#ifdef __cplusplus
}
#endif

#endif
