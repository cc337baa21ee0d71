/*
 * pulse_scene.h -- host-side producer of the scene buffer layout.
 *
 * The reference builds its BVHs on the CPU every time the scene changes and
 * keeps doing so here: these entry points restate its builders in C++ (there
 * is no Rust toolchain to run the originals) and emit byte-identical node
 * arrays for the same inputs:
 *   pulse_build_blas   <- build_blas, src/scene/blas.rs:44-87 (+ subdivide
 *                         89-161, find_best_split_plane 170-253,
 *                         calculate_node_aabb 296-314)
 *   pulse_build_tlas   <- build_tlas, src/scene/tlas.rs:29-136
 *   pulse_sb_*         <- prepare_mesh_data (src/scene/mod.rs:471-510),
 *                         prepare_mesh_instances (561-697), light CDFs
 *                         (700-742) and the buffer list of
 *                         queue_scene_bind_group (911-1090)
 * Meshes and materials are concatenated in insertion order (the reference
 * iterates a HashMap, whose order is unspecified).
 */
#ifndef PULSE_SCENE_H
#define PULSE_SCENE_H

#include "pulse_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PulsePrimitiveMeshInstance, src/scene/mod.rs:535-539: world AABB of an
 * instance and its centre, the TLAS builder's input. */
typedef struct pulse_instance_bounds {
    float bounds_min[3];
    float bounds_max[3];
    float center[3];
} pulse_instance_bounds;

/* Binned-SAH BLAS over `n` primitives.  nodes_out must hold 2*n-1 nodes
 * (n >= 1); tri_indices_out holds n entries. */
int pulse_build_blas(const pulse_primitive *prims, uint32_t n, pulse_blas_node *nodes_out,
                     uint32_t nodes_capacity, uint32_t *node_count, uint32_t *tri_indices_out);

/* Midpoint (2-bin SAH) TLAS over `n` instance bounds.  nodes_out must hold
 * max(1, 2*n-1) nodes; instance_indices_out holds n entries.  n == 0 yields
 * the single invalid root {0,0} (tlas.rs:13-20, 30-35). */
int pulse_build_tlas(const pulse_instance_bounds *inst, uint32_t n, pulse_tlas_node *nodes_out,
                     uint32_t nodes_capacity, uint32_t *node_count,
                     uint32_t *instance_indices_out);

/* Scene builder: meshes + materials + instances -> the 13 buffers. */
typedef struct pulse_scene_builder pulse_scene_builder;

typedef struct pulse_scene_build_info {
    uint32_t mesh_count;
    uint32_t instance_count;
    uint32_t light_count;
    uint32_t material_count;
    uint64_t unique_triangles;    /* sum of mesh triangle counts           */
    uint64_t instanced_triangles; /* sum over instances of triangle_count  */
    uint32_t blas_node_count;
    uint32_t tlas_node_count;
    uint32_t max_blas_depth;      /* edges from root to deepest leaf        */
    uint32_t tlas_depth;
    double ms_blas;               /* host build time of all BLAS            */
    double ms_instances_tlas;     /* instance prep + light CDFs + TLAS      */
} pulse_scene_build_info;

int pulse_sb_create(pulse_scene_builder **out);
int pulse_sb_destroy(pulse_scene_builder *sb);
/* positions/normals: vertex_count*3 floats; uvs: vertex_count*2 floats (may be
 * NULL: zeros); indices: index_count u32, a multiple of 3. */
int pulse_sb_add_mesh(pulse_scene_builder *sb, const float *positions, const float *normals,
                      const float *uvs, uint32_t vertex_count, const uint32_t *indices,
                      uint32_t index_count, uint32_t *mesh_id);
int pulse_sb_add_material(pulse_scene_builder *sb, const pulse_material *m, uint32_t *material_id);
/* transform: column-major object->world matrix (GlobalTransform::compute_matrix). */
int pulse_sb_add_instance(pulse_scene_builder *sb, uint32_t mesh_id, uint32_t material_id,
                          const float *transform16);
/* Builds every BLAS (threads > 1 builds meshes in parallel), then instances,
 * light tables and the TLAS. */
int pulse_sb_build(pulse_scene_builder *sb, int threads);
/* Moves instance `instance_id` (column-major object->world matrix); takes
 * effect at the next pulse_sb_update_instances / pulse_sb_build. */
int pulse_sb_set_instance_transform(pulse_scene_builder *sb, uint32_t instance_id, const float *transform16);
/* Recomputes the instance records, light tables and TLAS of a built scene
 * (the reference's per-frame prepare_mesh_instances + build_tlas,
 * src/scene/mod.rs:561-697) without touching meshes or BLAS; the result is
 * what pulse_sb_build would produce with the new transforms. */
int pulse_sb_update_instances(pulse_scene_builder *sb);
/* Views into builder-owned buffers, valid until the next build/destroy. */
int pulse_sb_get_desc(pulse_scene_builder *sb, pulse_scene_desc *out);
int pulse_sb_get_info(pulse_scene_builder *sb, pulse_scene_build_info *out);

#ifdef __cplusplus
}
#endif

#endif /* PULSE_SCENE_H */
