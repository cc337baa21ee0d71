/*
 * pulse_hip.h -- C ABI of the MI355X (gfx950) path-tracing hot path of pulse.
 *
 * The library replaces the wgpu dispatch of the `path_trace` compute shader
 * (src/path_tracer/node.rs:40-119 -> src/path_tracer/path_tracer.wgsl:58-124)
 * with hand-written HIP kernels.  It consumes the scene exactly as the
 * reference's host side lays it out: the 13 bindings of bind group 0
 * (src/scene/bindings.wgsl:14-28, built in src/scene/mod.rs:911-1090) in
 * encase/std430 byte layout (src/scene/types.wgsl:1-74).
 *
 * Conventions
 *   - Every entry point returns an int status: PULSE_OK (0) or a negative
 *     PULSE_E_* code.  pulse_last_error() returns a thread-local message.
 *   - No exception, abort or torch type crosses this boundary.
 *   - Device pointers are plain `void*`/`float*` obtained from hipMalloc (or
 *     pulse_device_alloc); host pointers are marked `host` in comments.
 *   - Calls on one context are externally serialized (one render-graph node
 *     drives one context, src/path_tracer/mod.rs:60-77).  Work is ordered on
 *     the context's stream; pulse_sync() waits for it.
 */
#ifndef PULSE_HIP_H
#define PULSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
#define PULSE_OK 0
#define PULSE_E_INVALID_ARG -1     /* null pointer, zero size, bad enum      */
#define PULSE_E_HIP -2             /* a HIP runtime call failed              */
#define PULSE_E_LAYOUT -3          /* buffer byte size not a multiple of the */
                                   /* element size, or an index out of range */
#define PULSE_E_STACK_OVERFLOW -4  /* traversal stack exceeded (device flag) */
#define PULSE_E_ITERATION_CAP -5   /* traversal hit its loop cap (100000    */
                                   /* steps per walk of the binary BVHs,     */
                                   /* loops 0/1; loop 2's 4-wide trees are   */
                                   /* checked acyclic at upload instead)     */
#define PULSE_E_NO_SCENE -6        /* trace called before scene upload       */
#define PULSE_E_UNSUPPORTED -7     /* parameter combination not supported    */
#define PULSE_E_BUILD -8           /* host BVH builder failed                */

/* ------------------------------------------------------------------------ */
/* Scene buffer layout (bind group 0).  Sizes are asserted in the sources.   */
/* ------------------------------------------------------------------------ */

/* binding 0: SceneUniform, src/scene/types.wgsl:3-6 (8 B) */
typedef struct pulse_scene_uniform {
    uint32_t instance_count;
    uint32_t light_count;
} pulse_scene_uniform;

/* binding 1: Primitive, types.wgsl:23-27; PulsePrimitive, scene/mod.rs:344-347.
 * Object-space vertices in original mesh order; vec3 at stride 16 (48 B). */
typedef struct pulse_primitive {
    float p_first[3];
    float _pad0;
    float p_second[3];
    float _pad1;
    float p_third[3];
    float _pad2;
} pulse_primitive;

/* binding 2: TriangleData, types.wgsl:29-36; PulseTriangleData,
 * scene/mod.rs:338-342 (80 B). */
typedef struct pulse_triangle_data {
    float n_first[3];
    float _pad0;
    float n_second[3];
    float _pad1;
    float n_third[3];
    float _pad2;
    float uv_first[2];
    float uv_second[2];
    float uv_third[2];
    float _pad3[2];
} pulse_triangle_data;

/* binding 4: BLASNode, types.wgsl:38-43; PulseBLASNode, blas.rs:5-13 (32 B).
 * tri_count > 0: leaf over triangle_indices[a_or_first_tri .. +tri_count];
 * tri_count == 0: interior with children a_or_first_tri and a_or_first_tri+1.
 * Indices are local to the mesh (node_offset/index_offset added per instance). */
typedef struct pulse_blas_node {
    float aabb_min[3];
    uint32_t a_or_first_tri;
    float aabb_max[3];
    uint32_t tri_count;
} pulse_blas_node;

/* binding 5: TLASNode, types.wgsl:45-50; PulseTLASNode, tlas.rs:4-10 (32 B). */
typedef struct pulse_tlas_node {
    float aabb_min[3];
    uint32_t a_or_first_instance;
    float aabb_max[3];
    uint32_t instance_count;
} pulse_tlas_node;

/* binding 7: MeshInstance, types.wgsl:52-60; PulseMeshInstance +
 * PulseMeshIndex, scene/mod.rs:460-466, 527-533 (160 B).  Matrices are
 * column-major (glam Mat4 / WGSL mat4x4f). */
typedef struct pulse_mesh_instance {
    float object_world[16];
    float world_object[16];
    uint32_t triangle_offset;
    uint32_t triangle_count;
    uint32_t index_offset;
    uint32_t node_offset;
    uint32_t material_index;
    uint32_t _pad[3];
} pulse_mesh_instance;

/* binding 8: Material, types.wgsl:62-68; PulseMaterial, scene/mod.rs:240-247 (48 B). */
typedef struct pulse_material {
    float base_color[4];
    float emissive[4];
    float perceptual_roughness;
    float reflectance;
    float metallic;
    float _pad;
} pulse_material;

/* binding 12: LightDataIndex, types.wgsl:70-74; PulseLightDataIndex,
 * scene/mod.rs:547-551 (8 B). */
typedef struct pulse_light_data_index {
    uint32_t cdf_offset;
    uint32_t mesh_instance_index;
} pulse_light_data_index;

/* A host buffer view (bytes must be a multiple of the element size). */
typedef struct pulse_buffer_view {
    const void *data; /* host */
    size_t bytes;
} pulse_buffer_view;

/* The 13 bindings of bind group 0 in binding order
 * (src/scene/mod.rs:1035-1088). */
typedef struct pulse_scene_desc {
    pulse_scene_uniform uniform;                   /* 0  */
    pulse_buffer_view primitives;                  /* 1  pulse_primitive[]       */
    pulse_buffer_view triangle_data;               /* 2  pulse_triangle_data[]   */
    pulse_buffer_view triangle_indices;            /* 3  uint32_t[]              */
    pulse_buffer_view blas_nodes;                  /* 4  pulse_blas_node[]       */
    pulse_buffer_view tlas_nodes;                  /* 5  pulse_tlas_node[]       */
    pulse_buffer_view instance_indices;            /* 6  uint32_t[]              */
    pulse_buffer_view instances;                   /* 7  pulse_mesh_instance[]   */
    pulse_buffer_view materials;                   /* 8  pulse_material[]        */
    pulse_buffer_view light_emission_strength_cdf; /* 9  float[] (unused, Q4)    */
    pulse_buffer_view light_triangle_area_cdfs;    /* 10 float[]                 */
    pulse_buffer_view light_mesh_areas;            /* 11 float[]                 */
    pulse_buffer_view light_indices;               /* 12 pulse_light_data_index[] */
} pulse_scene_desc;

/* ------------------------------------------------------------------------ */
/* View (bind group 1) and parameters                                        */
/* ------------------------------------------------------------------------ */

/* The parts of Bevy's `View` uniform the kernel reads
 * (path_tracer.wgsl:51, 70, 72) plus PathTracerUniform width/height
 * (path_tracer.wgsl:45-49). */
typedef struct pulse_camera {
    float inverse_view_proj[16]; /* column-major */
    float world_position[3];
    uint32_t width;
    uint32_t height;
} pulse_camera;

/* Path-loop modes (one path per pixel per sample). */
#define PULSE_MODE_REF_GGX 0        /* path_trace, path_tracer.wgsl:58-124          */
#define PULSE_MODE_REF_NEE_LAMBERT 1 /* path_traceBBB, path_tracer.wgsl:128-202, with */
                                     /* a camera primary ray instead of the G-buffer */
#define PULSE_MODE_GGX_NEE 2        /* GGX sampling + NEE; emissive at depth 0 only */

/* NEE evaluators for PULSE_MODE_GGX_NEE. */
#define PULSE_NEE_LAMBERT 0 /* sample_direct_light, utilities.wgsl:500-558     */
#define PULSE_NEE_GGX 1     /* sample_direct_light_ggx, utilities.wgsl:562-655 */

typedef struct pulse_params {
    uint32_t accumulation_count; /* frame index n of the running mean (Q18)       */
    uint32_t spp;                /* samples in this call, applied sequentially    */
                                 /* with accumulation counts n .. n+spp-1         */
    uint32_t max_depth;          /* path_trace uses 5 (path_tracer.wgsl:80)       */
    uint32_t mode;               /* PULSE_MODE_*                                  */
    uint32_t nee_eval;           /* PULSE_NEE_* (mode GGX_NEE only)               */
    uint32_t rr_start_depth;     /* Russian roulette from this depth on (ref: 0)  */
    uint32_t flags;              /* PULSE_FLAG_*                                  */
    uint32_t _reserved;
} pulse_params;

#define PULSE_FLAG_COLLECT_STATS 1u /* count rays per stage into pulse_stats */
#define PULSE_FLAG_STAGE_TIMING 2u  /* hipEvent timing per wavefront stage   */
#define PULSE_FLAG_HORIZON_CLAMP 4u /* not in the reference: a path ends when */
                                    /* n.wo <= 0 or the sampled n.wi <= 0 (the */
                                    /* reference keeps such samples, Q6, which */
                                    /* gives its GGX estimator unbounded       */
                                    /* weights); for the differentiable form   */
#define PULSE_FLAG_ADJOINT_REPARAM 8u /* adjoint: differentiate roughness through */
                                      /* the moving sampled directions instead of */
                                      /* at fixed directions (the default)        */
#define PULSE_FLAG_LIGHT_POWER 16u /* NEE picks the light by emission strength   */
                                   /* (light_emission_strength_cdf, pdf = its    */
                                   /* step) as the reference intended            */
                                   /* (utilities.wgsl:501, 541-546), instead of   */
                                   /* uniformly (its live code)                  */
#define PULSE_FLAG_COUNT_WORK 32u  /* traversal loop 2: before each trace launch,  */
                                   /* walk its queue once more and count node      */
                                   /* visits, triangle tests and instances entered */
                                   /* into pulse_stats.work (extra work: not for   */
                                   /* timed frames)                                */
#define PULSE_FLAG_ADJOINT_ACCUMULATE 64u /* adjoint: add to `grad` instead of     */
                                          /* overwriting it (the gradient of a sum */
                                          /* of per-view losses, one call a view)  */
#define PULSE_FLAG_RECORD_PATHS 128u /* pulse_trace: keep every path's closest hits */
                                     /* and shadow-ray visibilities (21 bytes per    */
                                     /* vertex) for a pulse_adjoint of the same       */
                                     /* image, which then reads them instead of       */
                                     /* tracing the paths again (same gradient).      */
                                     /* A call with accumulation_count 0 starts the   */
                                     /* record, later calls extend it.  One record    */
                                     /* per camera, up to 16 (least recently used     */
                                     /* replaced), 32 GB in all (older records give   */
                                     /* their memory back), none above 16 GB; a       */
                                     /* material update drops the scene's records     */

/* Ray / hit records of the debug trace entry point (RayHitRecord,
 * types.wgsl:8-21).  A miss leaves t = 1e30 and the ids at 0. */
typedef struct pulse_ray {
    float origin[3];
    float dir[3];
    float t_max; /* initial record.t; the reference uses 1e30 */
    uint32_t _pad;
} pulse_ray;

typedef struct pulse_hit {
    float t;
    uint32_t instance_index;
    uint32_t triangle_index; /* BVH-order local index (utilities.wgsl:277) */
    float u;
    float v;
} pulse_hit;

/* Per-call counters (exact ray counts come from the queue lengths), summed
 * over the samples of the call.  Stage times need PULSE_FLAG_STAGE_TIMING. */
typedef struct pulse_stats {
    uint64_t primary_rays;     /* closest-hit rays at trace index 0        */
    uint64_t extension_rays;   /* all closest-hit rays incl. primary       */
    uint64_t incoherent_rays;  /* extension rays at depth >= 1 + shadow    */
    uint64_t shadow_rays;
    uint32_t error_flags;      /* PULSE_ERR_* raised since the last read    */
    uint32_t stack_entries;    /* traversal stack depth used by the kernels */
    double ms_total;           /* generate .. accumulate, all samples      */
    double ms_primary_extend;  /* closest-hit stage at trace index 0       */
    double ms_incoherent;      /* wall time of the closest-hit stages at   */
                               /* trace index >= 1 and the occlusion       */
                               /* stages: their sum, or with concurrent    */
                               /* shadow launches the spans from each      */
                               /* shade stage's end until both launches it */
                               /* feeds are done                           */
    double ms_shade;           /* ms_total - primary - incoherent:         */
                               /* generate + shade + accumulate stages     */
    double ms_extend_bounce;   /* closest-hit launches at trace index >= 1 */
                               /* (each launch's own duration)             */
    double ms_shadow;          /* occlusion launches (own durations; they  */
                               /* overlap the former when concurrent)      */
    uint32_t extend_launches;  /* closest-hit stage launches               */
    uint32_t shadow_launches;  /* occlusion stage launches                 */
    uint64_t work[3][4];       /* PULSE_FLAG_COUNT_WORK: per ray category   */
                               /* (primary, bounce, shadow) rays, 4-wide    */
                               /* node visits, triangle tests, instances    */
} pulse_stats;

/* ------------------------------------------------------------------------ */
/* Context, memory and scene                                                 */
/* ------------------------------------------------------------------------ */
typedef struct pulse_ctx pulse_ctx;

const char *pulse_last_error(void);
const char *pulse_version(void);

/* Creates a context on HIP device `device` with its own stream. */
int pulse_create(int device, pulse_ctx **out);
int pulse_destroy(pulse_ctx *ctx);
/* The context's stream as a hipStream_t. */
void *pulse_stream(pulse_ctx *ctx);
int pulse_sync(pulse_ctx *ctx);
/* Self-test of the device arithmetic the triangle tests rely on (no
 * counterpart in the reference; for maintainers bringing the library up on
 * a new GPU or toolchain): counts the f32 a with |a| in [2^-100, 2^126),
 * every mantissa and both signs (3.8e9 values), for which the hardware
 * reciprocal refined by one fused Newton step differs from the correctly
 * rounded 1.0f / a.  The kernels assume 0.  Synchronous, about a second. */
int pulse_selftest_reciprocal(pulse_ctx *ctx, uint64_t *mismatches);

/* Device memory helpers for callers without a HIP binding (ctypes, FFI). */
int pulse_device_alloc(pulse_ctx *ctx, size_t bytes, void **out_device_ptr);
int pulse_device_free(pulse_ctx *ctx, void *device_ptr);
int pulse_copy_h2d(pulse_ctx *ctx, void *dst_device, const void *src_host, size_t bytes);
int pulse_copy_d2h(pulse_ctx *ctx, void *dst_host, const void *src_device, size_t bytes);
int pulse_memset_d(pulse_ctx *ctx, void *dst_device, int value, size_t bytes);

/* Replaces queue_scene_bind_group (src/scene/mod.rs:911-1090): copies the 13
 * host buffers into device buffers owned by the context, validates sizes and
 * index ranges, and derives the traversal-ordered triangle copy.  Re-upload
 * only when the scene changed; `generation` is returned by
 * pulse_scene_generation so callers can skip identical uploads. */
int pulse_scene_upload(pulse_ctx *ctx, const pulse_scene_desc *desc);
int pulse_scene_generation(pulse_ctx *ctx, uint64_t *out);

/* Replaces the material table only (binding 8) -- the per-iteration update
 * of an optimisation loop over material parameters.  Same count as the
 * uploaded scene. */
int pulse_scene_update_materials(pulse_ctx *ctx, const pulse_material *materials, uint32_t count);

/* Device-resident scenes beside the uploaded one, for rendering several
 * scenes from one context (the batched form; the reference keeps one scene
 * per render world, src/scene/mod.rs:911).  pulse_scene_create validates and
 * uploads like pulse_scene_upload; pulse_scene_bind makes a scene the one
 * that pulse_trace, pulse_trace_rays, pulse_adjoint and
 * pulse_scene_update_materials use (NULL: the uploaded scene again).  A
 * scene belongs to its context; destroy scenes before the context. */
typedef struct pulse_scene pulse_scene;
int pulse_scene_create(pulse_ctx *ctx, const pulse_scene_desc *desc, pulse_scene **out);
/* Per-frame instance update of a resident scene (`scene`, or the uploaded
 * one if NULL): replaces only the instance records, TLAS nodes and instance
 * indices (bindings 7, 5, 6) -- the buffers the reference's
 * prepare_mesh_instances and build_tlas (src/scene/mod.rs:561-697,
 * tlas.rs:29-136) rewrite every frame -- instead of all 13.  The meshes,
 * their BLAS and the light table must be unchanged (checked; otherwise
 * PULSE_E_LAYOUT: upload the scene).  Synchronises the context's stream. */
int pulse_scene_update_instances(pulse_ctx *ctx, pulse_scene *scene, const pulse_scene_desc *desc);
/* Instance motion without the host: `object_world` holds n = instance_count
 * new object-to-world matrices (16 floats each, column-major, host memory,
 * the MeshInstance field of scene/mod.rs:460-466).  On the device: the
 * matrices and their inverses (cofactor expansion in double, as
 * pulse_sb_build) go into the instance records, each instance's world box is
 * recomputed from its BLAS root (scene/mod.rs:624-680), and the 4-wide TLAS
 * is refit over the new boxes (the tree of the last build, new bounds) --
 * the part of the reference's per-frame prepare_mesh_instances / build_tlas
 * (scene/mod.rs:561-697, tlas.rs:29-136) that motion changes.  The
 * reference's binary TLAS is not rebuilt: afterwards only traversal loop 2
 * runs on the scene (loops 0 and 1 return PULSE_E_UNSUPPORTED until
 * pulse_scene_update_instances or pulse_scene_upload).  Requires the scene's
 * 4-wide BVHs.  Synchronises the context's stream. */
int pulse_scene_refit_instances(pulse_ctx *ctx, pulse_scene *scene, const float *object_world, uint32_t n);
/* As pulse_scene_refit_instances (same arguments, same instance records and
 * world boxes), with a new 4-wide TLAS built on the device over the moved
 * instances instead of a refit of the last one: a refit's boxes overlap
 * after large motion and traversal slows (the bench scene with all 800
 * instances moved: 12.9 ms per frame refit, 5.3 ms rebuilt); the rebuild
 * keeps it.  No counterpart in the reference, which rebuilds its TLAS on the
 * host every frame (build_tlas, tlas.rs:29-136): this is that rebuild, for
 * the 4-wide traversal, on the GPU.  One block sorts the entries' Morton
 * codes and builds a binary radix tree (Karras 2012) collapsed to 4-wide
 * nodes, so it takes scenes of at most 4096 instance_indices entries
 * (PULSE_E_UNSUPPORTED above) whose TLAS holds every entry once.  The binary
 * TLAS is stale afterwards, as after a refit.  Synchronous. */
int pulse_scene_rebuild_instances(pulse_ctx *ctx, pulse_scene *scene, const float *object_world, uint32_t n);
int pulse_scene_destroy(pulse_scene *scene);
int pulse_scene_bind(pulse_ctx *ctx, pulse_scene *scene);

/* ------------------------------------------------------------------------ */
/* Hot path                                                                  */
/* ------------------------------------------------------------------------ */

/* Replaces PulsePathTracerNode::run's dispatch of `path_trace`
 * (src/path_tracer/node.rs:104-116).  Renders params->spp samples per pixel
 * and folds each into `accum_rgba` (device, width*height*4 floats, rgba32f,
 * row-major) with the reference's running mean (path_tracer.wgsl:116-123).
 * Asynchronous on the context stream; errors raised on the device are
 * reported by pulse_sync / pulse_get_stats. */
int pulse_trace(pulse_ctx *ctx, const pulse_camera *camera, const pulse_params *params,
                float *accum_rgba);

/* No counterpart in the reference.  pulse_trace for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): the same image bit for bit, with the frame's queues and path state
 * in a fixed scratch array that the library holds on each device until the
 * process ends (about 10.8 MB) instead of buffers it allocates.  Calls no
 * hipMalloc/hipFree.  Limits: width*height <= pulse_trace_noalloc_max_pixels
 * (65536, e.g. 256x256), else PULSE_E_UNSUPPORTED; no PULSE_FLAG_STAGE_TIMING
 * or PULSE_FLAG_COUNT_WORK.  Launches run one after the other on the context
 * stream with one ray per lane.  Every call on a device shares that device's
 * scratch: calls that overlap on one GPU (two streams, two contexts, two
 * threads) race; calls ordered on one stream do not. */
int pulse_trace_noalloc(pulse_ctx *ctx, const pulse_camera *camera, const pulse_params *params,
                        float *accum_rgba);
int pulse_trace_noalloc_max_pixels(uint64_t *out);

/* The realtime GI pass (src/pulse/gi.wgsl:56-126, dispatched by the GI
 * plugin): paths start at the caller's G-buffer instead of camera rays.
 * Device arrays of width*height float4 each (w unused): world position,
 * world normal, base colour, emissive -- what the reference decodes from
 * Bevy's deferred prepass (pbr_input_from_deferred_gbuffer).  Per pixel: RNG
 * seeded with pixel_index * 5817321 and run on across params->spp samples
 * (the reference uses 2); each sample scatters (cosine) from the G-buffer
 * point with Lambert NEE there, then up to params->max_depth bounces (the
 * reference uses 5) with NEE and Russian roulette at every depth; gi_rgba
 * (device, width*height*4) = the average colour, alpha 1 (not accumulated
 * across calls).  params: spp, max_depth, flags (PULSE_FLAG_LIGHT_POWER);
 * nee_eval must be PULSE_NEE_LAMBERT; mode, rr_start_depth and
 * accumulation_count are not used. */
typedef struct {
    const float *position;
    const float *normal;
    const float *base_color;
    const float *emissive;
} pulse_gbuffer;
int pulse_trace_gbuffer(pulse_ctx *ctx, const pulse_gbuffer *gbuffer, uint32_t width, uint32_t height,
                        const pulse_params *params, float *gi_rgba);

/* Closest hit for a batch of world-space rays (trace_ray,
 * utilities.wgsl:112-114 -> traverse_tlas 127-188 -> traverse_blas
 * 259-326).  rays/hits are device arrays of n records. */
int pulse_trace_rays(pulse_ctx *ctx, const pulse_ray *rays, pulse_hit *hits, uint32_t n);

/* Occlusion for a batch of rays: occluded[i] = 1 iff the closest hit of ray i
 * has t < t_max (the shadow test of utilities.wgsl:517-525 with its
 * threshold |pl - origin| - 0.005 passed in t_max). */
int pulse_trace_shadow_rays(pulse_ctx *ctx, const pulse_ray *rays, uint32_t *occluded, uint32_t n);

/* Tuning options of a context (defaults are the measured best on MI355X). */
#define PULSE_OPT_TRAVERSAL_LOOP 1 /* 0: nested TLAS/BLAS loops as in the reference, */
                                   /* 1: one loop over both levels: 0 and 1 give the */
                                   /*    reference's results bit for bit              */
                                   /* 2 (default where the scene has them, else 1):  */
                                   /*    4-wide BVHs collapsed at upload from the     */
                                   /*    reference's BVHs: same leaves, fewer steps;  */
                                   /*    among hits with equal t (or within rounding  */
                                   /*    of each other) the first found is kept, so   */
                                   /*    such ties may resolve differently            */
#define PULSE_OPT_STACK_ENTRIES 2  /* read-only: LDS traversal stack entries in use  */
#define PULSE_OPT_FAST_DIVISION 3  /* read-only: 1 if the scene's coordinates allow   */
                                   /* the exact reciprocal-multiply slab division     */
#define PULSE_OPT_PERSISTENT_WAVES 4 /* single loop: 0 = one ray per lane; n = a     */
                                     /* persistent grid of n waves per CU whose lanes */
                                     /* take new rays as theirs finish               */
#define PULSE_OPT_REFILL_LANES 5   /* persistent: idle lanes that trigger a refill   */
                                   /* (0: the traversal loop's tuned value, 16 for   */
                                   /* loop 1, 24 for loop 2)                         */
#define PULSE_OPT_GATHER_LANES 6   /* persistent: lanes waiting for an instance change */
                                   /* before the wave runs it (1: no waiting; 0: the */
                                   /* loop's tuned value, 1 for loop 1, 64 for 2)    */
#define PULSE_OPT_LDS_STACK_ENTRIES 7 /* persistent: stack entries per lane kept in   */
                                      /* LDS (0: all); deeper entries go to device    */
                                      /* memory, so more waves fit per CU             */
#define PULSE_OPT_TRIS_PER_STEP 8  /* persistent: leaf triangles per traversal step  */
                                   /* (1 or 2; same order, same results)            */
#define PULSE_OPT_ADVANCE_PERIOD 9 /* persistent: every n-th step a wave also runs   */
                                   /* the instance changes its lanes wait for, however */
                                   /* few (1: every step; 0: the loop's tuned value,  */
                                   /* 1 for loop 1, 2 for loop 2)                     */
#define PULSE_OPT_CONCURRENT_SHADOW 10 /* persistent: 1 (default) = the shadow rays  */
                                       /* of bounce k trace on a second stream while */
                                       /* bounce k+1's extension rays trace (the two */
                                       /* launches fill each other's tails); 0 = one */
                                       /* after the other.  Same results either way  */
#define PULSE_OPT_PAIR_WAVES 11    /* persistent, concurrent shadow launches: waves */
                                   /* per CU of each of the two launches that run at */
                                   /* once (0, the default: half of                  */
                                   /* PULSE_OPT_PERSISTENT_WAVES), so that both start */
                                   /* at once; launches that run alone take all      */
#define PULSE_OPT_PAIR_WAVES_SHADOW 14 /* persistent, concurrent launches: waves */
                                       /* per CU of the shadow launch of a pair   */
                                       /* (0: PULSE_OPT_PAIR_WAVES).  A launch    */
                                       /* takes at most PULSE_OPT_PERSISTENT_WAVES */
#define PULSE_OPT_ADJOINT_REDUCE_ROUNDS 15 /* adjoint: per add of gradient terms, up to */
                                           /* n materials of a wave (in lane order)   */
                                           /* are summed over the wave and added by   */
                                           /* one atomic each; the other lanes add    */
                                           /* their own (0: per-lane atomics; 0..16). */
                                           /* -1 (default): 0 while the gradient fits */
                                           /* in LDS, 2 when it is in device memory   */
#define PULSE_OPT_RECORDED_SAMPLES 17 /* read-only: samples per pixel of the valid  */
                                      /* path record (PULSE_FLAG_RECORD_PATHS), 0   */
                                      /* if none                                    */
#define PULSE_OPT_ADJOINT_READ_RECORD 18 /* read-only: 1 if the last adjoint read   */
                                         /* the recorded paths, 0 if it traced them */
#define PULSE_OPT_INJECT_ERROR 12  /* write-only, for tests of error reporting: ORs */
                                   /* the value (PULSE_ERR_* bits) into the device */
                                   /* error word, as a kernel that met the          */
                                   /* condition would                                */
/* Bits of the device error word (pulse_stats.error_flags).  pulse_sync,
 * pulse_get_stats and pulse_batch_get_stats report the bits raised since the
 * word was last read and clear it: each condition is reported once. */
#define PULSE_ERR_STACK_OVERFLOW 1u
#define PULSE_ERR_ITERATION_CAP 2u
int pulse_set_option(pulse_ctx *ctx, int option, int64_t value);
int pulse_get_option(pulse_ctx *ctx, int option, int64_t *value);

/* Replaces the path tracer's upscaling pass (src/path_tracer/pt_upscaling/
 * node.rs:29-97, pt_upscaling.wgsl:39-50): draws the rgba32f accumulation
 * buffer (src_w x src_h) into a dst_w x dst_h target of the given format,
 * nearest texel at each target pixel centre, clamped to the edge (wgpu's
 * default sampler), converted as a target of that format stores it.  Both
 * buffers are device memory; asynchronous on the context stream. */
#define PULSE_PRESENT_RGBA32F 0
#define PULSE_PRESENT_RGBA16F 1     /* the HDR view target: IEEE half, round to nearest even */
#define PULSE_PRESENT_RGBA8_SRGB 2  /* linear -> sRGB on rgb, 8-bit, round to nearest        */
#define PULSE_PRESENT_RGBA8_UNORM 3
int pulse_present(pulse_ctx *ctx, const float *accum_rgba, uint32_t src_w, uint32_t src_h, void *dst,
                  uint32_t dst_w, uint32_t dst_h, int format);

/* ------------------------------------------------------------------------ */
/* Batched form: many views rendered by one captured launch sequence         */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference, which renders one camera per render-graph
 * node (src/path_tracer/node.rs:40-119); each view here is what one such
 * node renders: a scene, a camera (its own resolution) and an rgba32f
 * accumulation buffer (device, width*height*4 floats) updated with the
 * reference's running mean.  Every view's image is bit-identical to
 * pulse_trace of that view with the same params. */
typedef struct {
    pulse_scene *scene;
    pulse_camera camera;
    float *accum_rgba;
} pulse_view;

/* Launch shape of a batch; none of it changes the images. */
typedef struct {
    uint32_t persistent_waves_per_cu; /* 0: one wave per chunk of rays; n: a fixed */
                                      /* grid of n waves per CU drains the queues   */
    uint32_t chunk_rays;              /* rays per work item: a power of two, 64..4096 */
    uint32_t stack_entries;           /* LDS traversal stack (0: smallest that fits) */
    uint32_t traversal_loop;          /* as PULSE_OPT_TRAVERSAL_LOOP                */
    uint32_t regroup_every;           /* compact the extension queue every n bounces */
                                      /* (in between, rays stay in place)           */
    uint32_t tile_width;              /* primary-ray tile: width x (64 / width)      */
    uint32_t use_graph;               /* 1: replay one captured hipGraph per frame,  */
                                      /* its launches in sequence; 0 (default):     */
                                      /* direct launches, the shadow launches of    */
                                      /* bounce k beside bounce k+1's extension     */
                                      /* launch (PULSE_OPT_CONCURRENT_SHADOW)        */
    uint32_t lds_stack_entries;       /* persistent single loop: stack entries per   */
                                      /* lane in LDS, the rest in device memory     */
                                      /* (0: all in LDS)                            */
    uint32_t refill_lanes;            /* persistent single loop: idle lanes of a     */
                                      /* wave that take the next rays of its chunk  */
    uint32_t _reserved;
} pulse_batch_config;

typedef struct pulse_batch pulse_batch;
int pulse_batch_default_config(pulse_batch_config *out);
/* Allocates the per-view queues and path state (about 164 bytes per pixel
 * of all views).  Resolutions, scenes, accumulation buffers and params are
 * fixed for the batch's life. */
int pulse_batch_create(pulse_ctx *ctx, const pulse_view *views, uint32_t n_views, const pulse_params *params,
                       pulse_batch **out);
int pulse_batch_destroy(pulse_batch *batch);
/* Renders params->spp samples of every view, sample s with seed count
 * accumulation_count + s (asynchronous on the context stream). */
int pulse_batch_render(pulse_batch *batch, uint32_t accumulation_count);
/* New camera matrices for a view (same width and height). */
int pulse_batch_set_camera(pulse_batch *batch, uint32_t view, const pulse_camera *camera);
int pulse_batch_set_config(pulse_batch *batch, const pulse_batch_config *config);
int pulse_batch_get_config(pulse_batch *batch, pulse_batch_config *out);
/* Times frames under alternative launch shapes (coordinate descent over the
 * fields of pulse_batch_config, two passes, best of `reps` frames per shape,
 * each shape timed once) and keeps the fastest.  It renders into the views' accumulation buffers with
 * accumulation_count 0: tune before the frames that count. */
int pulse_batch_autotune(pulse_batch *batch, uint32_t reps, pulse_batch_config *best, double *best_ms);
/* Extension and shadow queue entries since the last call (synchronous). */
int pulse_batch_get_stats(pulse_batch *batch, pulse_stats *out);
/* One-shot: create, render params->accumulation_count, destroy. */
int pulse_trace_batch(pulse_ctx *ctx, const pulse_view *views, uint32_t n_views, const pulse_params *params);

/* ------------------------------------------------------------------------ */
/* Differentiable form: seed-replay adjoint of an L2 image loss              */
/* ------------------------------------------------------------------------ */
/* Gradient layout: 7 floats per material m at grad[7*m]:
 *   d/d base_color.r, .g, .b, d/d perceptual_roughness, d/d emissive.r, .g, .b
 * Loss: L = mean over pixels and rgb of (image - target)^2. */
#define PULSE_GRAD_PER_MATERIAL 7

/* L of `image` against `target` (device, npix*4 floats each); synchronous. */
int pulse_image_loss(pulse_ctx *ctx, const float *image, const float *target, uint32_t npix, double *loss);

/* dL/dparameters of the image rendered by pulse_trace with the same camera
 * and params: samples 0 .. accumulation_count + spp - 1 of every pixel,
 * i.e. the image pulse_trace accumulates over calls with these parameters
 * (accumulation_count 0 for a single call); any mode (PULSE_MODE_GGX_NEE,
 * the reference's path_trace PULSE_MODE_REF_GGX, or its path_traceBBB
 * PULSE_MODE_REF_NEE_LAMBERT, whose first hit takes one vertex more); at
 * most 8 path vertices -- max_depth <= 8, <= 7 in the Lambert mode; 4 / 3
 * with PULSE_FLAG_ADJOINT_REPARAM --, any rr_start_depth: Russian roulette is
 * replayed with its decisions; a random survival (p <= 1) contributes its
 * score, a certain one (p > 1) the derivative of its 1/p).
 * Replays every path from its seed; `grad`
 * (device, n_materials*7 floats) is overwritten, or added to with
 * PULSE_FLAG_ADJOINT_ACCUMULATE.  `replay_rgba` (device,
 * optional) receives the replayed image, which equals the forward image bit
 * for bit.  Metallic materials are differentiated in their albedo (which
 * moves f0 and with it the lobe probability; metallic itself is not a
 * parameter).  PULSE_FLAG_LIGHT_POWER is allowed: the light pick's pdf is
 * held at its value (the expected light sample does not depend on it). */
int pulse_adjoint(pulse_ctx *ctx, const pulse_camera *camera, const pulse_params *params, const float *image,
                  const float *target, float *grad, float *replay_rgba);

/* One view of a multi-view adjoint: the camera, the forward image rendered by
 * pulse_trace with it (device, width*height*4), the target (device, same
 * size) and an optional replay buffer as in pulse_adjoint. */
typedef struct pulse_adjoint_view {
    pulse_camera camera;
    const float *image;
    const float *target;
    float *replay_rgba;
} pulse_adjoint_view;

/* The batched adjoint (the reference has no differentiable form; its
 * per-view camera is the View uniform of path_tracer.wgsl:45-75): dL/dparams
 * of L = sum over the views of each view's loss, for n_views >= 1 views of
 * the bound scene that may differ in camera and resolution, in one launch
 * per PULSE_ADJOINT_MAX_VIEWS views (each launch adds to `grad`).  Same
 * requirements on params as pulse_adjoint; `grad` is overwritten, or added
 * to with PULSE_FLAG_ADJOINT_ACCUMULATE.  Equals the sum of pulse_adjoint
 * over the views up to the order of the f32 additions. */
#define PULSE_ADJOINT_MAX_VIEWS 16
int pulse_adjoint_views(pulse_ctx *ctx, const pulse_adjoint_view *views, uint32_t n_views, const pulse_params *params,
                        float *grad);

/* Forward + loss + adjoint in one call: renders into image_out (device,
 * width*height*4), writes the loss (host) and the gradient (device). */
int pulse_forward_adjoint(pulse_ctx *ctx, const pulse_camera *camera, const pulse_params *params,
                          const float *target, float *image_out, float *grad, double *loss);

/* Counters of the last pulse_trace call (synchronizes the stream). */
int pulse_get_stats(pulse_ctx *ctx, pulse_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* PULSE_HIP_H */
