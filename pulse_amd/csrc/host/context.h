// context.h -- host-side state of libpulse_hip shared by its translation
// units: device buffers, device-resident scenes and the context.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../device/pt_device.h"
#include "../device/queue.h"
#include "common.h"

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            PULSE_FAIL(PULSE_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                              \
    } while (0)

namespace pulse_host {

// One device allocation that grows on demand.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && ptr) return PULSE_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        size_t b = std::max<size_t>(bytes, 256);
        HIP_TRY(hipMalloc(&ptr, b));
        cap = b;
        return PULSE_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

enum SceneBuf {
    B_PRIMS, B_TRI_BVH, B_NRM_BVH, B_BLAS, B_TLAS, B_INST_IDX, B_INSTANCES, B_MATERIALS,
    B_CDFS, B_AREAS, B_LIGHTS, B_SCDF, B_WTLAS, B_WBLAS, B_WINST, B_WTRI, B_COUNT
};

// Queue counter words of the single-view path: extension and shadow queue
// lengths, then the persistent launches' taken-ray heads (per trace index,
// one per queue part).
constexpr int kCounterWords = 2 * (pt::kMaxTraces + 1) + 2 * pt::kHeadWords;

enum WfBuf {
    W_RAY_O0, W_RAY_O1, W_RAY_D0, W_RAY_D1, W_HIT, W_HIT_V, W_STATE_T, W_STATE_C, W_SH_O, W_SH_D, W_SH_C,
    W_COUNTERS, W_COUNT
};

// A validated scene in device memory (the 13 bindings, with the triangle
// records regathered in BVH order) and what the launches need to know of it.
struct SceneData {
    DevBuf buf[B_COUNT];
    pulse_scene_uniform uniform{};
    uint32_t stack_need = 0;     // TLAS depth + deepest BLAS + 1
    uint32_t wide_need = 0;      // the same bound for the 4-wide BVHs
    uint32_t wide_blas_need = 0; // deepest 4-wide BLAS part of wide_need
    bool has_wide = false;       // the 4-wide BVHs were built
    uint32_t n_wide_tlas = 0;    // quantised 4-wide TLAS nodes
    bool binary_stale = false;   // a device refit moved instances: the binary TLAS is out of date
    bool tlas_all_entries = false;  // the TLAS leaves hold every instance_indices entry once (device rebuild)
    DevBuf refit[10];            // refit: matrices, instance boxes, child boxes, node boxes, error word,
                                 // BLAS cut boxes, per-instance cut ranges, 4-wide TLAS levels;
                                 // rebuild: scratch, node / level counts
    // per instance {first, count} into cut_boxes: the boxes of its BLAS cut
    // at kCutDepth, from which the 4-wide TLAS's instance boxes are made
    std::vector<float> cut_boxes;
    std::vector<uint32_t> cut_range;
    uint32_t tlas_depth = 0, blas_depth = 0;
    uint32_t n_materials = 0;
    uint32_t n_strength_cdf = 0;  // entries of light_emission_strength_cdf
    bool fast_div = false;       // node coordinates allow the exact fast division
    bool fast_div_blas = false;  // the BLAS part of fast_div
    bool valid = false;
    uint64_t generation = 0;     // bumped by every upload or instance update
    // host copies checked by instance updates: per instance {triangle_offset,
    // triangle_count, index_offset, node_offset}, and the light table
    std::vector<uint32_t> inst_ranges;
    std::vector<pulse_light_data_index> light_indices;
    // per instance: the first 4-wide node of its BLAS (the instance records'
    // root word, rewritten by instance updates)
    std::vector<uint32_t> wide_root;
    uint32_t n_iidx = 0;         // instance_indices entries (instance records)
    void release() {
        for (auto &b : buf) b.release();
        for (auto &b : refit) b.release();
        valid = false;
    }
};

// Validates `d` and uploads it into `s` (left unchanged on failure).
int scene_upload(SceneData &s, const pulse_scene_desc *d);
// Replaces the instances, TLAS and instance indices of an uploaded scene
// (same meshes, same lights); validation failures leave `s` unchanged.
int scene_update_instances(SceneData &s, const pulse_scene_desc *d);
// Moves the instances of an uploaded scene on the device (refit.h): new
// object_world matrices (16 floats each, column-major, host memory), their
// inverses, world boxes and a refit of the 4-wide TLAS.
int scene_refit_instances(SceneData &s, const float *object_world, uint32_t n, hipStream_t stream);
// PULSE_E_UNSUPPORTED if traversal loop `loop` cannot run on `s` (the
// binary TLAS is stale after a device refit).
int check_traversal(const SceneData &s, int loop);
// Device view of a scene; `errors` is the context's device error word.
pt::DevScene dev_scene(const SceneData &s, uint32_t *errors);
// Reads the context's device error word into *out and clears it when set
// (each condition is reported once); the caller has synchronised the stream.
int take_device_errors(pulse_ctx *ctx, uint32_t *out);
// Smallest compiled stack size that holds `need` entries (0: none).
uint32_t pick_stack(uint32_t need, int loop_variant);
// Stack entries the traversal `loop_variant` needs on scene `s` (0: the
// variant cannot run on it).
uint32_t stack_for(const SceneData &s, int loop_variant);

}  // namespace pulse_host

struct pulse_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    pulse_host::SceneData own;            // the scene of pulse_scene_upload
    pulse_host::SceneData *cur = nullptr; // the scene the launches use
    pulse_host::DevBuf wf[pulse_host::W_COUNT];
    pulse_host::DevBuf errors;  // u32 device error word
    void *noalloc_scratch = nullptr;  // this device's address of pulse_trace_noalloc's scratch array
    size_t wf_pixels = 0;
    uint64_t generation = 0;
    uint32_t stack_entries = 0;  // traversal stack capacity in use
    int loop_variant = 2;        // in effect: 0 nested TLAS/BLAS loops, 1 single loop, 2 4-wide BVHs
    bool loop_explicit = false;  // set through PULSE_OPT_TRAVERSAL_LOOP; else 2 where the scene has
                                 // 4-wide BVHs and 1 where it has not
    int persistent = 24;         // persistent traversal waves per CU (0: one ray per lane)
    int refill = 0;              // idle lanes that make a persistent wave take new rays (0: tuned)
    int gather = 0;              // persistent: lanes gathered before an instance change runs (0: tuned)
    int advance_period = 0;      // persistent: instance changes run at least every n steps (0: tuned)
    // the values in effect: explicit settings, else the traversal loop's tuned ones
    int eff_refill() const { return refill ? refill : (loop_variant == 2 ? 24 : 16); }
    int eff_gather() const { return gather ? gather : (loop_variant == 2 ? 64 : 1); }
    int eff_period() const { return advance_period ? advance_period : (loop_variant == 2 ? 2 : 1); }
    int lds_entries = 24;        // persistent: stack entries kept in LDS (0: all); the rest spill
    int tris_per_step = 2;       // persistent: leaf triangles tested per traversal step (1 or 2)
    pulse_host::DevBuf spill;    // persistent: spilled stack entries
    pulse_host::DevBuf loss_acc; // f64 accumulator of pulse_image_loss
    pulse_host::DevBuf census;   // PULSE_FLAG_COUNT_WORK: u64 [3 categories][4 counters]
    pulse_stats stats{};
    hipEvent_t ev[3 * pt::kMaxTraces + 8] = {};
    bool events_created = false;
    // concurrent shadow launches (PULSE_OPT_CONCURRENT_SHADOW): a second
    // stream, the fork/join events that order it against `stream` (which
    // stays the one every call completes on), and the timing events
    int concurrent_shadow = 1;
    int pair_waves = 0;  // persistent waves per CU of each of two launches that run at once (0: half)
    int eff_pair_waves() const { return pair_waves ? pair_waves : std::max(1, persistent / 2); }
    int pair_waves_shadow = 0;  // the shadow launch's share of a concurrent pair (0: eff_pair_waves)
    int eff_pair_waves_shadow() const { return pair_waves_shadow ? pair_waves_shadow : eff_pair_waves(); }
    // distinct materials per wave summed before their atomic; -1: none for a
    // gradient in LDS, 2 for one in device memory (the measured best)
    int adjoint_reduce_rounds = -1;
    // PULSE_FLAG_RECORD_PATHS: up to kMaxRecords images (one per camera),
    // each per (sample, pixel, trace index) the closest hit, its v and the
    // shadow ray's visibility, of samples 0 .. samples - 1 of one camera,
    // parameter set and scene state
    struct PathRecord {
        pulse_host::DevBuf hit, v, vis;
        bool valid = false;
        uint32_t samples = 0, npix = 0, traces = 0;
        uint64_t generation = 0, used = 0;
        const void *scene = nullptr;
        int loop = -1;
        pulse_camera cam{};
        pulse_params prm{};
        void release() {
            hit.release();
            v.release();
            vis.release();
            valid = false;
        }
    };
    static constexpr int kMaxRecords = 16;
    PathRecord recs[kMaxRecords];
    int rec_pending = -1;          // the record the pulse_trace call in progress fills
    uint64_t rec_clock = 0;
    bool rec_force = false;        // pulse_forward_adjoint records its forward pass
    bool rec_read = false;         // the last adjoint read recorded paths for every view
    hipStream_t stream2 = nullptr;
    hipEvent_t fork_ev[pt::kMaxTraces] = {}, join_ev[pt::kMaxTraces] = {};
    hipEvent_t ev2[6 * pt::kMaxTraces + 4] = {};
    bool events2_created = false;
    bool has_scene() const { return cur && cur->valid; }
};

struct pulse_scene {
    pulse_ctx *ctx = nullptr;
    pulse_host::SceneData data;
};

namespace pulse_host {
// The context's second stream and its fork/join events
// (PULSE_OPT_CONCURRENT_SHADOW), made on first use at default priority;
// with `timing` also the concurrent loop's timing events.
int ensure_stream2(pulse_ctx *ctx, bool timing);
}  // namespace pulse_host
