// pulse_hip.hip -- context, scene upload, launch sequencing and the C ABI of
// libpulse_hip (declared in include/pulse_hip.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "device/adjoint.h"
#include "device/present.h"
#include "device/refit.h"
#include "device/wavefront.h"
#include "host/common.h"
#include "host/context.h"
#include "host/wide_bvh.h"

using namespace pt;

using namespace pulse_host;

// Scratch of pulse_trace_noalloc: the wavefront queues, path state and queue
// counters of one frame of at most kNoAllocMaxPixels pixels (ten float4
// arrays and one float array per pixel, then the counter words), one copy
// per device, held until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress gives for the current device
// (pulse_ctx::noalloc_scratch); every call on a device shares it.
constexpr size_t kNoAllocMaxPixels = size_t(1) << 16;
constexpr size_t kNoAllocF4 = 10;
constexpr size_t kNoAllocBytes =
    kNoAllocMaxPixels * (kNoAllocF4 * sizeof(float4) + sizeof(float)) + sizeof(uint32_t) * kCounterWords;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

int set_device(pulse_ctx *ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return PULSE_OK;
}

// Depth of the BLAS rooted at node_offset `root` (child indices are local to
// the mesh), checking every child index and leaf triangle range.
bool walk_blas(const pulse_blas_node *nodes, size_t n, uint32_t root, uint32_t tri_count, uint32_t &depth_out) {
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.push_back({0u, 0u});
    size_t visited = 0;
    while (!st.empty()) {
        auto [local, d] = st.back();
        st.pop_back();
        uint64_t ni = static_cast<uint64_t>(root) + local;
        if (ni >= n || ++visited > n) return false;
        best = std::max(best, d);
        const pulse_blas_node &x = nodes[ni];
        if (x.tri_count == 0) {
            if (x.a_or_first_tri == 0 || static_cast<uint64_t>(root) + x.a_or_first_tri + 1 >= n) return false;
            st.push_back({x.a_or_first_tri, d + 1});
            st.push_back({x.a_or_first_tri + 1, d + 1});
        } else if (static_cast<uint64_t>(x.a_or_first_tri) + x.tri_count > tri_count) {
            return false;
        }
    }
    depth_out = best;
    return true;
}

// Depth of the TLAS, checking child indices and leaf instance ranges.
bool walk_tlas(const pulse_tlas_node *nodes, size_t n, size_t n_iidx, uint32_t &depth_out) {
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.push_back({0u, 0u});
    size_t visited = 0;
    while (!st.empty()) {
        auto [ni, d] = st.back();
        st.pop_back();
        if (ni >= n || ++visited > n) return false;
        best = std::max(best, d);
        const pulse_tlas_node &x = nodes[ni];
        if (x.instance_count == 0) {
            if (x.a_or_first_instance == 0 || static_cast<uint64_t>(x.a_or_first_instance) + 1 >= n) return false;
            st.push_back({x.a_or_first_instance, d + 1});
            st.push_back({x.a_or_first_instance + 1, d + 1});
        } else if (static_cast<uint64_t>(x.a_or_first_instance) + x.instance_count > n_iidx) {
            return false;
        }
    }
    depth_out = best;
    return true;
}

// True if the leaves of `nodes` (walked by walk_tlas) hold each of the n_iidx
// instance_indices entries exactly once: the device rebuild makes a tree
// over all entries.
bool tlas_covers_entries(const pulse_tlas_node *nodes, size_t n_iidx) {
    std::vector<uint8_t> seen(n_iidx, 0);
    std::vector<uint32_t> st{0u};
    size_t held = 0;
    while (!st.empty()) {
        const pulse_tlas_node &x = nodes[st.back()];
        st.pop_back();
        if (x.instance_count == 0) {
            st.push_back(x.a_or_first_instance + 1u);
            st.push_back(x.a_or_first_instance);
            continue;
        }
        for (uint32_t c = 0; c < x.instance_count; ++c) {
            if (seen[x.a_or_first_instance + c]) return false;
            seen[x.a_or_first_instance + c] = 1;
            ++held;
        }
    }
    return held == n_iidx;
}

// The parameters a path record depends on (everything but the sample range
// and the flags that do not change paths).
bool same_paths(const pulse_params &a, const pulse_params &b) {
    const uint32_t keep = PULSE_FLAG_HORIZON_CLAMP | PULSE_FLAG_LIGHT_POWER;
    return a.max_depth == b.max_depth && a.mode == b.mode && a.nee_eval == b.nee_eval &&
           a.rr_start_depth == b.rr_start_depth && (a.flags & keep) == (b.flags & keep);
}
bool same_camera(const pulse_camera &a, const pulse_camera &b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

// A path record made with this camera and path parameters on the scene
// state now bound.
constexpr size_t kRecordBudget = 32ull << 30;  // bytes of path records per context
bool record_matches(const pulse_ctx *ctx, const pulse_ctx::PathRecord &r, const pulse_camera &cam,
                    const pulse_params &prm) {
    return r.scene == ctx->cur && r.generation == ctx->cur->generation && r.loop == ctx->loop_variant &&
           same_camera(r.cam, cam) && same_paths(r.prm, prm);
}
DevScene dev_scene(pulse_ctx *ctx) { return pulse_host::dev_scene(*ctx->cur, ctx->errors.as<uint32_t>()); }

// Makes `sd` the scene of the context's launches.
int bind_scene(pulse_ctx *ctx, SceneData *sd) {
    const int lv = ctx->loop_explicit ? ctx->loop_variant : (sd->has_wide ? 2 : 1);
    uint32_t st = stack_for(*sd, lv);
    if (st == 0) {
        if (lv == 2 && !sd->has_wide)
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "the scene has no 4-wide BVH (a leaf starts at slot 2^26 - 32 or later, "
                       "a node coordinate is not finite or beyond 2^40, or an instance matrix's bottom row is not "
                       "0 0 0 1)");
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "BVH too deep: TLAS depth %u + BLAS depth %u exceeds the %s stack",
                   sd->tlas_depth, sd->blas_depth, lv == 0 ? "nested loop's" : lv == 1 ? "single loop's" : "4-wide");
    }
    ctx->loop_variant = lv;
    ctx->cur = sd;
    ctx->stack_entries = st;
    ctx->generation += 1;
    return PULSE_OK;
}

constexpr int kTraceBlock = 64;  // one wave per block: LDS stack granularity

template <int STACK, int LOOP>
void launch_trace_rays(pulse_ctx *ctx, const pulse_ray *rays, pulse_hit *hits, uint32_t n) {
    uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_trace_rays<kTraceBlock, STACK, LOOP>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream,
                       dev_scene(ctx), rays, hits, n);
}
template <int STACK, int LOOP>
void launch_shadow_rays(pulse_ctx *ctx, const pulse_ray *rays, uint32_t *occ, uint32_t n) {
    uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_shadow_rays<kTraceBlock, STACK, LOOP>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream,
                       dev_scene(ctx), rays, occ, n);
}
// Persistent grid: `persistent` waves per CU (never more than the queue
// could need).
// Persistent grid: `waves` per CU (0: PULSE_OPT_PERSISTENT_WAVES), no more
// waves than the queue's capacity needs.
// A launch never takes more than PULSE_OPT_PERSISTENT_WAVES per CU: the
// spill area holds that grid's columns (spill_part_words).
uint32_t persistent_grid(pulse_ctx *ctx, uint32_t cap, int waves = 0) {
    const int w = waves > 0 ? std::min(waves, ctx->persistent) : ctx->persistent;
    uint32_t g = static_cast<uint32_t>(ctx->num_cus) * static_cast<uint32_t>(w);
    return std::max<uint32_t>(1, std::min<uint32_t>(g, (cap + kTraceBlock - 1) / kTraceBlock));
}
// LDS entries of the persistent kernels' stacks: the compiled size at most
// the requested one (and at most the stack's capacity).
constexpr int kLdsSizes[] = {8, 16, 24, 32, 40, 48, 64, 128};
int persistent_lds(pulse_ctx *ctx) {
    int want = ctx->lds_entries > 0 ? std::min<int>(ctx->lds_entries, ctx->stack_entries) : ctx->stack_entries;
    int got = kLdsSizes[0];
    for (int s : kLdsSizes)
        if (s <= want) got = s;
    return got;
}
template <int LDS, bool SHADOW, int TPS, int LOOP>
void launch_queue_k(pulse_ctx *ctx, const WfBuffers &wb, uint32_t grid, int q, int k, uint32_t *spill,
                    hipStream_t st) {
    hipLaunchKernelGGL((k_trace_queue<LDS, SHADOW, TPS, LOOP>), dim3(grid), dim3(kTraceBlock), 0, st,
                       dev_scene(ctx), wb, q, k, static_cast<uint32_t>(ctx->eff_refill()),
                       static_cast<uint32_t>(ctx->eff_gather()), static_cast<uint32_t>(ctx->eff_period()),
                       static_cast<int>(ctx->stack_entries), spill);
}
template <int LDS, int LOOP>
void launch_queue_loop(pulse_ctx *ctx, const WfBuffers &wb, uint32_t grid, bool shadow, int q, int k,
                       uint32_t *spill, hipStream_t st) {
    if (ctx->tris_per_step == 2) {
        if (shadow) launch_queue_k<LDS, true, 2, LOOP>(ctx, wb, grid, q, k, spill, st);
        else launch_queue_k<LDS, false, 2, LOOP>(ctx, wb, grid, q, k, spill, st);
    } else {
        if (shadow) launch_queue_k<LDS, true, 1, LOOP>(ctx, wb, grid, q, k, spill, st);
        else launch_queue_k<LDS, false, 1, LOOP>(ctx, wb, grid, q, k, spill, st);
    }
}
template <int LDS>
void launch_queue_lds(pulse_ctx *ctx, const WfBuffers &wb, uint32_t grid, bool shadow, int q, int k, uint32_t *spill,
                      hipStream_t st) {
    if (ctx->loop_variant == 2) launch_queue_loop<LDS, 2>(ctx, wb, grid, shadow, q, k, spill, st);
    else launch_queue_loop<LDS, 1>(ctx, wb, grid, shadow, q, k, spill, st);
}
// Spill area of the persistent launches over queues of up to `cap` rays:
// the stack entries below the LDS part, for every lane of the grid, twice
// (two launches may run at once).  Grows only after the stream has drained
// (no launch in flight uses it: the second stream's work is joined into the
// first before any call returns).
size_t spill_part_words(pulse_ctx *ctx, uint32_t cap) {
    const int lds = persistent_lds(ctx);
    if (static_cast<int>(ctx->stack_entries) <= lds) return 0;
    return static_cast<size_t>(persistent_grid(ctx, cap)) * kTraceBlock * (ctx->stack_entries - lds);
}
int ensure_spill(pulse_ctx *ctx, uint32_t cap) {
    const size_t bytes = 2 * sizeof(uint32_t) * spill_part_words(ctx, cap);
    if (bytes == 0 || (bytes <= ctx->spill.cap && ctx->spill.ptr)) return PULSE_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->spill.ensure(bytes);
}
// Persistent extend/shadow on stream `st` (default: the context's) with
// spill part `part` (0 or 1: launches that may run at once use different
// parts) and `waves` per CU (0: PULSE_OPT_PERSISTENT_WAVES); ensure_spill
// first.
int launch_queue(pulse_ctx *ctx, const WfBuffers &wb, uint32_t cap, bool shadow, int q, int k,
                 hipStream_t st = nullptr, int part = 0, int waves = 0) {
    const uint32_t grid = persistent_grid(ctx, cap, waves);
    const int lds = persistent_lds(ctx);
    if (!st) st = ctx->stream;
    uint32_t *spill = static_cast<int>(ctx->stack_entries) > lds
                          ? ctx->spill.as<uint32_t>() + static_cast<size_t>(part) * spill_part_words(ctx, cap)
                          : nullptr;
    switch (lds) {
        case 8: launch_queue_lds<8>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 16: launch_queue_lds<16>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 24: launch_queue_lds<24>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 32: launch_queue_lds<32>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 40: launch_queue_lds<40>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 48: launch_queue_lds<48>(ctx, wb, grid, shadow, q, k, spill, st); break;
        case 64: launch_queue_lds<64>(ctx, wb, grid, shadow, q, k, spill, st); break;
        default: launch_queue_lds<128>(ctx, wb, grid, shadow, q, k, spill, st); break;
    }
    return PULSE_OK;
}
// Work census (PULSE_FLAG_COUNT_WORK) of queue q / the shadow queue at k
// into category `cat` of ctx->census.
template <int STACK, int LOOP>
void launch_census(pulse_ctx *ctx, const WfBuffers &wb, uint32_t cap, bool shadow, int q, int k, int cat) {
    uint32_t grid = (cap + kTraceBlock - 1) / kTraceBlock;
    unsigned long long *out = ctx->census.as<unsigned long long>() + 4 * cat;
    if (shadow)
        hipLaunchKernelGGL((k_census<STACK, true>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream, dev_scene(ctx), wb, q,
                           k, out);
    else
        hipLaunchKernelGGL((k_census<STACK, false>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream, dev_scene(ctx), wb,
                           q, k, out);
}
template <int STACK, int LOOP>
void launch_extend(pulse_ctx *ctx, const WfBuffers &wb, uint32_t cap, int q, int k) {
    uint32_t grid = (cap + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_extend<kTraceBlock, STACK, LOOP>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream,
                       dev_scene(ctx), wb, q, k);
}
template <int STACK, int LOOP>
void launch_shadow(pulse_ctx *ctx, const WfBuffers &wb, uint32_t cap, int k) {
    uint32_t grid = (cap + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_shadow<kTraceBlock, STACK, LOOP>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream,
                       dev_scene(ctx), wb, k);
}

// Instantiates a launcher for the context's stack size and traversal loop.
#define PULSE_STACK_DISPATCH(fn, ctx, ...)                                           \
    do {                                                                             \
        if ((ctx)->loop_variant == 0) {                                              \
            switch ((ctx)->stack_entries) {                                          \
                case 32: fn<32, 0>(ctx, __VA_ARGS__); break;                         \
                case 48: fn<48, 0>(ctx, __VA_ARGS__); break;                         \
                case 64: fn<64, 0>(ctx, __VA_ARGS__); break;                         \
                default: fn<128, 0>(ctx, __VA_ARGS__); break;                        \
            }                                                                        \
        } else if ((ctx)->loop_variant == 1) {                                       \
            switch ((ctx)->stack_entries) {                                          \
                case 24: fn<24, 1>(ctx, __VA_ARGS__); break;                         \
                case 32: fn<32, 1>(ctx, __VA_ARGS__); break;                         \
                case 40: fn<40, 1>(ctx, __VA_ARGS__); break;                         \
                case 48: fn<48, 1>(ctx, __VA_ARGS__); break;                         \
                case 64: fn<64, 1>(ctx, __VA_ARGS__); break;                         \
                default: fn<128, 1>(ctx, __VA_ARGS__); break;                        \
            }                                                                        \
        } else {                                                                     \
            switch ((ctx)->stack_entries) {                                          \
                case 24: fn<24, 2>(ctx, __VA_ARGS__); break;                         \
                case 32: fn<32, 2>(ctx, __VA_ARGS__); break;                         \
                case 40: fn<40, 2>(ctx, __VA_ARGS__); break;                         \
                case 48: fn<48, 2>(ctx, __VA_ARGS__); break;                         \
                case 64: fn<64, 2>(ctx, __VA_ARGS__); break;                         \
                default: fn<128, 2>(ctx, __VA_ARGS__); break;                        \
            }                                                                        \
        }                                                                            \
    } while (0)

int ensure_wavefront(pulse_ctx *ctx, size_t npix) {
    if (npix <= ctx->wf_pixels) return PULSE_OK;
    const size_t f4 = sizeof(float4) * npix;
    int rc;
    for (int b : {W_RAY_O0, W_RAY_O1, W_RAY_D0, W_RAY_D1, W_HIT, W_STATE_T, W_STATE_C, W_SH_O, W_SH_D, W_SH_C})
        if ((rc = ctx->wf[b].ensure(f4)) != PULSE_OK) return rc;
    if ((rc = ctx->wf[W_HIT_V].ensure(sizeof(float) * npix)) != PULSE_OK) return rc;
    if ((rc = ctx->wf[W_COUNTERS].ensure(sizeof(uint32_t) * kCounterWords)) != PULSE_OK) return rc;
    ctx->wf_pixels = npix;
    return PULSE_OK;
}

// The same buffers laid out in one region of `max_pixels` slots each (the
// allocation-free entry point's scratch).
WfBuffers wf_buffers_in(unsigned char *base, size_t max_pixels) {
    float4 *f4 = reinterpret_cast<float4 *>(base);
    WfBuffers wb;
    wb.ray_o[0] = f4 + 0 * max_pixels;
    wb.ray_o[1] = f4 + 1 * max_pixels;
    wb.ray_d[0] = f4 + 2 * max_pixels;
    wb.ray_d[1] = f4 + 3 * max_pixels;
    wb.hit = f4 + 4 * max_pixels;
    wb.state_t = f4 + 5 * max_pixels;
    wb.state_c = f4 + 6 * max_pixels;
    wb.sh_o = f4 + 7 * max_pixels;
    wb.sh_d = f4 + 8 * max_pixels;
    wb.sh_c = f4 + 9 * max_pixels;
    wb.hit_v = reinterpret_cast<float *>(f4 + kNoAllocF4 * max_pixels);
    wb.ext_count = reinterpret_cast<uint32_t *>(wb.hit_v + max_pixels);
    wb.sh_count = wb.ext_count + (kMaxTraces + 1);
    wb.ext_head = wb.ext_count + 2 * (kMaxTraces + 1);
    wb.sh_head = wb.ext_head + kHeadWords;
    return wb;
}

WfBuffers wf_buffers(pulse_ctx *ctx) {
    WfBuffers wb;
    wb.ray_o[0] = ctx->wf[W_RAY_O0].as<float4>();
    wb.ray_o[1] = ctx->wf[W_RAY_O1].as<float4>();
    wb.ray_d[0] = ctx->wf[W_RAY_D0].as<float4>();
    wb.ray_d[1] = ctx->wf[W_RAY_D1].as<float4>();
    wb.hit = ctx->wf[W_HIT].as<float4>();
    wb.hit_v = ctx->wf[W_HIT_V].as<float>();
    wb.state_t = ctx->wf[W_STATE_T].as<float4>();
    wb.state_c = ctx->wf[W_STATE_C].as<float4>();
    wb.sh_o = ctx->wf[W_SH_O].as<float4>();
    wb.sh_d = ctx->wf[W_SH_D].as<float4>();
    wb.sh_c = ctx->wf[W_SH_C].as<float4>();
    wb.ext_count = ctx->wf[W_COUNTERS].as<uint32_t>();
    wb.sh_count = wb.ext_count + (kMaxTraces + 1);
    wb.ext_head = wb.ext_count + 2 * (kMaxTraces + 1);
    wb.sh_head = wb.ext_head + kHeadWords;
    return wb;
}

}  // namespace

namespace pulse_host {

// The second stream and its events (PULSE_OPT_CONCURRENT_SHADOW), made on
// first use; default priority like the context's stream.
int ensure_stream2(pulse_ctx *ctx, bool timing) {
    if (!ctx->stream2) {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
        for (int i = 0; i < pt::kMaxTraces; ++i) {
            HIP_TRY(hipEventCreateWithFlags(&ctx->fork_ev[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming));
        }
    }
    if (timing && !ctx->events2_created) {
        for (auto &e : ctx->ev2) HIP_TRY(hipEventCreate(&e));
        ctx->events2_created = true;
    }
    return PULSE_OK;
}

constexpr int kHostThreads = 8;  // host threads for the per-instance TLAS boxes and the wide BLASes

// A depth-first traversal that pushes at most one far child per level never
// holds more entries than the depth of the node it is at.
uint32_t pick_stack(uint32_t need, int loop_variant) {
    static const uint32_t nested[] = {32, 48, 64, 128};
    static const uint32_t unified[] = {24, 32, 40, 48, 64, 128};
    if (loop_variant == 0) {
        for (uint32_t s : nested)
            if (s >= need) return s;
    } else {
        for (uint32_t s : unified)
            if (s >= need) return s;
    }
    return 0;
}

uint32_t stack_for(const SceneData &s, int loop_variant) {
    if (loop_variant == 2) return s.has_wide ? pick_stack(s.wide_need, 2) : 0;
    return pick_stack(s.stack_need, loop_variant);
}

int take_device_errors(pulse_ctx *ctx, uint32_t *out) {
    *out = 0;
    HIP_TRY(hipMemcpy(out, ctx->errors.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (*out != 0) HIP_TRY(hipMemset(ctx->errors.ptr, 0, sizeof(uint32_t)));
    return PULSE_OK;
}

DevScene dev_scene(const SceneData &sd, uint32_t *errors) {
    DevScene s;
    s.tri_bvh = sd.buf[B_TRI_BVH].as<float4>();
    s.nrm_bvh = sd.buf[B_NRM_BVH].as<float4>();
    s.primitives = sd.buf[B_PRIMS].as<pulse_primitive>();
    s.blas_nodes = sd.buf[B_BLAS].as<float4>();
    s.tlas_nodes = sd.buf[B_TLAS].as<float4>();
    s.instance_indices = sd.buf[B_INST_IDX].as<uint32_t>();
    s.instances = sd.buf[B_INSTANCES].as<pulse_mesh_instance>();
    s.materials = sd.buf[B_MATERIALS].as<pulse_material>();
    s.light_cdfs = sd.buf[B_CDFS].as<float>();
    s.light_areas = sd.buf[B_AREAS].as<float>();
    s.lights = sd.buf[B_LIGHTS].as<pulse_light_data_index>();
    s.strength_cdf = sd.buf[B_SCDF].as<float>();
    s.wide_tlas = sd.has_wide ? sd.buf[B_WTLAS].as<float4>() : nullptr;
    s.wide_blas = sd.has_wide ? sd.buf[B_WBLAS].as<float4>() : nullptr;
    s.wide_inst = sd.has_wide ? sd.buf[B_WINST].as<uint4>() : nullptr;
    s.wide_tri = sd.has_wide ? sd.buf[B_WTRI].as<float4>() : nullptr;
    s.instance_count = sd.uniform.instance_count;
    s.light_count = sd.uniform.light_count;
    s.error_flags = errors;
    s.fast_div = sd.fast_div;
    return s;
}

// Validation failures leave `sd` unchanged.
// The 4-wide traversal's instance records (kInstRecordWords per
// instance_indices entry, device/pt_device.h), from instance matrices and
// per-instance BLAS roots.  False if a matrix's bottom row is not 0 0 0 1
// (the records hold rows 0-2 only).  iidx entries are below n_inst.
bool instance_records(const pulse_mesh_instance *inst, size_t n_inst, const uint32_t *iidx, size_t n_iidx,
                      const uint32_t *root, std::vector<uint32_t> &out) {
    for (size_t i = 0; i < n_inst; ++i)
        if (!affine_matrix(inst[i].object_world) || !affine_matrix(inst[i].world_object)) return false;
    out.assign(kInstRecordWords * std::max<size_t>(n_iidx, 1), 0u);
    for (size_t k = 0; k < n_iidx; ++k) {
        const uint32_t i = iidx[k];
        uint32_t *r = &out[kInstRecordWords * k];
        record_matrix(inst[i].world_object, r + kRecWorldObject);
        r[kRecRoot] = root[i];
        r[kRecMissOverflows] = miss_overflows(inst[i].object_world, inst[i].world_object) ? 1u : 0u;
        record_matrix(inst[i].object_world, r + kRecObjectWorld);
        r[kRecIndexOffset] = inst[i].index_offset;
        r[kRecInstance] = i;
    }
    return true;
}

int scene_upload(SceneData &sd, const pulse_scene_desc *d) {
    if (!d) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_upload: null argument");
    int rc = PULSE_OK;
    auto count_of = [](const pulse_buffer_view &v, size_t elem, const char *name, size_t &n) -> int {
        if (v.bytes % elem != 0)
            PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_upload: %s has %zu bytes, not a multiple of %zu", name, v.bytes, elem);
        if (v.bytes && !v.data) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_upload: %s data is null", name);
        n = v.bytes / elem;
        return PULSE_OK;
    };
    size_t n_prim, n_tdata, n_tidx, n_blas, n_tlas, n_iidx, n_inst, n_mat, n_scdf, n_cdf, n_area, n_light;
    if ((rc = count_of(d->primitives, sizeof(pulse_primitive), "primitives", n_prim)) ||
        (rc = count_of(d->triangle_data, sizeof(pulse_triangle_data), "triangle_data", n_tdata)) ||
        (rc = count_of(d->triangle_indices, 4, "triangle_indices", n_tidx)) ||
        (rc = count_of(d->blas_nodes, sizeof(pulse_blas_node), "blas_nodes", n_blas)) ||
        (rc = count_of(d->tlas_nodes, sizeof(pulse_tlas_node), "tlas_nodes", n_tlas)) ||
        (rc = count_of(d->instance_indices, 4, "instance_indices", n_iidx)) ||
        (rc = count_of(d->instances, sizeof(pulse_mesh_instance), "instances", n_inst)) ||
        (rc = count_of(d->materials, sizeof(pulse_material), "materials", n_mat)) ||
        (rc = count_of(d->light_emission_strength_cdf, 4, "light_emission_strength_cdf", n_scdf)) ||
        (rc = count_of(d->light_triangle_area_cdfs, 4, "light_triangle_area_cdfs", n_cdf)) ||
        (rc = count_of(d->light_mesh_areas, 4, "light_mesh_areas", n_area)) ||
        (rc = count_of(d->light_indices, sizeof(pulse_light_data_index), "light_indices", n_light)))
        return rc;
    if (n_tlas == 0) PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_upload: tlas_nodes is empty (needs at least the root)");
    if (n_blas >= (1u << 27) || n_tlas >= (1u << 27))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_scene_upload: node arrays of 2^27 entries or more (traversal stack "
                   "entries hold node indices in 27 bits)");
    if (d->uniform.instance_count > n_inst)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_upload: instance_count %u > %zu instances", d->uniform.instance_count, n_inst);
    if (d->uniform.light_count > n_light || d->uniform.light_count > n_area)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_upload: light_count %u exceeds the light tables", d->uniform.light_count);
    if (n_tdata != n_prim) PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_upload: %zu triangle_data for %zu primitives", n_tdata, n_prim);

    const auto *prims = static_cast<const pulse_primitive *>(d->primitives.data);
    const auto *tdata = static_cast<const pulse_triangle_data *>(d->triangle_data.data);
    const auto *tidx = static_cast<const uint32_t *>(d->triangle_indices.data);
    const auto *blas = static_cast<const pulse_blas_node *>(d->blas_nodes.data);
    const auto *tlas = static_cast<const pulse_tlas_node *>(d->tlas_nodes.data);
    const auto *iidx = static_cast<const uint32_t *>(d->instance_indices.data);
    const auto *inst = static_cast<const pulse_mesh_instance *>(d->instances.data);
    const auto *lights = static_cast<const pulse_light_data_index *>(d->light_indices.data);

    // --- validate references and build the traversal-ordered copies ---
    std::vector<float4> tri_bvh(3 * std::max<size_t>(n_tidx, 1), make_float4(0, 0, 0, 0));
    std::vector<float4> nrm_bvh(3 * std::max<size_t>(n_tidx, 1), make_float4(0, 0, 0, 0));
    std::vector<int64_t> slot_owner(n_tidx, -1);  // triangle_offset that filled a slot
    // every edge component finite and at most 2^60: the 4-wide triangle test's
    // reciprocal (device/wide.h) is exact for such triangles
    bool edges_small = true;
    uint32_t blas_depth = 0;
    std::vector<uint32_t> seen_roots;
    for (size_t i = 0; i < n_inst; ++i) {
        const pulse_mesh_instance &mi = inst[i];
        if (mi.material_index >= n_mat)
            PULSE_FAIL(PULSE_E_LAYOUT, "instance %zu: material_index %u >= %zu", i, mi.material_index, n_mat);
        if (static_cast<uint64_t>(mi.triangle_offset) + mi.triangle_count > n_prim ||
            static_cast<uint64_t>(mi.index_offset) + mi.triangle_count > n_tidx || mi.node_offset >= n_blas ||
            mi.triangle_count == 0)
            PULSE_FAIL(PULSE_E_LAYOUT, "instance %zu: triangle/index/node range outside the buffers", i);
        if (std::find(seen_roots.begin(), seen_roots.end(), mi.node_offset) == seen_roots.end()) {
            seen_roots.push_back(mi.node_offset);
            uint32_t dep = 0;
            if (!walk_blas(blas, n_blas, mi.node_offset, mi.triangle_count, dep))
                PULSE_FAIL(PULSE_E_LAYOUT, "instance %zu: BLAS at node_offset %u has an out-of-range child or leaf", i,
                           mi.node_offset);
            blas_depth = std::max(blas_depth, dep);
        }
        for (uint32_t k = 0; k < mi.triangle_count; ++k) {
            size_t slot = mi.index_offset + k;
            if (slot_owner[slot] == static_cast<int64_t>(mi.triangle_offset)) continue;
            if (slot_owner[slot] != -1)
                PULSE_FAIL(PULSE_E_LAYOUT, "triangle_indices slot %zu is shared by meshes with different triangle_offset", slot);
            slot_owner[slot] = mi.triangle_offset;
            uint32_t local = tidx[slot];
            if (local >= mi.triangle_count)
                PULSE_FAIL(PULSE_E_LAYOUT, "triangle_indices[%zu] = %u >= mesh triangle count %u", slot, local, mi.triangle_count);
            const pulse_primitive &p = prims[mi.triangle_offset + local];
            // e1 = p1 - p0, e2 = p2 - p0 in f32 (utilities.wgsl:432-433)
            float e1x = p.p_second[0] - p.p_first[0], e1y = p.p_second[1] - p.p_first[1], e1z = p.p_second[2] - p.p_first[2];
            float e2x = p.p_third[0] - p.p_first[0], e2y = p.p_third[1] - p.p_first[1], e2z = p.p_third[2] - p.p_first[2];
            for (float e : {e1x, e1y, e1z, e2x, e2y, e2z}) edges_small = edges_small && std::fabs(e) <= 0x1p60f;
            tri_bvh[3 * slot] = make_float4(p.p_first[0], p.p_first[1], p.p_first[2], e1x);
            tri_bvh[3 * slot + 1] = make_float4(e1y, e1z, e2x, e2y);
            tri_bvh[3 * slot + 2] = make_float4(e2z, 0.f, 0.f, 0.f);
            const pulse_triangle_data &t = tdata[mi.triangle_offset + local];
            nrm_bvh[3 * slot] = make_float4(t.n_first[0], t.n_first[1], t.n_first[2], 0.f);
            nrm_bvh[3 * slot + 1] = make_float4(t.n_second[0], t.n_second[1], t.n_second[2], 0.f);
            nrm_bvh[3 * slot + 2] = make_float4(t.n_third[0], t.n_third[1], t.n_third[2], 0.f);
        }
    }
    uint32_t tlas_depth = 0;
    bool empty_tlas = tlas[0].a_or_first_instance == 0 && tlas[0].instance_count == 0;
    if (!empty_tlas) {
        if (!walk_tlas(tlas, n_tlas, n_iidx, tlas_depth))
            PULSE_FAIL(PULSE_E_LAYOUT, "tlas_nodes: child index or instance range out of range");
        for (size_t k = 0; k < n_iidx; ++k)
            if (iidx[k] >= n_inst) PULSE_FAIL(PULSE_E_LAYOUT, "instance_indices[%zu] = %u >= %zu", k, iidx[k], n_inst);
    }
    for (size_t l = 0; l < d->uniform.light_count; ++l) {
        if (lights[l].mesh_instance_index >= n_inst)
            PULSE_FAIL(PULSE_E_LAYOUT, "light %zu: mesh_instance_index out of range", l);
        const pulse_mesh_instance &mi = inst[lights[l].mesh_instance_index];
        if (static_cast<uint64_t>(lights[l].cdf_offset) + mi.triangle_count > n_cdf)
            PULSE_FAIL(PULSE_E_LAYOUT, "light %zu: cdf range outside light_triangle_area_cdfs", l);
    }
    uint32_t need = tlas_depth + blas_depth + 1;
    // node coordinates: 0 or magnitude in [2^-56, 2^40] for the fast division
    auto coord_ok = [](float c) {
        float a = std::fabs(c);
        return c == 0.0f || (a >= 0x1p-56f && a <= 0x1p40f);
    };
    bool fast_div = true;
    for (size_t k = 0; k < n_blas && fast_div; ++k)
        for (int c = 0; c < 3; ++c) fast_div = fast_div && coord_ok(blas[k].aabb_min[c]) && coord_ok(blas[k].aabb_max[c]);
    const bool fast_div_blas = fast_div;
    for (size_t k = 0; k < n_tlas && fast_div; ++k)
        for (int c = 0; c < 3; ++c) fast_div = fast_div && coord_ok(tlas[k].aabb_min[c]) && coord_ok(tlas[k].aabb_max[c]);
    if (pick_stack(need, 0) == 0 && pick_stack(need, 1) == 0)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "BVH too deep: TLAS depth %u + BLAS depth %u exceeds the 128-entry stack",
                   tlas_depth, blas_depth);

    // --- 4-wide BVHs over the same leaves (traversal loop 2), quantised ---
    // instance boxes for its TLAS: the transformed BLAS cut boxes
    std::vector<float> cut_boxes, ibox(6 * std::max<size_t>(n_inst, 1));
    std::vector<uint32_t> cut_range(2 * std::max<size_t>(n_inst, 1), 0u);
    {
        std::vector<std::array<uint32_t, 3>> seen;  // node_offset, first box, count
        for (size_t i = 0; i < n_inst; ++i) {
            const uint32_t off = inst[i].node_offset;
            auto it = std::find_if(seen.begin(), seen.end(), [&](const auto &p) { return p[0] == off; });
            if (it == seen.end()) {
                const uint32_t first = static_cast<uint32_t>(cut_boxes.size() / 6);
                blas_cut_boxes(reinterpret_cast<const uint32_t *>(blas + off), kCutDepth, cut_boxes);
                seen.push_back({off, first, static_cast<uint32_t>(cut_boxes.size() / 6) - first});
                it = seen.end() - 1;
            }
            cut_range[2 * i] = (*it)[1];
            cut_range[2 * i + 1] = (*it)[2];
        }
        if (n_inst)
            instance_world_boxes(inst[0].object_world, sizeof(pulse_mesh_instance) / sizeof(float), cut_boxes.data(),
                                 cut_range.data(), n_inst, ibox.data(), kHostThreads);
    }
    WideTree wtlas;
    std::vector<uint32_t> qtlas, wblas;
    bool has_wide = edges_small && (empty_tlas ? collapse_wide(reinterpret_cast<const uint32_t *>(tlas), n_tlas, 0u, wtlas)
                                               : wide_tlas(reinterpret_cast<const uint32_t *>(tlas), iidx, ibox.data(), wtlas));
    has_wide = has_wide && quantize_wide(wtlas, qtlas);
    if (cut_boxes.empty()) cut_boxes.assign(6, 0.0f);
    std::vector<uint32_t> wroot(std::max<size_t>(n_inst, 1), 0u);
    uint32_t wide_blas_need = 0;
    // the refined leaves' triangle order: wperm[g] = tri_bvh slot at wide slot g
    std::vector<uint32_t> wperm(std::max<size_t>(n_tidx, 1)), leaf_end(std::max<size_t>(n_tidx, 1), 0u);
    for (size_t g = 0; g < wperm.size(); ++g) wperm[g] = static_cast<uint32_t>(g);
    if (has_wide) {
        // one tree per (node_offset, index_offset): leaf entries hold global slots
        std::vector<std::array<uint32_t, 2>> trees;
        std::vector<uint32_t> tree_of(n_inst);
        for (size_t i = 0; i < n_inst; ++i) {
            const std::array<uint32_t, 2> key{inst[i].node_offset, inst[i].index_offset};
            auto it = std::find(trees.begin(), trees.end(), key);
            tree_of[i] = static_cast<uint32_t>(it - trees.begin());
            if (it == trees.end()) trees.push_back(key);
        }
        std::vector<uint8_t> refs(n_tidx, 0);
        for (const auto &t : trees)
            count_leaf_refs(reinterpret_cast<const uint32_t *>(blas + t[0]), n_blas - t[0], t[1], refs, &leaf_end);
        std::vector<std::vector<uint32_t>> q(trees.size());
        std::vector<uint32_t> t_need(trees.size(), 0);
        std::vector<char> t_ok(trees.size(), 0);
        std::atomic<size_t> next{0};
        auto work = [&]() {
            std::vector<uint32_t> nodes;
            for (size_t k; (k = next.fetch_add(1)) < trees.size();) {
                refine_leaves(reinterpret_cast<const uint32_t *>(blas + trees[k][0]), n_blas - trees[k][0], trees[k][1],
                              reinterpret_cast<const float *>(tri_bvh.data()), kRefineLeaf, refs, nodes, wperm.data());
                WideTree w;
                t_ok[k] = collapse_wide(nodes.data(), nodes.size() / 8, trees[k][1], w) && quantize_wide(w, q[k]);
                t_need[k] = w.stack_need;
            }
        };
        std::vector<std::thread> pool;
        const size_t nth = std::min<size_t>(kHostThreads, trees.size());
        for (size_t k = 1; k < nth; ++k) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        std::vector<uint32_t> root(trees.size());
        for (size_t k = 0; k < trees.size() && has_wide; ++k) {
            has_wide = t_ok[k] != 0;
            root[k] = static_cast<uint32_t>(wblas.size() / kQuantNodeWords);
            wblas.insert(wblas.end(), q[k].begin(), q[k].end());
            wide_blas_need = std::max(wide_blas_need, t_need[k]);
        }
        for (size_t i = 0; i < n_inst && has_wide; ++i) wroot[i] = root[tree_of[i]];
    }
    if (wblas.empty()) wblas.assign(kQuantNodeWords, 0u);
    // wide_tri: tri_bvh's rows in the refined order, each with its tri_bvh
    // slot and its reference leaf's end slot in the third row
    std::vector<float4> wtri;
    if (has_wide) {
        wtri = tri_bvh;
        for (size_t g = 0; g < n_tidx; ++g) {
            const size_t p = wperm[g];
            wtri[3 * g] = tri_bvh[3 * p];
            wtri[3 * g + 1] = tri_bvh[3 * p + 1];
            uint32_t bits[2] = {static_cast<uint32_t>(p), leaf_end[p]};
            float f[2];
            std::memcpy(f, bits, 8);
            wtri[3 * g + 2] = make_float4(tri_bvh[3 * p + 2].x, f[0], f[1], 0.0f);
        }
    }
    std::vector<uint32_t> wrec;
    if (has_wide && !instance_records(inst, n_inst, iidx, n_iidx, wroot.data(), wrec)) has_wide = false;
    const uint32_t wide_need = wtlas.stack_need + wide_blas_need + 1;
    if (pick_stack(wide_need, 2) == 0) has_wide = false;

    // --- copy to device ---
    sd.valid = false;
    auto up = [&](int b, const void *src, size_t bytes) -> int {
        int r = sd.buf[b].ensure(bytes);
        if (r) return r;
        if (bytes) HIP_TRY(hipMemcpy(sd.buf[b].ptr, src, bytes, hipMemcpyHostToDevice));
        return PULSE_OK;
    };
    if ((rc = up(B_PRIMS, prims, d->primitives.bytes)) || (rc = up(B_TRI_BVH, tri_bvh.data(), tri_bvh.size() * sizeof(float4))) ||
        (rc = up(B_NRM_BVH, nrm_bvh.data(), nrm_bvh.size() * sizeof(float4))) || (rc = up(B_BLAS, blas, d->blas_nodes.bytes)) ||
        (rc = up(B_TLAS, tlas, d->tlas_nodes.bytes)) || (rc = up(B_INST_IDX, iidx, d->instance_indices.bytes)) ||
        (rc = up(B_INSTANCES, inst, d->instances.bytes)) || (rc = up(B_MATERIALS, d->materials.data, d->materials.bytes)) ||
        (rc = up(B_CDFS, d->light_triangle_area_cdfs.data, d->light_triangle_area_cdfs.bytes)) ||
        (rc = up(B_AREAS, d->light_mesh_areas.data, d->light_mesh_areas.bytes)) ||
        (rc = up(B_LIGHTS, lights, d->light_indices.bytes)) ||
        (rc = up(B_SCDF, d->light_emission_strength_cdf.data, d->light_emission_strength_cdf.bytes)))
        return rc;
    // room for a device rebuild's tree (at most one node per internal node of
    // a binary tree over the entries), so that one never moves the buffer
    if (has_wide && n_iidx <= kRebuildMax &&
        (rc = sd.buf[B_WTLAS].ensure(std::max<size_t>(qtlas.size() * 4,
                                                      64 * std::max<size_t>(n_iidx > 0 ? n_iidx - 1 : 1, 1)))))
        return rc;
    if (has_wide && ((rc = up(B_WTLAS, qtlas.data(), qtlas.size() * 4)) ||
                     (rc = up(B_WBLAS, wblas.data(), wblas.size() * 4)) ||
                     (rc = up(B_WINST, wrec.data(), wrec.size() * 4)) ||
                     (rc = up(B_WTRI, wtri.data(), wtri.size() * sizeof(float4)))))
        return rc;
    {
        auto up_refit = [&](int b, const void *src, size_t bytes) -> int {
            int r = sd.refit[b].ensure(bytes);
            if (r) return r;
            HIP_TRY(hipMemcpy(sd.refit[b].ptr, src, bytes, hipMemcpyHostToDevice));
            return PULSE_OK;
        };
        std::vector<uint32_t> levels;
        wide_levels(wtlas, levels);
        if ((rc = up_refit(5, cut_boxes.data(), cut_boxes.size() * 4)) ||
            (rc = up_refit(6, cut_range.data(), cut_range.size() * 4)) ||
            (rc = up_refit(7, levels.data(), levels.size() * 4)))
            return rc;
    }
    sd.cut_boxes = std::move(cut_boxes);
    sd.cut_range = std::move(cut_range);
    sd.wide_root = std::move(wroot);
    sd.n_iidx = static_cast<uint32_t>(n_iidx);
    sd.tlas_all_entries = !empty_tlas && tlas_covers_entries(tlas, n_iidx);
    sd.has_wide = has_wide;
    sd.wide_need = wide_need;
    sd.wide_blas_need = wide_blas_need;
    sd.n_wide_tlas = static_cast<uint32_t>(qtlas.size() / kQuantNodeWords);
    sd.binary_stale = false;
    sd.uniform = d->uniform;
    sd.n_strength_cdf = static_cast<uint32_t>(n_scdf);
    sd.stack_need = need;
    sd.n_materials = static_cast<uint32_t>(n_mat);
    sd.fast_div = fast_div;
    sd.fast_div_blas = fast_div_blas;
    sd.tlas_depth = tlas_depth;
    sd.blas_depth = blas_depth;
    sd.inst_ranges.resize(4 * n_inst);
    for (size_t i = 0; i < n_inst; ++i) {
        sd.inst_ranges[4 * i] = inst[i].triangle_offset;
        sd.inst_ranges[4 * i + 1] = inst[i].triangle_count;
        sd.inst_ranges[4 * i + 2] = inst[i].index_offset;
        sd.inst_ranges[4 * i + 3] = inst[i].node_offset;
    }
    sd.light_indices.assign(lights, lights + n_light);
    sd.generation += 1;
    sd.valid = true;
    return PULSE_OK;
}

int scene_update_instances(SceneData &sd, const pulse_scene_desc *d) {
    if (!d) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_update_instances: null argument");
    if (!sd.valid) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_scene_update_instances: no scene uploaded");
    const size_t n_inst = d->instances.bytes / sizeof(pulse_mesh_instance);
    const size_t n_tlas = d->tlas_nodes.bytes / sizeof(pulse_tlas_node);
    const size_t n_iidx = d->instance_indices.bytes / 4;
    if (d->instances.bytes % sizeof(pulse_mesh_instance) || d->tlas_nodes.bytes % sizeof(pulse_tlas_node) ||
        d->instance_indices.bytes % 4 || (d->instances.bytes && !d->instances.data) || !d->tlas_nodes.data ||
        (d->instance_indices.bytes && !d->instance_indices.data) || n_tlas == 0)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_update_instances: instances / tlas_nodes / instance_indices layout");
    if (n_tlas >= (1u << 27)) PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_scene_update_instances: TLAS of 2^27 nodes or more");
    if (4 * n_inst != sd.inst_ranges.size() || d->uniform.instance_count != sd.uniform.instance_count ||
        d->uniform.light_count != sd.uniform.light_count)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_update_instances: instance or light count changed (upload the scene)");
    const auto *inst = static_cast<const pulse_mesh_instance *>(d->instances.data);
    for (size_t i = 0; i < n_inst; ++i) {
        if (inst[i].triangle_offset != sd.inst_ranges[4 * i] || inst[i].triangle_count != sd.inst_ranges[4 * i + 1] ||
            inst[i].index_offset != sd.inst_ranges[4 * i + 2] || inst[i].node_offset != sd.inst_ranges[4 * i + 3])
            PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_update_instances: instance %zu changed mesh (upload the scene)", i);
        if (inst[i].material_index >= sd.n_materials)
            PULSE_FAIL(PULSE_E_LAYOUT, "instance %zu: material_index %u >= %u", i, inst[i].material_index, sd.n_materials);
    }
    if (d->light_indices.bytes != sd.light_indices.size() * sizeof(pulse_light_data_index) ||
        (d->light_indices.bytes &&
         std::memcmp(d->light_indices.data, sd.light_indices.data(), d->light_indices.bytes) != 0))
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_update_instances: the light table changed (upload the scene)");
    const auto *tlas = static_cast<const pulse_tlas_node *>(d->tlas_nodes.data);
    const auto *iidx = static_cast<const uint32_t *>(d->instance_indices.data);
    uint32_t tlas_depth = 0;
    bool empty_tlas = tlas[0].a_or_first_instance == 0 && tlas[0].instance_count == 0;
    if (!empty_tlas) {
        if (!walk_tlas(tlas, n_tlas, n_iidx, tlas_depth))
            PULSE_FAIL(PULSE_E_LAYOUT, "tlas_nodes: child index or instance range out of range");
        for (size_t k = 0; k < n_iidx; ++k)
            if (iidx[k] >= n_inst) PULSE_FAIL(PULSE_E_LAYOUT, "instance_indices[%zu] = %u >= %zu", k, iidx[k], n_inst);
    }
    const uint32_t need = tlas_depth + sd.blas_depth + 1;
    if (pick_stack(need, 0) == 0 && pick_stack(need, 1) == 0)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "BVH too deep: TLAS depth %u + BLAS depth %u exceeds the 128-entry stack",
                   tlas_depth, sd.blas_depth);
    auto coord_ok = [](float c) {
        float a = std::fabs(c);
        return c == 0.0f || (a >= 0x1p-56f && a <= 0x1p40f);
    };
    bool fast_div = sd.fast_div_blas;
    for (size_t k = 0; k < n_tlas && fast_div; ++k)
        for (int c = 0; c < 3; ++c) fast_div = fast_div && coord_ok(tlas[k].aabb_min[c]) && coord_ok(tlas[k].aabb_max[c]);
    auto up = [&](int b, const void *src, size_t bytes) -> int {
        int r = sd.buf[b].ensure(bytes);
        if (r) return r;
        if (bytes) HIP_TRY(hipMemcpy(sd.buf[b].ptr, src, bytes, hipMemcpyHostToDevice));
        return PULSE_OK;
    };
    WideTree wtlas;
    std::vector<uint32_t> qtlas;
    bool has_wide = sd.has_wide;
    if (has_wide && empty_tlas) {
        has_wide = collapse_wide(reinterpret_cast<const uint32_t *>(tlas), n_tlas, 0u, wtlas);
    } else if (has_wide) {
        std::vector<float> ibox(6 * n_inst);
        instance_world_boxes(inst[0].object_world, sizeof(pulse_mesh_instance) / sizeof(float), sd.cut_boxes.data(),
                             sd.cut_range.data(), n_inst, ibox.data(), kHostThreads);
        has_wide = wide_tlas(reinterpret_cast<const uint32_t *>(tlas), iidx, ibox.data(), wtlas);
    }
    has_wide = has_wide && quantize_wide(wtlas, qtlas);
    std::vector<uint32_t> wrec;
    if (has_wide && !instance_records(inst, n_inst, iidx, n_iidx, sd.wide_root.data(), wrec))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_scene_update_instances: an instance matrix's bottom row is not 0 0 0 1; "
                   "the scene's 4-wide BVHs hold affine transforms only (upload the scene instead)");
    const uint32_t wide_need = wtlas.stack_need + sd.wide_blas_need + 1;
    if (pick_stack(wide_need, 2) == 0) has_wide = false;
    int rc;
    std::vector<uint32_t> levels;
    if (has_wide) wide_levels(wtlas, levels);
    auto up_levels = [&]() -> int {
        int r = sd.refit[7].ensure(levels.size() * 4);
        if (r) return r;
        HIP_TRY(hipMemcpy(sd.refit[7].ptr, levels.data(), levels.size() * 4, hipMemcpyHostToDevice));
        return PULSE_OK;
    };
    if ((rc = up(B_INSTANCES, inst, d->instances.bytes)) || (rc = up(B_TLAS, tlas, d->tlas_nodes.bytes)) ||
        (rc = up(B_INST_IDX, iidx, d->instance_indices.bytes)) ||
        (has_wide && ((rc = up(B_WTLAS, qtlas.data(), qtlas.size() * 4)) ||
                      (rc = up(B_WINST, wrec.data(), wrec.size() * 4)) || (rc = up_levels())))) {
        sd.valid = false;  // partially replaced: unusable until the next upload
        return rc;
    }
    sd.has_wide = has_wide;
    sd.wide_need = wide_need;
    sd.n_wide_tlas = static_cast<uint32_t>(qtlas.size() / kQuantNodeWords);
    sd.n_iidx = static_cast<uint32_t>(n_iidx);
    sd.tlas_all_entries = !empty_tlas && tlas_covers_entries(tlas, n_iidx);
    sd.binary_stale = false;
    sd.tlas_depth = tlas_depth;
    sd.stack_need = need;
    sd.fast_div = fast_div;
    sd.generation += 1;
    return PULSE_OK;
}

int check_traversal(const SceneData &s, int loop) {
    if (loop != 2 && s.binary_stale)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "traversal loop %d: the scene's binary TLAS is out of date after a device "
                   "refit (pulse_scene_refit_instances); use loop 2 or pulse_scene_update_instances", loop);
    return PULSE_OK;
}

// Instance motion with a new 4-wide TLAS built on the device
// (k_rebuild_tlas) instead of a refit of the last one.
int scene_rebuild_instances(SceneData &sd, const float *object_world, uint32_t n, hipStream_t stream) {
    const char *who = "pulse_scene_rebuild_instances";
    if (!sd.valid) PULSE_FAIL(PULSE_E_NO_SCENE, "%s: no scene uploaded", who);
    if (!sd.has_wide) PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: the scene has no 4-wide BVHs", who);
    if (n != sd.uniform.instance_count || 4ull * n != sd.inst_ranges.size())
        PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: %u matrices for %u instances", who, n, sd.uniform.instance_count);
    if (n == 0) return PULSE_OK;
    if (!object_world) PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: null matrices", who);
    if (!sd.tlas_all_entries)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: the scene's TLAS does not hold every instance_indices entry once", who);
    const uint32_t N = sd.n_iidx;
    if (N == 0 || N > kRebuildMax)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: %u instance_indices entries (the device rebuild takes 1..%u; "
                   "pulse_scene_refit_instances or pulse_scene_update_instances otherwise)", who, N, kRebuildMax);
    const size_t M1 = std::max<uint32_t>(N - 1, 1);
    if (sd.buf[B_WTLAS].cap < 64 * M1)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: the TLAS buffer has no room for a rebuilt tree", who);
    int rc;
    if ((rc = sd.refit[0].ensure(sizeof(float) * 16 * n)) || (rc = sd.refit[1].ensure(sizeof(float) * 6 * n)) ||
        (rc = sd.refit[2].ensure(sizeof(float) * 24 * M1)) || (rc = sd.refit[4].ensure(sizeof(uint32_t))) ||
        (rc = sd.refit[7].ensure(sizeof(uint32_t) * (2 + 2 * M1))) ||
        (rc = sd.refit[8].ensure(sizeof(uint32_t) * kRebuildScratchWords * N)) ||
        (rc = sd.refit[9].ensure(sizeof(uint32_t) * 2)))
        return rc;
    HIP_TRY(hipStreamSynchronize(stream));  // launches in flight read the current instances and TLAS
    HIP_TRY(hipMemcpy(sd.refit[0].ptr, object_world, sizeof(float) * 16 * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(sd.refit[4].ptr, 0, sizeof(uint32_t), stream));
    uint32_t *err = sd.refit[4].as<uint32_t>();
    hipLaunchKernelGGL(k_refit_instances, dim3(n), dim3(64), 0, stream, sd.refit[0].as<float>(),
                       sd.buf[B_INSTANCES].as<pulse_mesh_instance>(), sd.refit[5].as<float>(),
                       sd.refit[6].as<uint32_t>(), sd.refit[1].as<float>(), n, err);
    hipLaunchKernelGGL(k_refit_records, dim3((N + 63) / 64), dim3(64), 0, stream,
                       sd.buf[B_INSTANCES].as<pulse_mesh_instance>(), sd.buf[B_INST_IDX].as<uint32_t>(), N,
                       sd.buf[B_WINST].as<uint32_t>(), err);
    hipLaunchKernelGGL(k_rebuild_tlas, dim3(1), dim3(kRebuildThreads), 0, stream, sd.buf[B_INST_IDX].as<uint32_t>(), N,
                       sd.refit[1].as<float>(), sd.refit[8].as<uint32_t>(), sd.refit[2].as<float>(),
                       sd.buf[B_WTLAS].as<uint32_t>(), sd.refit[7].as<uint32_t>(), sd.refit[9].as<uint32_t>(), err);
    HIP_TRY(hipGetLastError());
    uint32_t flags = 0, info[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&flags, err, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(info, sd.refit[9].ptr, sizeof(info), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (flags || info[0] == 0 || info[0] > M1 || info[1] == 0 || info[1] > info[0]) {
        sd.valid = false;  // partly rewritten: unusable until the next upload
        PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: a singular transform, a bottom row other than 0 0 0 1, or a world box "
                   "beyond 2^40 (the scene must be uploaded again)", who);
    }
    sd.n_wide_tlas = info[0];
    // a node pushes at most 3 entries per level of the tree
    sd.wide_need = 3 * info[1] + sd.wide_blas_need + 1;
    sd.binary_stale = true;
    sd.generation += 1;
    return PULSE_OK;
}

int scene_refit_instances(SceneData &sd, const float *object_world, uint32_t n, hipStream_t stream) {
    if (!sd.valid) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_scene_refit_instances: no scene uploaded");
    if (!sd.has_wide)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_scene_refit_instances: the scene has no 4-wide BVHs to refit");
    if (n != sd.uniform.instance_count || 4ull * n != sd.inst_ranges.size())
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_refit_instances: %u matrices for %u instances", n,
                   sd.uniform.instance_count);
    if (n == 0) return PULSE_OK;
    if (!object_world) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_refit_instances: null matrices");
    const uint32_t nodes = sd.n_wide_tlas;
    int rc;
    if ((rc = sd.refit[0].ensure(sizeof(float) * 16 * n)) || (rc = sd.refit[1].ensure(sizeof(float) * 6 * n)) ||
        (rc = sd.refit[2].ensure(sizeof(float) * 24 * nodes)) || (rc = sd.refit[3].ensure(sizeof(float) * 6 * nodes)) ||
        (rc = sd.refit[4].ensure(sizeof(uint32_t))))
        return rc;
    HIP_TRY(hipStreamSynchronize(stream));  // launches in flight read the current instances
    HIP_TRY(hipMemcpy(sd.refit[0].ptr, object_world, sizeof(float) * 16 * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(sd.refit[4].ptr, 0, sizeof(uint32_t), stream));
    uint32_t *err = sd.refit[4].as<uint32_t>();
    hipLaunchKernelGGL(k_refit_instances, dim3(n), dim3(64), 0, stream, sd.refit[0].as<float>(),
                       sd.buf[B_INSTANCES].as<pulse_mesh_instance>(), sd.refit[5].as<float>(),
                       sd.refit[6].as<uint32_t>(), sd.refit[1].as<float>(), n, err);
    if (sd.n_iidx)
        hipLaunchKernelGGL(k_refit_records, dim3((sd.n_iidx + 63) / 64), dim3(64), 0, stream,
                           sd.buf[B_INSTANCES].as<pulse_mesh_instance>(), sd.buf[B_INST_IDX].as<uint32_t>(),
                           sd.n_iidx, sd.buf[B_WINST].as<uint32_t>(), err);
    hipLaunchKernelGGL(k_refit_boxes, dim3(1), dim3(256), 0, stream, sd.buf[B_WTLAS].as<uint32_t>(),
                       sd.refit[7].as<uint32_t>(), sd.buf[B_INST_IDX].as<uint32_t>(), sd.refit[1].as<float>(),
                       sd.refit[2].as<float>(), sd.refit[3].as<float>());
    hipLaunchKernelGGL(k_refit_quantize, dim3((nodes + 63) / 64), dim3(64), 0, stream, sd.buf[B_WTLAS].as<uint32_t>(),
                       nodes, sd.refit[2].as<float>(), err);
    HIP_TRY(hipGetLastError());
    uint32_t flags = 0;
    HIP_TRY(hipMemcpyAsync(&flags, err, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (flags) {
        sd.valid = false;  // partly rewritten: unusable until the next upload
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_refit_instances: a singular transform, a bottom row other than "
                   "0 0 0 1, or a world box beyond 2^40 (the scene must be uploaded again)");
    }
    sd.binary_stale = true;
    sd.generation += 1;
    return PULSE_OK;
}

}  // namespace pulse_host

namespace {
constexpr int kAdjBlock = 128;
// The view table travels in the kernel arguments (no allocation, nothing to
// copy before the launch); HIP takes at most 4 KB of them.
static_assert(sizeof(AdjArgs) + sizeof(DevScene) <= 4096, "adjoint kernel arguments exceed 4 KB");
template <int STACK, int LOOP>
void launch_adjoint_loop(pulse_ctx *ctx, const AdjArgs &a, size_t lds, uint32_t grid) {
    if (a.vertices > static_cast<uint32_t>(kAdjMaxVertices)) {
        // deep paths: fixed-direction roughness derivative, the 64-entry stack
        hipLaunchKernelGGL((k_adjoint<kAdjBlock, 64, true, LOOP, kAdjMaxVerticesDeep>), dim3(grid), dim3(kAdjBlock),
                           lds, ctx->stream, dev_scene(ctx), a);
        return;
    }
    if (a.detached)
        hipLaunchKernelGGL((k_adjoint<kAdjBlock, STACK, true, LOOP>), dim3(grid), dim3(kAdjBlock), lds, ctx->stream,
                           dev_scene(ctx), a);
    else
        hipLaunchKernelGGL((k_adjoint<kAdjBlock, STACK, false, LOOP>), dim3(grid), dim3(kAdjBlock), lds, ctx->stream,
                           dev_scene(ctx), a);
}
// The forward pass's traversal: the 4-wide BVHs for loop 2, the binary ones
// (same results as loops 0 and 1) otherwise.
template <int STACK>
void launch_adjoint(pulse_ctx *ctx, const AdjArgs &a, size_t lds, uint32_t grid) {
    if (ctx->loop_variant == 2) launch_adjoint_loop<STACK, 2>(ctx, a, lds, grid);
    else launch_adjoint_loop<STACK, 1>(ctx, a, lds, grid);
}

// What the adjoint supports (params, scene, context); `who` names the entry
// point in the message.
int adjoint_check(pulse_ctx *ctx, const pulse_params *prm, const char *who) {
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "%s: no scene uploaded", who);
    if (int rc_ = check_traversal(*ctx->cur, ctx->loop_variant)) return rc_;
    if (prm->mode != PULSE_MODE_GGX_NEE && prm->mode != PULSE_MODE_REF_GGX && prm->mode != PULSE_MODE_REF_NEE_LAMBERT)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: unknown mode %u", who, prm->mode);
    // path vertices: the Lambert mode's first hit comes before its max_depth bounces
    const uint32_t extra = prm->mode == PULSE_MODE_REF_NEE_LAMBERT ? 1u : 0u;
    if (prm->max_depth + extra == 0 || prm->max_depth > static_cast<uint32_t>(kAdjMaxVerticesDeep) - extra)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: max_depth must be %u..%u in this mode", who, 1u - extra,
                   static_cast<uint32_t>(kAdjMaxVerticesDeep) - extra);
    if (prm->max_depth + extra > static_cast<uint32_t>(kAdjMaxVertices) && (prm->flags & PULSE_FLAG_ADJOINT_REPARAM))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: PULSE_FLAG_ADJOINT_REPARAM needs max_depth <= %u in this mode", who,
                   static_cast<uint32_t>(kAdjMaxVertices) - extra);
    if (prm->spp == 0 || prm->spp > 0xffffffffu - prm->accumulation_count)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: needs spp >= 1 and accumulation_count + spp below 2^32", who);
    if (ctx->stack_entries > 64) PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: BVH too deep for the adjoint stack", who);
    return PULSE_OK;
}

// The adjoint over views[0..n): one launch per kAdjMaxViews views, every
// launch adding to `grad` (zeroed first unless the params ask to accumulate).
// The views have been checked (pointers, sizes).
int adjoint_run(pulse_ctx *ctx, const pulse_params *prm, const pulse_adjoint_view *views, uint32_t n, float *grad) {
    int rc = set_device(ctx);
    if (rc) return rc;
    const uint32_t nent = ctx->cur->n_materials * kGradPerMaterial;
    if (nent && !(prm->flags & PULSE_FLAG_ADJOINT_ACCUMULATE)) {
        hipLaunchKernelGGL(k_fill_zero, dim3((nent + 255) / 256), dim3(256), 0, ctx->stream, grad, nent);
    }
    AdjArgs a{};
    a.lambert = prm->mode == PULSE_MODE_REF_NEE_LAMBERT ? 1u : 0u;
    a.vertices = prm->max_depth + a.lambert;
    a.nee_ggx = prm->nee_eval == PULSE_NEE_GGX && !a.lambert ? 1u : 0u;
    a.spp = prm->accumulation_count + prm->spp;  // every sample the accumulated image holds
    a.light_power = (prm->flags & PULSE_FLAG_LIGHT_POWER) ? 1u : 0u;
    a.ref_ggx = prm->mode == PULSE_MODE_REF_GGX ? 1u : 0u;
    a.grad = grad;
    a.n_materials = ctx->cur->n_materials;
    a.horizon_clamp = (prm->flags & PULSE_FLAG_HORIZON_CLAMP) ? 1u : 0u;
    a.detached = (prm->flags & PULSE_FLAG_ADJOINT_REPARAM) ? 0u : 1u;
    a.rr_start = prm->rr_start_depth;
    ctx->rec_read = true;
    const size_t lds_bytes = static_cast<size_t>(nent) * sizeof(float);
    a.use_lds = lds_bytes <= 16384 ? 1u : 0u;
    const size_t lds = a.use_lds ? lds_bytes : 0;
    const int rounds = ctx->adjoint_reduce_rounds;
    a.reduce_rounds = rounds >= 0 ? static_cast<uint32_t>(rounds) : (a.use_lds ? 0u : 2u);
    for (uint32_t first = 0; first < n; first += kAdjMaxViews) {
        const uint32_t cnt = std::min<uint32_t>(n - first, kAdjMaxViews);
        uint64_t blocks = 0;
        a.n_views = cnt;
        for (uint32_t i = 0; i < cnt; ++i) {
            const pulse_adjoint_view &v = views[first + i];
            AdjView &w = a.views[i];
            w.cam = v.camera;
            w.image = reinterpret_cast<const float4 *>(v.image);
            w.target = reinterpret_cast<const float4 *>(v.target);
            w.replay = reinterpret_cast<float4 *>(v.replay_rgba);
            const uint64_t npix = static_cast<uint64_t>(v.camera.width) * v.camera.height;
            if (npix > 0xffffffffull) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint: view of %llu pixels",
                                                 static_cast<unsigned long long>(npix));
            w.grad_scale = static_cast<float>(2.0 / (3.0 * static_cast<double>(npix)));
            w.npix = static_cast<uint32_t>(npix);
            // the forward pass's recorded paths, when they are this view's image
            w.rec_hit = nullptr;
            w.rec_v = nullptr;
            w.rec_vis = nullptr;
            for (auto &r : ctx->recs)
                if (r.valid && record_matches(ctx, r, v.camera, *prm) && r.samples >= a.spp &&
                    r.traces == a.vertices && r.npix == npix) {
                    w.rec_hit = r.hit.as<float4>();
                    w.rec_v = r.v.as<float>();
                    w.rec_vis = r.vis.as<uint8_t>();
                    r.used = ++ctx->rec_clock;
                }
            if (!w.rec_hit) ctx->rec_read = false;
            w.block0 = static_cast<uint32_t>(blocks);
            blocks += (npix + kAdjBlockPixels - 1) / kAdjBlockPixels;
        }
        if (blocks > 0x7fffffffu) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint: %llu blocks in one launch",
                                             static_cast<unsigned long long>(blocks));
        const uint32_t grid = static_cast<uint32_t>(blocks);
        switch (ctx->stack_entries) {
            case 24: launch_adjoint<24>(ctx, a, lds, grid); break;
            case 32: launch_adjoint<32>(ctx, a, lds, grid); break;
            case 40: launch_adjoint<40>(ctx, a, lds, grid); break;
            case 48: launch_adjoint<48>(ctx, a, lds, grid); break;
            default: launch_adjoint<64>(ctx, a, lds, grid); break;
        }
        HIP_TRY(hipGetLastError());
    }
    return PULSE_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *pulse_version(void) { return "pulse_hip 0.1 (gfx950)"; }

int pulse_create(int device, pulse_ctx **out) {
    if (!out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_create: null out");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_create: device %d of %d", device, n);
    auto *ctx = new pulse_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        delete ctx;
        PULSE_FAIL(PULSE_E_HIP, "pulse_create: %s", hipGetErrorString(e));
    }
    if (ctx->errors.ensure(sizeof(uint32_t)) != PULSE_OK || hipMemset(ctx->errors.ptr, 0, sizeof(uint32_t)) != hipSuccess) {
        pulse_destroy(ctx);
        PULSE_FAIL(PULSE_E_HIP, "pulse_create: cannot allocate the error word");
    }
    // this device's copy of the allocation-free entry point's scratch
    e = hipGetSymbolAddress(&ctx->noalloc_scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ctx->noalloc_scratch) {
        pulse_destroy(ctx);
        PULSE_FAIL(PULSE_E_HIP, "pulse_create: no address for the scratch array: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return PULSE_OK;
}

int pulse_destroy(pulse_ctx *ctx) {
    if (!ctx) return PULSE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->own.release();
    for (auto &b : ctx->wf) b.release();
    ctx->errors.release();
    ctx->loss_acc.release();
    ctx->spill.release();
    ctx->census.release();
    for (auto &r : ctx->recs) r.release();
    if (ctx->events_created)
        for (auto &e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->events2_created)
        for (auto &e : ctx->ev2) (void)hipEventDestroy(e);
    if (ctx->stream2) {
        (void)hipStreamSynchronize(ctx->stream2);
        for (int i = 0; i < pt::kMaxTraces; ++i) {
            (void)hipEventDestroy(ctx->fork_ev[i]);
            (void)hipEventDestroy(ctx->join_ev[i]);
        }
        (void)hipStreamDestroy(ctx->stream2);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PULSE_OK;
}

void *pulse_stream(pulse_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int pulse_sync(pulse_ctx *ctx) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sync: null ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t err = 0;
    int rc = pulse_host::take_device_errors(ctx, &err);
    if (rc != PULSE_OK) return rc;
    if (err & kErrStackOverflow) PULSE_FAIL(PULSE_E_STACK_OVERFLOW, "traversal stack overflow (capacity %u)", ctx->stack_entries);
    if (err & kErrIterationCap)
        PULSE_FAIL(PULSE_E_ITERATION_CAP, "traversal reached its loop cap (100000 steps per walk of the "
                   "reference's binary BVHs)");
    return PULSE_OK;
}

namespace {
// Every f32 a with biased exponent in [27, 253) (|a| in [2^-100, 2^126)),
// every mantissa, both signs: counts those where rcp_newton(a) and the
// correctly rounded 1.0f / a differ.
__global__ void k_selftest_rcp(unsigned long long *bad) {
    constexpr uint64_t kExps = 253u - 27u;
    constexpr uint64_t kN = 2ull * kExps << 23;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long local = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < kN; i += stride) {
        const uint32_t m = static_cast<uint32_t>(i & ((1u << 23) - 1u));
        const uint32_t e = 27u + static_cast<uint32_t>((i >> 23) % kExps);
        const uint32_t s = static_cast<uint32_t>((i >> 23) / kExps);
        const float a = __uint_as_float((s << 31) | (e << 23) | m);
        if (__float_as_uint(rcp_newton(a)) != __float_as_uint(1.0f / a)) ++local;
    }
    if (local) atomicAdd(bad, local);
}
}  // namespace

int pulse_selftest_reciprocal(pulse_ctx *ctx, uint64_t *mismatches) {
    if (!ctx || !mismatches) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_selftest_reciprocal: null argument");
    int rc = set_device(ctx);
    if (rc) return rc;
    // the counter: the context's 8-byte loss accumulator (cleared before each use)
    if ((rc = ctx->loss_acc.ensure(sizeof(unsigned long long))) != PULSE_OK) return rc;
    HIP_TRY(hipMemsetAsync(ctx->loss_acc.ptr, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL(k_selftest_rcp, dim3(static_cast<uint32_t>(ctx->num_cus) * 16u), dim3(256), 0, ctx->stream,
                       ctx->loss_acc.as<unsigned long long>());
    HIP_TRY(hipGetLastError());
    unsigned long long n = 0;
    HIP_TRY(hipMemcpyAsync(&n, ctx->loss_acc.ptr, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *mismatches = n;
    return PULSE_OK;
}

int pulse_device_alloc(pulse_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out || bytes == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_device_alloc: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(out, bytes));
    return PULSE_OK;
}
int pulse_device_free(pulse_ctx *ctx, void *p) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_device_free: null ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    if (p) HIP_TRY(hipFree(p));
    return PULSE_OK;
}
int pulse_copy_h2d(pulse_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_copy_h2d: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PULSE_OK;
}
int pulse_copy_d2h(pulse_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_copy_d2h: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PULSE_OK;
}
int pulse_memset_d(pulse_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!ctx || (bytes && !dst)) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_memset_d: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return PULSE_OK;
}

int pulse_scene_upload(pulse_ctx *ctx, const pulse_scene_desc *d) {
    if (!ctx || !d) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_upload: null argument");
    int rc = set_device(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = scene_upload(ctx->own, d)) != PULSE_OK) return rc;
    return bind_scene(ctx, &ctx->own);
}

int pulse_scene_generation(pulse_ctx *ctx, uint64_t *out) {
    if (!ctx || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_generation: null argument");
    *out = ctx->generation;
    return PULSE_OK;
}

int pulse_scene_update_materials(pulse_ctx *ctx, const pulse_material *materials, uint32_t count) {
    if (!ctx || (count && !materials)) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_update_materials: bad argument");
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_scene_update_materials: no scene uploaded");
    if (count != ctx->cur->n_materials)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_scene_update_materials: %u materials, scene has %u", count, ctx->cur->n_materials);
    int rc = set_device(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (count) HIP_TRY(hipMemcpy(ctx->cur->buf[B_MATERIALS].ptr, materials, count * sizeof(pulse_material), hipMemcpyHostToDevice));
    for (auto &r : ctx->recs)
        if (r.scene == ctx->cur) r.valid = false;  // materials steer the paths
    return PULSE_OK;
}

int pulse_scene_update_instances(pulse_ctx *ctx, pulse_scene *scene, const pulse_scene_desc *desc) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_update_instances: null ctx");
    if (scene && scene->ctx != ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_update_instances: scene of another context");
    int rc = set_device(ctx);
    if (rc) return rc;
    SceneData *sd = scene ? &scene->data : &ctx->own;
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // launches in flight read the old buffers
    if ((rc = scene_update_instances(*sd, desc)) != PULSE_OK) return rc;
    if (ctx->cur == sd) {
        uint32_t st = stack_for(*sd, ctx->loop_variant);
        if (st == 0) {
            ctx->cur = nullptr;
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "BVH too deep for the current traversal loop");
        }
        ctx->stack_entries = st;
        ctx->generation += 1;
    }
    return PULSE_OK;
}

int pulse_scene_refit_instances(pulse_ctx *ctx, pulse_scene *scene, const float *object_world, uint32_t n) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_refit_instances: null ctx");
    if (scene && scene->ctx != ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_refit_instances: scene of another context");
    int rc = set_device(ctx);
    if (rc) return rc;
    SceneData *sd = scene ? &scene->data : &ctx->own;
    if ((rc = scene_refit_instances(*sd, object_world, n, ctx->stream)) != PULSE_OK) {
        if (ctx->cur == sd && !sd->valid) ctx->cur = nullptr;
        return rc;
    }
    if (ctx->cur == sd) ctx->generation += 1;
    return PULSE_OK;
}

int pulse_scene_rebuild_instances(pulse_ctx *ctx, pulse_scene *scene, const float *object_world, uint32_t n) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_rebuild_instances: null ctx");
    if (scene && scene->ctx != ctx)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_rebuild_instances: scene of another context");
    int rc = set_device(ctx);
    if (rc) return rc;
    SceneData *sd = scene ? &scene->data : &ctx->own;
    if ((rc = scene_rebuild_instances(*sd, object_world, n, ctx->stream)) != PULSE_OK) {
        if (ctx->cur == sd && !sd->valid) ctx->cur = nullptr;
        return rc;
    }
    if (ctx->cur == sd) {
        uint32_t st = stack_for(*sd, ctx->loop_variant);
        if (st == 0) {
            ctx->cur = nullptr;
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_scene_rebuild_instances: TLAS too deep for the traversal stack");
        }
        ctx->stack_entries = st;
        ctx->generation += 1;
    }
    return PULSE_OK;
}

int pulse_scene_create(pulse_ctx *ctx, const pulse_scene_desc *desc, pulse_scene **out) {
    if (!ctx || !desc || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_create: null argument");
    int rc = set_device(ctx);
    if (rc) return rc;
    auto *sc = new pulse_scene();
    sc->ctx = ctx;
    if ((rc = scene_upload(sc->data, desc)) != PULSE_OK) {
        sc->data.release();
        delete sc;
        return rc;
    }
    *out = sc;
    return PULSE_OK;
}

int pulse_scene_destroy(pulse_scene *scene) {
    if (!scene) return PULSE_OK;
    pulse_ctx *ctx = scene->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->cur == &scene->data) ctx->cur = ctx->own.valid ? &ctx->own : nullptr;
    for (auto &r : ctx->recs)
        if (r.scene == &scene->data) r.release();
    scene->data.release();
    delete scene;
    return PULSE_OK;
}

int pulse_scene_bind(pulse_ctx *ctx, pulse_scene *scene) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_bind: null ctx");
    if (scene && scene->ctx != ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_scene_bind: scene of another context");
    SceneData *sd = scene ? &scene->data : &ctx->own;
    if (!sd->valid) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_scene_bind: no scene uploaded");
    if (ctx->cur == sd) return PULSE_OK;
    return bind_scene(ctx, sd);
}

int pulse_trace_rays(pulse_ctx *ctx, const pulse_ray *rays, pulse_hit *hits, uint32_t n) {
    if (!ctx || (n && (!rays || !hits))) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_trace_rays: bad argument");
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_trace_rays: no scene uploaded");
    if (int rc_ = check_traversal(*ctx->cur, ctx->loop_variant)) return rc_;
    int rc = set_device(ctx);
    if (rc) return rc;
    if (n == 0) return PULSE_OK;
    PULSE_STACK_DISPATCH(launch_trace_rays, ctx, rays, hits, n);
    HIP_TRY(hipGetLastError());
    return PULSE_OK;
}

int pulse_trace_shadow_rays(pulse_ctx *ctx, const pulse_ray *rays, uint32_t *occluded, uint32_t n) {
    if (!ctx || (n && (!rays || !occluded))) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_trace_shadow_rays: bad argument");
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_trace_shadow_rays: no scene uploaded");
    if (int rc_ = check_traversal(*ctx->cur, ctx->loop_variant)) return rc_;
    int rc = set_device(ctx);
    if (rc) return rc;
    if (n == 0) return PULSE_OK;
    PULSE_STACK_DISPATCH(launch_shadow_rays, ctx, rays, occluded, n);
    HIP_TRY(hipGetLastError());
    return PULSE_OK;
}

}  // extern "C"

namespace {
// Checks of pulse_trace's arguments and the frame arguments they give; `who`
// names the entry point in messages.
int frame_setup(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, const float *accum_rgba,
                const char *who, FrameArgs &fa, uint32_t &traces) {
    if (!ctx || !cam || !prm || !accum_rgba) PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: null argument", who);
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "%s: no scene uploaded", who);
    if (int rc_ = check_traversal(*ctx->cur, ctx->loop_variant)) return rc_;
    if (cam->width == 0 || cam->height == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: empty image", who);
    if (prm->mode > PULSE_MODE_GGX_NEE) PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: unknown mode %u", who, prm->mode);
    if (prm->nee_eval > PULSE_NEE_GGX) PULSE_FAIL(PULSE_E_INVALID_ARG, "%s: unknown nee_eval %u", who, prm->nee_eval);
    traces = prm->mode == PULSE_MODE_REF_NEE_LAMBERT ? prm->max_depth + 1u : prm->max_depth;
    if (traces > static_cast<uint32_t>(kMaxTraces))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: max_depth %u above %d", who, prm->max_depth, kMaxTraces - 1);
    const size_t npix = static_cast<size_t>(cam->width) * cam->height;
    if (npix > 0xffffffffull / 4) PULSE_FAIL(PULSE_E_UNSUPPORTED, "%s: image too large", who);
    fa = FrameArgs{};
    fa.cam = *cam;
    fa.mode = prm->mode;
    fa.nee_ggx = prm->nee_eval == PULSE_NEE_GGX ? 1u : 0u;
    fa.max_traces = traces;
    fa.rr_start = prm->rr_start_depth;
    fa.horizon_clamp = (prm->flags & PULSE_FLAG_HORIZON_CLAMP) ? 1u : 0u;
    fa.light_power = (prm->flags & PULSE_FLAG_LIGHT_POWER) ? 1u : 0u;
    fa.gi = 0;
    if (fa.light_power && ctx->cur->n_strength_cdf < ctx->cur->uniform.light_count)
        PULSE_FAIL(PULSE_E_LAYOUT, "%s: PULSE_FLAG_LIGHT_POWER needs one light_emission_strength_cdf entry "
                   "per light (%u for %u lights)", who, ctx->cur->n_strength_cdf, ctx->cur->uniform.light_count);
    return PULSE_OK;
}
}  // namespace

// Path record of a pulse_trace call (PULSE_FLAG_RECORD_PATHS, or the forward
// pass of pulse_forward_adjoint), one per camera: accumulation_count 0
// starts one (in the slot of the same camera, else a free slot, else the
// least recently used), a call that continues a camera's recorded samples
// with the same parameters and scene state extends it (the older samples
// are kept when the buffers grow); anything else leaves that camera without
// a valid record.  Sets wb's record pointers (its rec_base is set per
// sample) or leaves them null.
int record_setup(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, uint32_t traces, size_t npix,
                 WfBuffers &wb) {
    ctx->rec_pending = -1;
    const bool want = (prm->flags & PULSE_FLAG_RECORD_PATHS) || ctx->rec_force;
    if (!want) return PULSE_OK;
    const uint64_t first = prm->accumulation_count, total = first + std::max(prm->spp, 1u);
    int slot = -1;
    for (int i = 0; i < pulse_ctx::kMaxRecords; ++i)
        if (ctx->recs[i].valid && same_camera(ctx->recs[i].cam, *cam)) slot = i;
    const uint64_t entries = total * npix * traces;
    bool extend = false;
    if (slot >= 0) {
        pulse_ctx::PathRecord &r = ctx->recs[slot];
        extend = first > 0 && record_matches(ctx, r, *cam, *prm) && r.samples == first && r.npix == npix &&
                 r.traces == traces;
        r.valid = false;
    }
    if ((first > 0 && !extend) || entries >= (1ull << 32) || entries * 21 > (16ull << 30)) return PULSE_OK;
    if (slot < 0) {
        for (int i = 0; i < pulse_ctx::kMaxRecords && slot < 0; ++i)
            if (!ctx->recs[i].valid) slot = i;
        if (slot < 0) {
            slot = 0;
            for (int i = 1; i < pulse_ctx::kMaxRecords; ++i)
                if (ctx->recs[i].used < ctx->recs[slot].used) slot = i;
        }
    }
    pulse_ctx::PathRecord &r = ctx->recs[slot];
    r.valid = false;
    // all records together stay within kRecordBudget: the least recently
    // used other records give their memory back first
    const size_t need = static_cast<size_t>(entries) * 21;
    for (;;) {
        size_t held = 0;
        int lru = -1;
        for (int i = 0; i < pulse_ctx::kMaxRecords; ++i) {
            if (i == slot || !ctx->recs[i].hit.ptr) continue;
            held += ctx->recs[i].hit.cap + ctx->recs[i].v.cap + ctx->recs[i].vis.cap;
            if (lru < 0 || ctx->recs[i].used < ctx->recs[lru].used) lru = i;
        }
        if (held + need <= kRecordBudget || lru < 0) break;
        ctx->recs[lru].release();
    }
    const size_t keep = extend ? static_cast<size_t>(first) * npix * traces : 0;
    auto grow = [&](pulse_host::DevBuf &b, size_t elem) -> int {
        const size_t bytes = static_cast<size_t>(entries) * elem;
        if (b.ptr && b.cap >= bytes) return PULSE_OK;
        void *np = nullptr;
        HIP_TRY(hipMalloc(&np, bytes));
        if (keep && b.ptr) HIP_TRY(hipMemcpyAsync(np, b.ptr, keep * elem, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        b.release();
        b.ptr = np;
        b.cap = bytes;
        return PULSE_OK;
    };
    int rc;
    if ((rc = grow(r.hit, sizeof(float4))) || (rc = grow(r.v, sizeof(float))) || (rc = grow(r.vis, sizeof(uint8_t))))
        return rc;
    wb.rec_hit = r.hit.as<float4>();
    wb.rec_v = r.v.as<float>();
    wb.rec_vis = r.vis.as<uint8_t>();
    wb.rec_traces = traces;
    r.samples = static_cast<uint32_t>(total);
    r.npix = static_cast<uint32_t>(npix);
    r.traces = traces;
    r.scene = ctx->cur;
    r.generation = ctx->cur->generation;
    r.loop = ctx->loop_variant;
    r.cam = *cam;
    r.prm = *prm;
    r.used = ++ctx->rec_clock;
    ctx->rec_pending = slot;  // valid once the call has queued every sample
    return PULSE_OK;
}

extern "C" {

int pulse_trace(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, float *accum_rgba) {
    FrameArgs fa;
    uint32_t traces = 0;
    int rc = frame_setup(ctx, cam, prm, accum_rgba, "pulse_trace", fa, traces);
    if (rc != PULSE_OK) return rc;
    if ((rc = set_device(ctx)) != PULSE_OK) return rc;
    const size_t npix = static_cast<size_t>(cam->width) * cam->height;
    if ((rc = ensure_wavefront(ctx, npix)) != PULSE_OK) return rc;
    WfBuffers wb = wf_buffers(ctx);
    if ((rc = record_setup(ctx, cam, prm, traces, npix, wb)) != PULSE_OK) return rc;
    const bool timing = (prm->flags & PULSE_FLAG_STAGE_TIMING) != 0;
    if (timing && !ctx->events_created) {
        for (auto &e : ctx->ev) HIP_TRY(hipEventCreate(&e));
        ctx->events_created = true;
    }
    const uint32_t cap = static_cast<uint32_t>(npix);
    const uint32_t tiles = ((cam->width + 7) / 8) * ((cam->height + 7) / 8);
    const uint32_t gen_grid = (tiles * 64 + kShadeBlock - 1) / kShadeBlock;
    const uint32_t shade_grid = (cap + kShadeBlock - 1) / kShadeBlock;

    const bool persistent = ctx->loop_variant != 0 && ctx->persistent > 0;
    if (persistent && (rc = ensure_spill(ctx, cap)) != PULSE_OK) return rc;
    const bool census = (prm->flags & PULSE_FLAG_COUNT_WORK) && ctx->loop_variant == 2;
    if (census) {
        if ((rc = ctx->census.ensure(sizeof(uint64_t) * 12)) != PULSE_OK) return rc;
        HIP_TRY(hipMemsetAsync(ctx->census.ptr, 0, sizeof(uint64_t) * 12, ctx->stream));
    }
    // shadow rays of bounce k on the second stream beside bounce k+1's
    // extension rays: both queues come from shade(k) and neither launch
    // touches what the other writes (hits vs path colours); shade(k+1) waits
    // for both, since it rewrites the shadow queue and the path colours
    const bool concurrent = persistent && ctx->concurrent_shadow && !census;
    if (concurrent && (rc = ensure_stream2(ctx, timing)) != PULSE_OK) return rc;
    ctx->stats = pulse_stats{};
    double ms_primary = 0, ms_ext_bounce = 0, ms_shadow = 0, ms_total = 0, ms_incoherent = 0;
    uint64_t ext_total = 0, ext_primary = 0, sh_total = 0;
    std::vector<uint32_t> counts(2 * (kMaxTraces + 1));
    for (uint32_t s = 0; s < std::max(prm->spp, 1u); ++s) {
        fa.accumulation_count = prm->accumulation_count + s;
        wb.rec_base = fa.accumulation_count * static_cast<uint32_t>(npix);
        HIP_TRY(hipMemsetAsync(wb.ext_count, 0, sizeof(uint32_t) * kCounterWords, ctx->stream));
        if (concurrent) {
            // timing events (ctx->ev2): [0] start; extend k: [1 + 2k], [2 + 2k];
            // after shade k: [1 + 2T + k]; both of bounce k's launches done:
            // [1 + 3T + k]; shadow k (second stream): [1 + 4T + k], [1 + 5T + k];
            // end [1 + 6T]
            const int T = static_cast<int>(traces);
            hipEvent_t *E = ctx->ev2;
            hipStream_t A = ctx->stream, B = ctx->stream2;
            if (timing) HIP_TRY(hipEventRecord(E[0], A));
            hipLaunchKernelGGL(k_generate, dim3(gen_grid), dim3(kShadeBlock), 0, A, wb, fa);
            int q = 0;
            if (timing) HIP_TRY(hipEventRecord(E[1], A));
            if ((rc = launch_queue(ctx, wb, cap, false, q, 0, A, 0)) != PULSE_OK) return rc;
            if (timing) HIP_TRY(hipEventRecord(E[2], A));
            for (int k = 0; k < T; ++k) {
                hipLaunchKernelGGL(k_shade, dim3(shade_grid), dim3(kShadeBlock), 0, A, dev_scene(ctx), wb, fa, q, k);
                if (timing) HIP_TRY(hipEventRecord(E[1 + 2 * T + k], A));
                q ^= 1;
                HIP_TRY(hipEventRecord(ctx->fork_ev[k], A));
                HIP_TRY(hipStreamWaitEvent(B, ctx->fork_ev[k], 0));
                // a pair of launches shares the CUs from the start: each
                // takes the paired share of waves; the last shadow launch,
                // alone, all of them
                const int pair_waves = k + 1 < T ? ctx->eff_pair_waves() : 0;
                const int pair_waves_sh = k + 1 < T ? ctx->eff_pair_waves_shadow() : 0;
                if (timing) HIP_TRY(hipEventRecord(E[1 + 4 * T + k], B));
                if ((rc = launch_queue(ctx, wb, cap, true, 0, k, B, 1, pair_waves_sh)) != PULSE_OK) return rc;
                if (timing) HIP_TRY(hipEventRecord(E[1 + 5 * T + k], B));
                HIP_TRY(hipEventRecord(ctx->join_ev[k], B));
                if (k + 1 < T) {
                    if (timing) HIP_TRY(hipEventRecord(E[1 + 2 * (k + 1)], A));
                    if ((rc = launch_queue(ctx, wb, cap, false, q, k + 1, A, 0, pair_waves)) != PULSE_OK) return rc;
                    if (timing) HIP_TRY(hipEventRecord(E[2 + 2 * (k + 1)], A));
                }
                HIP_TRY(hipStreamWaitEvent(A, ctx->join_ev[k], 0));
                if (timing) HIP_TRY(hipEventRecord(E[1 + 3 * T + k], A));
            }
            hipLaunchKernelGGL(k_accumulate, dim3(shade_grid), dim3(kShadeBlock), 0, A, wb,
                               reinterpret_cast<float4 *>(accum_rgba), cap, fa.accumulation_count);
            if (timing) HIP_TRY(hipEventRecord(E[1 + 6 * T], A));
            HIP_TRY(hipGetLastError());
            if (timing || (prm->flags & PULSE_FLAG_COLLECT_STATS)) {
                HIP_TRY(hipMemcpyAsync(counts.data(), wb.ext_count, sizeof(uint32_t) * counts.size(),
                                       hipMemcpyDeviceToHost, A));
                HIP_TRY(hipStreamSynchronize(A));
                for (uint32_t k = 0; k < traces; ++k) {
                    ext_total += counts[k];
                    sh_total += counts[kMaxTraces + 1 + k];
                }
                ext_primary += counts[0];
            }
            if (timing) {
                auto el = [&](int a, int b) {
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, E[a], E[b]);
                    return static_cast<double>(ms);
                };
                ms_total += el(0, 1 + 6 * T);
                ms_primary += el(1, 2);
                for (int k = 0; k < T; ++k) {
                    ms_incoherent += el(1 + 2 * T + k, 1 + 3 * T + k);
                    ms_shadow += el(1 + 4 * T + k, 1 + 5 * T + k);
                    if (k > 0) ms_ext_bounce += el(1 + 2 * k, 2 + 2 * k);
                }
            }
            continue;
        }
        int e = 0;
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
        hipLaunchKernelGGL(k_generate, dim3(gen_grid), dim3(kShadeBlock), 0, ctx->stream, wb, fa);
        int q = 0;
        for (uint32_t k = 0; k < traces; ++k) {
            if (census) PULSE_STACK_DISPATCH(launch_census, ctx, wb, cap, false, q, static_cast<int>(k), k == 0 ? 0 : 1);
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
            if (persistent) {
                if ((rc = launch_queue(ctx, wb, cap, false, q, static_cast<int>(k))) != PULSE_OK) return rc;
            } else {
                PULSE_STACK_DISPATCH(launch_extend, ctx, wb, cap, q, static_cast<int>(k));
            }
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
            hipLaunchKernelGGL(k_shade, dim3(shade_grid), dim3(kShadeBlock), 0, ctx->stream, dev_scene(ctx), wb, fa, q,
                               static_cast<int>(k));
            if (census) PULSE_STACK_DISPATCH(launch_census, ctx, wb, cap, true, 0, static_cast<int>(k), 2);
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
            if (persistent) {
                if ((rc = launch_queue(ctx, wb, cap, true, 0, static_cast<int>(k))) != PULSE_OK) return rc;
            } else {
                PULSE_STACK_DISPATCH(launch_shadow, ctx, wb, cap, static_cast<int>(k));
            }
            q ^= 1;
        }
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
        hipLaunchKernelGGL(k_accumulate, dim3(shade_grid), dim3(kShadeBlock), 0, ctx->stream, wb,
                           reinterpret_cast<float4 *>(accum_rgba), cap, fa.accumulation_count);
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[e++], ctx->stream));
        HIP_TRY(hipGetLastError());
        if (timing || (prm->flags & PULSE_FLAG_COLLECT_STATS)) {
            HIP_TRY(hipMemcpyAsync(counts.data(), wb.ext_count, sizeof(uint32_t) * counts.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            for (uint32_t k = 0; k < traces; ++k) {
                ext_total += counts[k];
                sh_total += counts[kMaxTraces + 1 + k];
            }
            ext_primary += counts[0];
        }
        if (timing) {
            auto el = [&](int a, int b) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]);
                return static_cast<double>(ms);
            };
            ms_total += el(0, e - 1);
            // events: 0 gen, then per k: [ext start, shade start, shadow start], end-of-shadow = next ext start
            for (uint32_t k = 0; k < traces; ++k) {
                int ext0 = 1 + 3 * static_cast<int>(k);
                int shade0 = ext0 + 1, shadow0 = ext0 + 2, next0 = ext0 + 3;
                double t_ext = el(ext0, shade0), t_sh = el(shadow0, next0);
                if (k == 0) ms_primary += t_ext;
                else ms_ext_bounce += t_ext;
                ms_shadow += t_sh;
                ms_incoherent += (k == 0 ? 0.0 : t_ext) + t_sh;
            }
        }
    }
    if (census) {
        HIP_TRY(hipMemcpyAsync(ctx->stats.work, ctx->census.ptr, sizeof(uint64_t) * 12, hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    ctx->stats.primary_rays = ext_primary;
    ctx->stats.extension_rays = ext_total;
    ctx->stats.shadow_rays = sh_total;
    ctx->stats.incoherent_rays = (ext_total - ext_primary) + sh_total;
    ctx->stats.stack_entries = ctx->stack_entries;
    ctx->stats.ms_total = ms_total;
    ctx->stats.ms_primary_extend = ms_primary;
    ctx->stats.ms_incoherent = ms_incoherent;
    ctx->stats.ms_extend_bounce = ms_ext_bounce;
    ctx->stats.ms_shadow = ms_shadow;
    ctx->stats.ms_shade = ms_total - ms_primary - ms_incoherent;
    ctx->stats.extend_launches = traces * std::max(prm->spp, 1u);
    ctx->stats.shadow_launches = traces * std::max(prm->spp, 1u);
    if (ctx->rec_pending >= 0) ctx->recs[ctx->rec_pending].valid = true;
    ctx->rec_pending = -1;
    return PULSE_OK;
}

int pulse_trace_noalloc(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, float *accum_rgba) {
    FrameArgs fa;
    uint32_t traces = 0;
    int rc = frame_setup(ctx, cam, prm, accum_rgba, "pulse_trace_noalloc", fa, traces);
    if (rc != PULSE_OK) return rc;
    const size_t npix = static_cast<size_t>(cam->width) * cam->height;
    if (npix > kNoAllocMaxPixels)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_trace_noalloc: %ux%u = %zu pixels, above the scratch's %zu", cam->width,
                   cam->height, npix, kNoAllocMaxPixels);
    if (prm->flags & (PULSE_FLAG_STAGE_TIMING | PULSE_FLAG_COUNT_WORK))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_trace_noalloc: stage timing and the work census need buffers of "
                   "their own; use pulse_trace");
    if (!ctx->noalloc_scratch) PULSE_FAIL(PULSE_E_HIP, "pulse_trace_noalloc: no scratch address for this device");
    if ((rc = set_device(ctx)) != PULSE_OK) return rc;
    WfBuffers wb = wf_buffers_in(static_cast<unsigned char *>(ctx->noalloc_scratch), kNoAllocMaxPixels);
    const uint32_t cap = static_cast<uint32_t>(npix);
    const uint32_t tiles = ((cam->width + 7) / 8) * ((cam->height + 7) / 8);
    const uint32_t gen_grid = (tiles * 64 + kShadeBlock - 1) / kShadeBlock;
    const uint32_t shade_grid = (cap + kShadeBlock - 1) / kShadeBlock;
    const uint32_t clear_grid = (static_cast<uint32_t>(kCounterWords) + 255u) / 256u;
    std::array<uint32_t, 2 * (kMaxTraces + 1)> counts{};
    uint64_t ext_total = 0, ext_primary = 0, sh_total = 0;
    // one ray per lane, the whole traversal stack in LDS (no spill area);
    // every launch on the context's stream
    for (uint32_t s = 0; s < std::max(prm->spp, 1u); ++s) {
        fa.accumulation_count = prm->accumulation_count + s;
        hipLaunchKernelGGL(k_fill_zero, dim3(clear_grid), dim3(256), 0, ctx->stream,
                           reinterpret_cast<float *>(wb.ext_count), static_cast<uint32_t>(kCounterWords));
        hipLaunchKernelGGL(k_generate, dim3(gen_grid), dim3(kShadeBlock), 0, ctx->stream, wb, fa);
        int q = 0;
        for (uint32_t k = 0; k < traces; ++k) {
            PULSE_STACK_DISPATCH(launch_extend, ctx, wb, cap, q, static_cast<int>(k));
            hipLaunchKernelGGL(k_shade, dim3(shade_grid), dim3(kShadeBlock), 0, ctx->stream, dev_scene(ctx), wb, fa, q,
                               static_cast<int>(k));
            PULSE_STACK_DISPATCH(launch_shadow, ctx, wb, cap, static_cast<int>(k));
            q ^= 1;
        }
        hipLaunchKernelGGL(k_accumulate, dim3(shade_grid), dim3(kShadeBlock), 0, ctx->stream, wb,
                           reinterpret_cast<float4 *>(accum_rgba), cap, fa.accumulation_count);
        HIP_TRY(hipGetLastError());
        if (prm->flags & PULSE_FLAG_COLLECT_STATS) {
            HIP_TRY(hipMemcpyAsync(counts.data(), wb.ext_count, sizeof(uint32_t) * counts.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            for (uint32_t k = 0; k < traces; ++k) {
                ext_total += counts[k];
                sh_total += counts[kMaxTraces + 1 + k];
            }
            ext_primary += counts[0];
        }
    }
    ctx->stats = pulse_stats{};
    ctx->stats.primary_rays = ext_primary;
    ctx->stats.extension_rays = ext_total;
    ctx->stats.shadow_rays = sh_total;
    ctx->stats.incoherent_rays = (ext_total - ext_primary) + sh_total;
    ctx->stats.stack_entries = ctx->stack_entries;
    ctx->stats.extend_launches = traces * std::max(prm->spp, 1u);
    ctx->stats.shadow_launches = traces * std::max(prm->spp, 1u);
    return PULSE_OK;
}

int pulse_trace_noalloc_max_pixels(uint64_t *out) {
    if (!out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_trace_noalloc_max_pixels: null out");
    *out = kNoAllocMaxPixels;
    return PULSE_OK;
}

int pulse_trace_gbuffer(pulse_ctx *ctx, const pulse_gbuffer *gb, uint32_t width, uint32_t height,
                        const pulse_params *prm, float *gi_rgba) {
    if (!ctx || !gb || !prm || !gi_rgba || !gb->position || !gb->normal || !gb->base_color || !gb->emissive)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_trace_gbuffer: null argument");
    if (!ctx->has_scene()) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_trace_gbuffer: no scene uploaded");
    if (int rc_ = check_traversal(*ctx->cur, ctx->loop_variant)) return rc_;
    if (width == 0 || height == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_trace_gbuffer: empty image");
    if (prm->max_depth > static_cast<uint32_t>(kMaxTraces - 1))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_trace_gbuffer: max_depth %u above %d", prm->max_depth, kMaxTraces - 1);
    if (prm->nee_eval != PULSE_NEE_LAMBERT)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_trace_gbuffer: the GI pass evaluates NEE with the Lambert BRDF only");
    const uint64_t npix64 = static_cast<uint64_t>(width) * height;
    if (npix64 > 0xffffffffull / 4) PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_trace_gbuffer: image too large");
    if ((prm->flags & PULSE_FLAG_LIGHT_POWER) && ctx->cur->n_strength_cdf < ctx->cur->uniform.light_count)
        PULSE_FAIL(PULSE_E_LAYOUT, "pulse_trace_gbuffer: PULSE_FLAG_LIGHT_POWER needs one light_emission_strength_cdf "
                   "entry per light");
    int rc = set_device(ctx);
    if (rc) return rc;
    const uint32_t npix = static_cast<uint32_t>(npix64);
    if ((rc = ensure_wavefront(ctx, npix)) != PULSE_OK) return rc;
    WfBuffers wb = wf_buffers(ctx);
    FrameArgs fa{};
    fa.mode = PULSE_MODE_REF_NEE_LAMBERT;
    fa.nee_ggx = 0;
    fa.max_traces = prm->max_depth;
    fa.rr_start = 0;
    fa.horizon_clamp = 0;
    fa.light_power = (prm->flags & PULSE_FLAG_LIGHT_POWER) ? 1u : 0u;
    fa.gi = 1;
    GBufferArgs g;
    g.position = reinterpret_cast<const float4 *>(gb->position);
    g.normal = reinterpret_cast<const float4 *>(gb->normal);
    g.base_color = reinterpret_cast<const float4 *>(gb->base_color);
    g.emissive = reinterpret_cast<const float4 *>(gb->emissive);
    g.npix = npix;
    const bool persistent = ctx->loop_variant != 0 && ctx->persistent > 0;
    if (persistent && (rc = ensure_spill(ctx, npix)) != PULSE_OK) return rc;
    const uint32_t grid = (npix + kShadeBlock - 1) / kShadeBlock;
    const uint32_t spp = std::max(prm->spp, 1u);
    const float spp_inv = 1.0f / static_cast<float>(spp);
    auto shadow = [&](int k) -> int {
        if (persistent) return launch_queue(ctx, wb, npix, true, 0, k);
        PULSE_STACK_DISPATCH(launch_shadow, ctx, wb, npix, k);
        return PULSE_OK;
    };
    // as in pulse_trace: each shadow launch on the second stream beside the
    // next extension launch (generate feeds the first pair like a shade)
    const bool concurrent = persistent && ctx->concurrent_shadow;
    if (concurrent && (rc = ensure_stream2(ctx, false)) != PULSE_OK) return rc;
    for (uint32_t s = 0; s < spp; ++s) {
        HIP_TRY(hipMemsetAsync(wb.ext_count, 0, sizeof(uint32_t) * kCounterWords, ctx->stream));
        hipLaunchKernelGGL(k_gi_generate, dim3(grid), dim3(kShadeBlock), 0, ctx->stream, dev_scene(ctx), wb, fa, g, s);
        if (concurrent) {
            hipStream_t A = ctx->stream, B = ctx->stream2;
            const int T = static_cast<int>(fa.max_traces);
            int q = 0;
            for (int k = -1; k < T; ++k) {
                if (k >= 0) {
                    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kShadeBlock), 0, A, dev_scene(ctx), wb, fa, q, k);
                    q ^= 1;
                }
                HIP_TRY(hipEventRecord(ctx->fork_ev[k + 1], A));
                HIP_TRY(hipStreamWaitEvent(B, ctx->fork_ev[k + 1], 0));
                // k = -1: NEE at the G-buffer points
                const int pair_waves = k + 1 < T ? ctx->eff_pair_waves() : 0;
                const int pair_waves_sh = k + 1 < T ? ctx->eff_pair_waves_shadow() : 0;
                if ((rc = launch_queue(ctx, wb, npix, true, 0, k < 0 ? kMaxTraces : k, B, 1, pair_waves_sh)) != PULSE_OK)
                    return rc;
                HIP_TRY(hipEventRecord(ctx->join_ev[k + 1], B));
                if (k + 1 < T && (rc = launch_queue(ctx, wb, npix, false, q, k + 1, A, 0, pair_waves)) != PULSE_OK)
                    return rc;
                HIP_TRY(hipStreamWaitEvent(A, ctx->join_ev[k + 1], 0));
            }
            hipLaunchKernelGGL(k_gi_accumulate, dim3(grid), dim3(kShadeBlock), 0, A, wb,
                               reinterpret_cast<float4 *>(gi_rgba), npix, s, spp_inv);
            continue;
        }
        if ((rc = shadow(kMaxTraces)) != PULSE_OK) return rc;  // NEE at the G-buffer points
        int q = 0;
        for (uint32_t k = 0; k < fa.max_traces; ++k) {
            if (persistent) {
                if ((rc = launch_queue(ctx, wb, npix, false, q, static_cast<int>(k))) != PULSE_OK) return rc;
            } else {
                PULSE_STACK_DISPATCH(launch_extend, ctx, wb, npix, q, static_cast<int>(k));
            }
            hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kShadeBlock), 0, ctx->stream, dev_scene(ctx), wb, fa, q,
                               static_cast<int>(k));
            if ((rc = shadow(static_cast<int>(k))) != PULSE_OK) return rc;
            q ^= 1;
        }
        hipLaunchKernelGGL(k_gi_accumulate, dim3(grid), dim3(kShadeBlock), 0, ctx->stream, wb,
                           reinterpret_cast<float4 *>(gi_rgba), npix, s, spp_inv);
    }
    HIP_TRY(hipGetLastError());
    return PULSE_OK;
}

int pulse_present(pulse_ctx *ctx, const float *accum_rgba, uint32_t src_w, uint32_t src_h, void *dst, uint32_t dst_w,
                  uint32_t dst_h, int format) {
    if (!ctx || !accum_rgba || !dst) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_present: null argument");
    if (src_w == 0 || src_h == 0 || dst_w == 0 || dst_h == 0)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_present: empty image");
    if (format < PULSE_PRESENT_RGBA32F || format > PULSE_PRESENT_RGBA8_UNORM)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_present: unknown format %d", format);
    const uint64_t n = static_cast<uint64_t>(dst_w) * dst_h;
    if (n > 0xffffffffull || static_cast<uint64_t>(src_w) * src_h > 0xffffffffull)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_present: image too large");
    int rc = set_device(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_present, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const float4 *>(accum_rgba), src_w, src_h, dst, dst_w, dst_h, format);
    HIP_TRY(hipGetLastError());
    return PULSE_OK;
}

int pulse_image_loss(pulse_ctx *ctx, const float *image, const float *target, uint32_t npix, double *loss) {
    if (!ctx || !image || !target || !loss || npix == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_image_loss: bad argument");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ctx->loss_acc.ensure(sizeof(double))) != PULSE_OK) return rc;
    HIP_TRY(hipMemsetAsync(ctx->loss_acc.ptr, 0, sizeof(double), ctx->stream));
    uint32_t grid = std::min<uint32_t>((npix + 255) / 256, 1024);
    hipLaunchKernelGGL(k_image_loss, dim3(grid), dim3(256), 0, ctx->stream, reinterpret_cast<const float4 *>(image),
                       reinterpret_cast<const float4 *>(target), npix, ctx->loss_acc.as<double>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(loss, ctx->loss_acc.ptr, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PULSE_OK;
}

int pulse_adjoint(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, const float *image,
                  const float *target, float *grad, float *replay_rgba) {
    if (!ctx || !cam || !prm || !image || !target || !grad) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint: null argument");
    if (int rc = adjoint_check(ctx, prm, "pulse_adjoint")) return rc;
    if (cam->width == 0 || cam->height == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint: empty image");
    pulse_adjoint_view v;
    v.camera = *cam;
    v.image = image;
    v.target = target;
    v.replay_rgba = replay_rgba;
    return adjoint_run(ctx, prm, &v, 1, grad);
}

int pulse_adjoint_views(pulse_ctx *ctx, const pulse_adjoint_view *views, uint32_t n_views, const pulse_params *prm,
                        float *grad) {
    if (!ctx || !views || !prm || !grad) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint_views: null argument");
    if (n_views == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint_views: no views");
    if (int rc = adjoint_check(ctx, prm, "pulse_adjoint_views")) return rc;
    for (uint32_t i = 0; i < n_views; ++i) {
        const pulse_adjoint_view &v = views[i];
        if (!v.image || !v.target) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint_views: view %u: null image or target", i);
        if (v.camera.width == 0 || v.camera.height == 0)
            PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_adjoint_views: view %u: empty image", i);
    }
    return adjoint_run(ctx, prm, views, n_views, grad);
}

int pulse_forward_adjoint(pulse_ctx *ctx, const pulse_camera *cam, const pulse_params *prm, const float *target,
                          float *image_out, float *grad, double *loss) {
    // the forward pass keeps its paths: the adjoint reads them back
    if (ctx) ctx->rec_force = true;
    int rc = pulse_trace(ctx, cam, prm, image_out);
    if (ctx) ctx->rec_force = false;
    if (rc) return rc;
    if (loss) {
        if ((rc = pulse_image_loss(ctx, image_out, target, cam->width * cam->height, loss)) != PULSE_OK) return rc;
    }
    return pulse_adjoint(ctx, cam, prm, image_out, target, grad, nullptr);
}

int pulse_set_option(pulse_ctx *ctx, int option, int64_t value) {
    if (!ctx) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_set_option: null ctx");
    switch (option) {
        case PULSE_OPT_TRAVERSAL_LOOP: {
            if (value < 0 || value > 2) PULSE_FAIL(PULSE_E_INVALID_ARG, "traversal loop must be 0, 1 or 2");
            uint32_t st = ctx->has_scene() ? stack_for(*ctx->cur, static_cast<int>(value)) : 1;
            if (st == 0) PULSE_FAIL(PULSE_E_UNSUPPORTED, "no stack size for this variant and scene");
            ctx->loop_variant = static_cast<int>(value);
            ctx->loop_explicit = true;
            if (ctx->has_scene()) ctx->stack_entries = st;
            return PULSE_OK;
        }
        case PULSE_OPT_PERSISTENT_WAVES:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "persistent waves per CU must be 0..64");
            ctx->persistent = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_REFILL_LANES:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "refill lanes must be 0..64");
            ctx->refill = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_TRIS_PER_STEP:
            if (value != 1 && value != 2) PULSE_FAIL(PULSE_E_INVALID_ARG, "triangles per step must be 1 or 2");
            ctx->tris_per_step = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_LDS_STACK_ENTRIES:
            if (value < 0 || value > 128) PULSE_FAIL(PULSE_E_INVALID_ARG, "LDS stack entries must be 0..128");
            ctx->lds_entries = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_GATHER_LANES:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "gather lanes must be 0..64");
            ctx->gather = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_ADVANCE_PERIOD:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "advance period must be 0..64");
            ctx->advance_period = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_CONCURRENT_SHADOW:
            if (value != 0 && value != 1) PULSE_FAIL(PULSE_E_INVALID_ARG, "concurrent shadow must be 0 or 1");
            ctx->concurrent_shadow = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_PAIR_WAVES_SHADOW:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "pair waves must be 0..64");
            ctx->pair_waves_shadow = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_ADJOINT_REDUCE_ROUNDS:
            if (value < -1 || value > 16) PULSE_FAIL(PULSE_E_INVALID_ARG, "adjoint reduce rounds must be -1..16");
            ctx->adjoint_reduce_rounds = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_PAIR_WAVES:
            if (value < 0 || value > 64) PULSE_FAIL(PULSE_E_INVALID_ARG, "pair waves must be 0..64");
            ctx->pair_waves = static_cast<int>(value);
            return PULSE_OK;
        case PULSE_OPT_INJECT_ERROR: {
            if (value < 0 || value > 0xffffffffll) PULSE_FAIL(PULSE_E_INVALID_ARG, "error bits must fit 32 bits");
            HIP_TRY(hipSetDevice(ctx->device));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            uint32_t err = 0;
            HIP_TRY(hipMemcpy(&err, ctx->errors.ptr, sizeof(err), hipMemcpyDeviceToHost));
            err |= static_cast<uint32_t>(value);
            HIP_TRY(hipMemcpy(ctx->errors.ptr, &err, sizeof(err), hipMemcpyHostToDevice));
            return PULSE_OK;
        }
        default:
            PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_set_option: unknown or read-only option %d", option);
    }
}

int pulse_get_option(pulse_ctx *ctx, int option, int64_t *value) {
    if (!ctx || !value) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_get_option: null argument");
    switch (option) {
        case PULSE_OPT_TRAVERSAL_LOOP: *value = ctx->loop_variant; return PULSE_OK;
        case PULSE_OPT_PERSISTENT_WAVES: *value = ctx->persistent; return PULSE_OK;
        case PULSE_OPT_REFILL_LANES: *value = ctx->eff_refill(); return PULSE_OK;
        case PULSE_OPT_GATHER_LANES: *value = ctx->eff_gather(); return PULSE_OK;
        case PULSE_OPT_ADVANCE_PERIOD: *value = ctx->eff_period(); return PULSE_OK;
        case PULSE_OPT_CONCURRENT_SHADOW: *value = ctx->concurrent_shadow; return PULSE_OK;
        case PULSE_OPT_PAIR_WAVES: *value = ctx->eff_pair_waves(); return PULSE_OK;
        case PULSE_OPT_PAIR_WAVES_SHADOW: *value = ctx->eff_pair_waves_shadow(); return PULSE_OK;
        case PULSE_OPT_ADJOINT_REDUCE_ROUNDS: *value = ctx->adjoint_reduce_rounds; return PULSE_OK;
        case PULSE_OPT_RECORDED_SAMPLES: {
            // of the most recently recorded camera
            const pulse_ctx::PathRecord *last = nullptr;
            for (const auto &r : ctx->recs)
                if (r.used && (!last || r.used > last->used)) last = &r;
            *value = last && last->valid ? last->samples : 0;
            return PULSE_OK;
        }
        case PULSE_OPT_ADJOINT_READ_RECORD: *value = ctx->rec_read ? 1 : 0; return PULSE_OK;
        case PULSE_OPT_LDS_STACK_ENTRIES: *value = ctx->lds_entries; return PULSE_OK;
        case PULSE_OPT_TRIS_PER_STEP: *value = ctx->tris_per_step; return PULSE_OK;
        case PULSE_OPT_STACK_ENTRIES: *value = ctx->stack_entries; return PULSE_OK;
        case PULSE_OPT_FAST_DIVISION: *value = ctx->cur->fast_div ? 1 : 0; return PULSE_OK;
        default: PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_get_option: unknown option %d", option);
    }
}

int pulse_get_stats(pulse_ctx *ctx, pulse_stats *out) {
    if (!ctx || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_get_stats: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t err = 0;
    int rc = pulse_host::take_device_errors(ctx, &err);
    if (rc != PULSE_OK) return rc;
    *out = ctx->stats;
    out->error_flags = err;
    return PULSE_OK;
}

}  // extern "C"
