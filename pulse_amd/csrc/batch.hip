// batch.hip -- the batched form: many views (scenes, cameras, resolutions)
// rendered by one sequence of stage launches per sample, captured once into
// a hipGraph and replayed per frame, with an autotuner over the launch
// shape.  C ABI declared in include/pulse_hip.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "device/batch.h"
#include "host/common.h"
#include "host/context.h"

using namespace pt;
using namespace pulse_host;

struct pulse_batch {
    pulse_ctx *ctx = nullptr;
    pulse_params prm{};
    uint32_t n_views = 0;
    uint32_t traces = 0;        // extension traces per path
    uint32_t cstride = 0;
    uint32_t stack_need = 0;    // deepest TLAS + BLAS over the views' scenes
    uint32_t wide_need = 0;     // the same for the 4-wide BVHs
    bool has_wide = true;       // every view's scene has its 4-wide BVHs
    uint32_t total_slots = 0;
    uint32_t total_tiles = 0;
    uint32_t max_chunks = 0;    // sum over views of ceil(npix / chunk)
    std::vector<BatchView> views;
    std::vector<pulse_scene *> scenes;
    std::vector<uint64_t> scene_gen;  // scene generations the view table was written for
    DevBuf d_views, counts, work, totals, frame, spill;
    DevBuf slot_buf[11];
    pulse_batch_config cfg{};
    uint32_t stack_entries = 0;  // resolved from cfg
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    // shadow launches on the context's second stream beside the next
    // bounce's extension launches (the context's PULSE_OPT_CONCURRENT_SHADOW
    // when the configuration was applied; persistent traversal only)
    bool concurrent = false;
};

namespace {

enum { S_RAY_O0, S_RAY_O1, S_RAY_D0, S_RAY_D1, S_HIT, S_STATE_T, S_STATE_C, S_SH_O, S_SH_D, S_SH_C, S_HIT_V };

constexpr uint32_t kStacks[] = {24, 32, 40, 48, 64, 128};

bool stack_allowed(uint32_t s, uint32_t loop) {
    if (loop == 0) return s == 32 || s == 48 || s == 64 || s == 128;
    for (uint32_t x : kStacks)
        if (x == s) return true;
    return false;
}

uint32_t resolve_stack(const pulse_batch *b, const pulse_batch_config &c) {
    if (c.traversal_loop == 2 && !b->has_wide) return 0;
    const uint32_t need = c.traversal_loop == 2 ? b->wide_need : b->stack_need;
    if (c.stack_entries == 0) return pick_stack(need, static_cast<int>(c.traversal_loop));
    if (c.stack_entries < need || !stack_allowed(c.stack_entries, c.traversal_loop)) return 0;
    return c.stack_entries;
}

int check_config(const pulse_batch *b, const pulse_batch_config &c) {
    if (c.chunk_rays < 64 || c.chunk_rays > 4096 || (c.chunk_rays & (c.chunk_rays - 1)) != 0)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: chunk_rays must be a power of two in 64..4096 (got %u)",
                   c.chunk_rays);
    if (c.refill_lanes < 1 || c.refill_lanes > 64)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: refill_lanes in 1..64 (got %u)", c.refill_lanes);
    if (c.lds_stack_entries > 128)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: lds_stack_entries in 0..128 (got %u)", c.lds_stack_entries);
    if (c.persistent_waves_per_cu > 64)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: persistent_waves_per_cu above 64 (got %u)", c.persistent_waves_per_cu);
    if (c.traversal_loop > 2) PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: traversal_loop must be 0, 1 or 2");
    if (c.traversal_loop == 2 && !b->has_wide)
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "batch config: traversal_loop 2 needs 4-wide BVHs in every view's scene");
    if (c.regroup_every < 1 || c.regroup_every > 8) PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: regroup_every in 1..8");
    if (c.tile_width != 4 && c.tile_width != 8 && c.tile_width != 16 && c.tile_width != 32 && c.tile_width != 64)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: tile_width must be 4, 8, 16, 32 or 64 (got %u)", c.tile_width);
    if (c.use_graph > 1) PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: use_graph must be 0 or 1");
    if (resolve_stack(b, c) == 0)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "batch config: stack_entries %u unusable (needs >= %u, a compiled size)",
                   c.stack_entries, c.traversal_loop == 2 ? b->wide_need : b->stack_need);
    return PULSE_OK;
}

void drop_graph(pulse_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

BatchArgs batch_args(const pulse_batch *b) {
    BatchArgs a;
    a.views = b->d_views.as<BatchView>();
    a.n_views = b->n_views;
    a.cstride = b->cstride;
    a.ray_o[0] = b->slot_buf[S_RAY_O0].as<float4>();
    a.ray_o[1] = b->slot_buf[S_RAY_O1].as<float4>();
    a.ray_d[0] = b->slot_buf[S_RAY_D0].as<float4>();
    a.ray_d[1] = b->slot_buf[S_RAY_D1].as<float4>();
    a.hit = b->slot_buf[S_HIT].as<float4>();
    a.hit_v = b->slot_buf[S_HIT_V].as<float>();
    a.state_t = b->slot_buf[S_STATE_T].as<float4>();
    a.state_c = b->slot_buf[S_STATE_C].as<float4>();
    a.sh_o = b->slot_buf[S_SH_O].as<float4>();
    a.sh_d = b->slot_buf[S_SH_D].as<float4>();
    a.sh_c = b->slot_buf[S_SH_C].as<float4>();
    a.counts = b->counts.as<uint32_t>();
    a.work = b->work.as<uint32_t>();
    a.totals = b->totals.as<uint64_t>();
    a.frame = b->frame.as<uint32_t>();
    a.mode = b->prm.mode;
    a.nee_ggx = b->prm.nee_eval == PULSE_NEE_GGX ? 1u : 0u;
    a.max_traces = b->traces;
    a.rr_start = b->prm.rr_start_depth;
    a.horizon_clamp = (b->prm.flags & PULSE_FLAG_HORIZON_CLAMP) ? 1u : 0u;
    a.light_power = (b->prm.flags & PULSE_FLAG_LIGHT_POWER) ? 1u : 0u;
    a.chunk = b->cfg.chunk_rays;
    a.persistent = b->cfg.persistent_waves_per_cu;
    a.refill = b->cfg.refill_lanes;
    // the context's instance-change gathering, tuned per traversal loop as
    // for single views (PULSE_OPT_GATHER_LANES, PULSE_OPT_ADVANCE_PERIOD)
    const bool wide = b->cfg.traversal_loop == 2;
    a.gather = static_cast<uint32_t>(b->ctx->gather ? b->ctx->gather : (wide ? 64 : 1));
    a.period = static_cast<uint32_t>(b->ctx->advance_period ? b->ctx->advance_period : (wide ? 2 : 1));
    a.stack_cap = static_cast<int>(b->stack_entries);
    a.spill = b->spill.as<uint32_t>();
    a.regroup = b->cfg.regroup_every;
    a.tile_w = b->cfg.tile_width;
    return a;
}

// Per-view generate tiling and chunk totals for the current config; uploads
// the view table.
int layout_views(pulse_batch *b) {
    const uint32_t tw = b->cfg.tile_width, th = 64u / tw;
    uint32_t tile = 0, chunks = 0;
    for (auto &v : b->views) {
        v.tiles_x = (v.cam.width + tw - 1) / tw;
        uint32_t tiles_y = (v.cam.height + th - 1) / th;
        v.tile_base = tile;
        tile += v.tiles_x * tiles_y;
        chunks += (v.npix + b->cfg.chunk_rays - 1) / b->cfg.chunk_rays;
    }
    b->total_tiles = tile;
    b->max_chunks = chunks;
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    HIP_TRY(hipMemcpy(b->d_views.ptr, b->views.data(), sizeof(BatchView) * b->views.size(), hipMemcpyHostToDevice));
    return PULSE_OK;
}

uint32_t trace_grid(const pulse_batch *b);
template <int STACK, int LOOP>
void launch_trace(pulse_batch *b, const BatchArgs &a, bool shadow, int q, int k, hipStream_t st) {
    const uint32_t grid = trace_grid(b);
    if (shadow)
        hipLaunchKernelGGL((k_b_trace<STACK, LOOP, true>), dim3(grid), dim3(64), 0, st, a, q, k);
    else
        hipLaunchKernelGGL((k_b_trace<STACK, LOOP, false>), dim3(grid), dim3(64), 0, st, a, q, k);
}

// Persistent single-loop traversal with per-lane refill: the compiled LDS
// stack size at most the requested one (0: the whole stack).
constexpr int kLdsSizes[] = {8, 16, 24, 32, 40, 48, 64, 128};
bool refill_path(const pulse_batch *b) { return b->cfg.persistent_waves_per_cu > 0 && b->cfg.traversal_loop != 0; }
int refill_lds(const pulse_batch *b) {
    int want = b->cfg.lds_stack_entries > 0 ? std::min<int>(b->cfg.lds_stack_entries, b->stack_entries)
                                            : static_cast<int>(b->stack_entries);
    int got = kLdsSizes[0];
    for (int s : kLdsSizes)
        if (s <= want) got = s;
    return got;
}
uint32_t trace_grid(const pulse_batch *b) {
    uint32_t grid = b->max_chunks;
    if (b->cfg.persistent_waves_per_cu)
        grid = std::min<uint32_t>(grid, static_cast<uint32_t>(b->ctx->num_cus) * b->cfg.persistent_waves_per_cu);
    return std::max<uint32_t>(grid, 1);
}
// A persistent launch that runs beside another (the shadow launch of bounce
// k and the extension launch of bounce k+1) takes half the waves, so that
// both are resident at once; a launch running alone takes them all.
uint32_t pair_grid(const pulse_batch *b) {
    const uint32_t half = std::max<uint32_t>(b->cfg.persistent_waves_per_cu / 2, 1);
    return std::max<uint32_t>(std::min<uint32_t>(b->max_chunks, static_cast<uint32_t>(b->ctx->num_cus) * half), 1);
}
// Spilled stack words of one persistent launch (0: the whole stack in LDS
// or not persistent); concurrent launches take one part each.
size_t spill_part_words(const pulse_batch *b) {
    if (!refill_path(b) || static_cast<int>(b->stack_entries) <= refill_lds(b)) return 0;
    return static_cast<size_t>(trace_grid(b)) * 64u * (b->stack_entries - static_cast<uint32_t>(refill_lds(b)));
}
template <int LDS, int LOOP>
void launch_refill_loop(const BatchArgs &a, bool shadow, int q, int k, hipStream_t st, uint32_t grid) {
    if (shadow)
        hipLaunchKernelGGL((k_b_trace_refill<LDS, true, LOOP>), dim3(grid), dim3(64), 0, st, a,
                           static_cast<const BatchView *>(a.views), q, k);
    else
        hipLaunchKernelGGL((k_b_trace_refill<LDS, false, LOOP>), dim3(grid), dim3(64), 0, st, a,
                           static_cast<const BatchView *>(a.views), q, k);
}
template <int LDS>
void launch_refill(pulse_batch *b, const BatchArgs &a, bool shadow, int q, int k, hipStream_t st, uint32_t grid) {
    if (b->cfg.traversal_loop == 2) launch_refill_loop<LDS, 2>(a, shadow, q, k, st, grid);
    else launch_refill_loop<LDS, 1>(a, shadow, q, k, st, grid);
}

// paired: the launch runs beside another one (persistent launches only).
void dispatch_trace(pulse_batch *b, const BatchArgs &a, bool shadow, int q, int k, hipStream_t st,
                    bool paired = false) {
    if (refill_path(b)) {
        const uint32_t g = paired ? pair_grid(b) : trace_grid(b);
        switch (refill_lds(b)) {
            case 8: launch_refill<8>(b, a, shadow, q, k, st, g); break;
            case 16: launch_refill<16>(b, a, shadow, q, k, st, g); break;
            case 24: launch_refill<24>(b, a, shadow, q, k, st, g); break;
            case 32: launch_refill<32>(b, a, shadow, q, k, st, g); break;
            case 40: launch_refill<40>(b, a, shadow, q, k, st, g); break;
            case 48: launch_refill<48>(b, a, shadow, q, k, st, g); break;
            case 64: launch_refill<64>(b, a, shadow, q, k, st, g); break;
            default: launch_refill<128>(b, a, shadow, q, k, st, g); break;
        }
        return;
    }
    if (b->cfg.traversal_loop == 0) {
        switch (b->stack_entries) {
            case 32: launch_trace<32, 0>(b, a, shadow, q, k, st); break;
            case 48: launch_trace<48, 0>(b, a, shadow, q, k, st); break;
            case 64: launch_trace<64, 0>(b, a, shadow, q, k, st); break;
            default: launch_trace<128, 0>(b, a, shadow, q, k, st); break;
        }
    } else if (b->cfg.traversal_loop == 1) {
        switch (b->stack_entries) {
            case 24: launch_trace<24, 1>(b, a, shadow, q, k, st); break;
            case 32: launch_trace<32, 1>(b, a, shadow, q, k, st); break;
            case 40: launch_trace<40, 1>(b, a, shadow, q, k, st); break;
            case 48: launch_trace<48, 1>(b, a, shadow, q, k, st); break;
            case 64: launch_trace<64, 1>(b, a, shadow, q, k, st); break;
            default: launch_trace<128, 1>(b, a, shadow, q, k, st); break;
        }
    } else {
        switch (b->stack_entries) {
            case 24: launch_trace<24, 2>(b, a, shadow, q, k, st); break;
            case 32: launch_trace<32, 2>(b, a, shadow, q, k, st); break;
            case 40: launch_trace<40, 2>(b, a, shadow, q, k, st); break;
            case 48: launch_trace<48, 2>(b, a, shadow, q, k, st); break;
            case 64: launch_trace<64, 2>(b, a, shadow, q, k, st); break;
            default: launch_trace<128, 2>(b, a, shadow, q, k, st); break;
        }
    }
}

// All launches of one frame (prm.spp samples) on the context stream.
void enqueue_frame(pulse_batch *b) {
    const BatchArgs a = batch_args(b);
    hipStream_t st = b->ctx->stream;
    const uint32_t n_work = 3u * b->traces;
    const uint32_t clear_n = b->n_views * b->cstride + n_work;
    uint32_t shade_waves = b->max_chunks;
    if (b->cfg.persistent_waves_per_cu)
        shade_waves = std::min<uint32_t>(shade_waves, static_cast<uint32_t>(b->ctx->num_cus) * b->cfg.persistent_waves_per_cu);
    const uint32_t shade_grid = std::max<uint32_t>((shade_waves + 3) / 4, 1);
    // concurrent: the shadow launches take the second spill part and the
    // second stream, forked after each shade and joined before the next
    BatchArgs a2 = a;
    if (b->concurrent && a2.spill) a2.spill += spill_part_words(b);
    for (uint32_t s = 0; s < std::max(b->prm.spp, 1u); ++s) {
        hipLaunchKernelGGL(k_b_clear, dim3((clear_n + 255) / 256), dim3(256), 0, st, a, n_work);
        hipLaunchKernelGGL(k_b_generate, dim3((b->total_tiles + 3) / 4), dim3(256), 0, st, a, s);
        int q = 0;
        if (b->concurrent) {
            hipStream_t st2 = b->ctx->stream2;
            dispatch_trace(b, a, false, q, 0, st);
            for (uint32_t k = 0; k < b->traces; ++k) {
                hipLaunchKernelGGL(k_b_shade, dim3(shade_grid), dim3(256), 0, st, a, q, static_cast<int>(k));
                (void)hipEventRecord(b->ctx->fork_ev[k], st);
                (void)hipStreamWaitEvent(st2, b->ctx->fork_ev[k], 0);
                const bool pair = k + 1 < b->traces;  // the last shadow launch runs alone
                dispatch_trace(b, a2, true, q, static_cast<int>(k), st2, pair);
                (void)hipEventRecord(b->ctx->join_ev[k], st2);
                q ^= 1;
                if (pair) dispatch_trace(b, a, false, q, static_cast<int>(k + 1), st, true);
                (void)hipStreamWaitEvent(st, b->ctx->join_ev[k], 0);
            }
        } else {
            for (uint32_t k = 0; k < b->traces; ++k) {
                dispatch_trace(b, a, false, q, static_cast<int>(k), st);
                hipLaunchKernelGGL(k_b_shade, dim3(shade_grid), dim3(256), 0, st, a, q, static_cast<int>(k));
                dispatch_trace(b, a, true, q, static_cast<int>(k), st);
                q ^= 1;
            }
        }
        hipLaunchKernelGGL(k_b_accumulate, dim3(b->total_slots / 256), dim3(256), 0, st, a, s);
    }
}

int build_graph(pulse_batch *b) {
    drop_graph(b);
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueue_frame(b);
    hipError_t e = hipStreamEndCapture(st, &b->graph);
    if (e != hipSuccess) {
        b->graph = nullptr;
        PULSE_FAIL(PULSE_E_HIP, "pulse_batch: graph capture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        PULSE_FAIL(PULSE_E_HIP, "pulse_batch: graph instantiation failed: %s", hipGetErrorString(e));
    }
    b->graph_valid = true;
    return PULSE_OK;
}

int apply_config(pulse_batch *b, const pulse_batch_config &c) {
    int rc = check_config(b, c);
    if (rc) return rc;
    b->cfg = c;
    b->stack_entries = resolve_stack(b, c);
    drop_graph(b);
    if ((rc = layout_views(b)) != PULSE_OK) return rc;  // (drains the stream)
    // concurrent shadow launches with direct launches only: a graph replay
    // runs the two streams' branches one after the other (measured on the
    // 32-view workload: replay with the fork/join 88.9 ms, replay in sequence
    // 87.6 ms, direct launches with concurrent pairs 83.0 ms)
    b->concurrent = refill_path(b) && b->ctx->concurrent_shadow != 0 && c.use_graph == 0;
    if (b->concurrent && (rc = ensure_stream2(b->ctx, false)) != PULSE_OK) return rc;
    if (const size_t words = spill_part_words(b)) {
        if ((rc = b->spill.ensure(sizeof(uint32_t) * words * (b->concurrent ? 2 : 1))) != PULSE_OK) return rc;
    }
    return PULSE_OK;
}

// Picks up instance updates of the views' scenes (new TLAS, maybe moved
// buffers): refreshes the view table read by the kernels.
int refresh_scenes(pulse_batch *b) {
    bool changed = false;
    for (uint32_t i = 0; i < b->n_views; ++i) {
        const SceneData &sd = b->scenes[i]->data;
        if (int rc = check_traversal(sd, static_cast<int>(b->cfg.traversal_loop))) return rc;
        if (sd.generation == b->scene_gen[i]) continue;
        if (!sd.valid) PULSE_FAIL(PULSE_E_NO_SCENE, "pulse_batch_render: view %u: its scene is not valid", i);
        const bool wide = b->cfg.traversal_loop == 2;
        if (wide && !sd.has_wide)
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_batch_render: view %u: the updated scene has no 4-wide BVH", i);
        const uint32_t need = wide ? sd.wide_need : sd.stack_need;
        if (need > b->stack_entries)
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_batch_render: view %u: the updated scene needs %u stack entries, the "
                       "batch has %u (set a larger stack_entries)", i, need, b->stack_entries);
        b->views[i].sc = dev_scene(sd, b->ctx->errors.as<uint32_t>());
        b->scene_gen[i] = sd.generation;
        changed = true;
    }
    if (!changed) return PULSE_OK;
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    HIP_TRY(hipMemcpy(b->d_views.ptr, b->views.data(), sizeof(BatchView) * b->views.size(), hipMemcpyHostToDevice));
    return PULSE_OK;
}

int render(pulse_batch *b, uint32_t acc) {
    int rc0 = refresh_scenes(b);
    if (rc0) return rc0;
    hipStream_t st = b->ctx->stream;
    hipLaunchKernelGGL(k_b_set_frame, dim3(1), dim3(1), 0, st, b->frame.as<uint32_t>(), acc);
    if (b->cfg.use_graph) {
        int rc;
        if (!b->graph_valid && (rc = build_graph(b)) != PULSE_OK) return rc;
        HIP_TRY(hipGraphLaunch(b->exec, st));
    } else {
        enqueue_frame(b);
    }
    HIP_TRY(hipGetLastError());
    return PULSE_OK;
}

}  // namespace

extern "C" {

int pulse_batch_default_config(pulse_batch_config *out) {
    if (!out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_default_config: null out");
    *out = pulse_batch_config{};
    out->persistent_waves_per_cu = 24;
    out->chunk_rays = 256;
    out->stack_entries = 0;
    out->traversal_loop = 2;
    out->regroup_every = 1;
    out->tile_width = 8;
    out->use_graph = 0;
    out->lds_stack_entries = 24;
    out->refill_lanes = 16;
    return PULSE_OK;
}

int pulse_batch_create(pulse_ctx *ctx, const pulse_view *views, uint32_t n_views, const pulse_params *prm,
                       pulse_batch **out) {
    if (!ctx || !views || !prm || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_create: null argument");
    if (n_views == 0 || n_views > kBatchMaxViews)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_create: %u views (1..%u)", n_views, kBatchMaxViews);
    if (prm->mode > PULSE_MODE_GGX_NEE) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_create: unknown mode %u", prm->mode);
    if (prm->nee_eval > PULSE_NEE_GGX) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_create: unknown nee_eval %u", prm->nee_eval);
    const uint32_t traces = prm->mode == PULSE_MODE_REF_NEE_LAMBERT ? prm->max_depth + 1u : prm->max_depth;
    if (traces > static_cast<uint32_t>(kMaxTraces))
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_batch_create: max_depth %u above %d", prm->max_depth, kMaxTraces - 1);
    HIP_TRY(hipSetDevice(ctx->device));
    auto *b = new pulse_batch();
    b->ctx = ctx;
    b->prm = *prm;
    b->n_views = n_views;
    b->traces = traces;
    b->cstride = 2u * (traces + 1u);
    uint64_t slots = 0;
    auto fail = [&](int rc) {
        pulse_batch_destroy(b);
        return rc;
    };
    for (uint32_t i = 0; i < n_views; ++i) {
        const pulse_view &v = views[i];
        if (v.scene && (prm->flags & PULSE_FLAG_LIGHT_POWER) && v.scene->data.valid &&
            v.scene->data.n_strength_cdf < v.scene->data.uniform.light_count) {
            pulse_batch_destroy(b);
            PULSE_FAIL(PULSE_E_LAYOUT, "pulse_batch_create: view %u: PULSE_FLAG_LIGHT_POWER needs one "
                       "light_emission_strength_cdf entry per light", i);
        }
        if (!v.scene || v.scene->ctx != ctx || !v.scene->data.valid || !v.accum_rgba || v.camera.width == 0 ||
            v.camera.height == 0) {
            pulse_batch_destroy(b);
            PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_create: view %u: needs a scene of this context, a camera and "
                       "an accumulation buffer", i);
        }
        BatchView bv{};
        bv.sc = dev_scene(v.scene->data, ctx->errors.as<uint32_t>());
        bv.cam = v.camera;
        bv.accum = reinterpret_cast<float4 *>(v.accum_rgba);
        bv.npix = v.camera.width * v.camera.height;
        if (static_cast<uint64_t>(v.camera.width) * v.camera.height > (1ull << 28)) {
            pulse_batch_destroy(b);
            PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_batch_create: view %u too large", i);
        }
        bv.pix_base = static_cast<uint32_t>(slots);
        slots += (bv.npix + 255u) / 256u * 256u;
        b->views.push_back(bv);
        b->scenes.push_back(v.scene);
        b->scene_gen.push_back(v.scene->data.generation);
        b->stack_need = std::max(b->stack_need, v.scene->data.stack_need);
        b->wide_need = std::max(b->wide_need, v.scene->data.wide_need);
        b->has_wide = b->has_wide && v.scene->data.has_wide;
    }
    if (slots > (1ull << 31)) {
        pulse_batch_destroy(b);
        PULSE_FAIL(PULSE_E_UNSUPPORTED, "pulse_batch_create: %llu pixel slots above 2^31", (unsigned long long)slots);
    }
    b->total_slots = static_cast<uint32_t>(slots);
    int rc;
    for (int i = 0; i < 11; ++i) {
        size_t bytes = (i == S_HIT_V ? sizeof(float) : sizeof(float4)) * slots;
        if ((rc = b->slot_buf[i].ensure(bytes)) != PULSE_OK) return fail(rc);
    }
    if ((rc = b->d_views.ensure(sizeof(BatchView) * n_views)) || (rc = b->counts.ensure(4ull * n_views * b->cstride)) ||
        (rc = b->work.ensure(4ull * 3u * traces)) || (rc = b->totals.ensure(2 * sizeof(uint64_t))) ||
        (rc = b->frame.ensure(sizeof(uint32_t))))
        return fail(rc);
    if (hipMemset(b->totals.ptr, 0, 2 * sizeof(uint64_t)) != hipSuccess) {
        pulse_batch_destroy(b);
        PULSE_FAIL(PULSE_E_HIP, "pulse_batch_create: memset failed");
    }
    pulse_batch_config c;
    pulse_batch_default_config(&c);
    if (!b->has_wide) c.traversal_loop = 1;
    if ((rc = apply_config(b, c)) != PULSE_OK) return fail(rc);
    *out = b;
    return PULSE_OK;
}

int pulse_batch_destroy(pulse_batch *b) {
    if (!b) return PULSE_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (auto &x : b->slot_buf) x.release();
    b->d_views.release();
    b->counts.release();
    b->work.release();
    b->totals.release();
    b->frame.release();
    b->spill.release();
    delete b;
    return PULSE_OK;
}

int pulse_batch_set_camera(pulse_batch *b, uint32_t view, const pulse_camera *cam) {
    if (!b || !cam || view >= b->n_views) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_set_camera: bad argument");
    if (cam->width != b->views[view].cam.width || cam->height != b->views[view].cam.height)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_set_camera: resolution is fixed at creation (%ux%u)",
                   b->views[view].cam.width, b->views[view].cam.height);
    HIP_TRY(hipSetDevice(b->ctx->device));
    b->views[view].cam = *cam;
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    HIP_TRY(hipMemcpy(b->d_views.as<BatchView>() + view, &b->views[view], sizeof(BatchView), hipMemcpyHostToDevice));
    return PULSE_OK;
}

int pulse_batch_set_config(pulse_batch *b, const pulse_batch_config *cfg) {
    if (!b || !cfg) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_set_config: null argument");
    HIP_TRY(hipSetDevice(b->ctx->device));
    return apply_config(b, *cfg);
}

int pulse_batch_get_config(pulse_batch *b, pulse_batch_config *out) {
    if (!b || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_get_config: null argument");
    *out = b->cfg;
    out->stack_entries = b->stack_entries;
    return PULSE_OK;
}

int pulse_batch_render(pulse_batch *b, uint32_t accumulation_count) {
    if (!b) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_render: null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    return render(b, accumulation_count);
}

int pulse_batch_get_stats(pulse_batch *b, pulse_stats *out) {
    if (!b || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_get_stats: null argument");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    uint64_t t[2] = {0, 0};
    uint32_t err = 0;
    HIP_TRY(hipMemcpy(t, b->totals.ptr, sizeof(t), hipMemcpyDeviceToHost));
    int rc = pulse_host::take_device_errors(b->ctx, &err);
    if (rc != PULSE_OK) return rc;
    HIP_TRY(hipMemset(b->totals.ptr, 0, sizeof(t)));
    *out = pulse_stats{};
    out->extension_rays = t[0];
    out->shadow_rays = t[1];
    out->error_flags = err;
    out->stack_entries = b->stack_entries;
    return PULSE_OK;
}

int pulse_batch_autotune(pulse_batch *b, uint32_t reps, pulse_batch_config *best_out, double *best_ms_out) {
    if (!b) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_batch_autotune: null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    reps = std::max(reps, 1u);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = PULSE_OK;
    std::vector<std::pair<pulse_batch_config, double>> seen;
    auto time_config = [&](const pulse_batch_config &c, double &ms) -> int {
        for (const auto &x : seen)
            if (std::memcmp(&x.first, &c, sizeof(c)) == 0) {
                ms = x.second;
                return PULSE_OK;
            }
        int r = apply_config(b, c);
        if (r) return r;
        if (b->cfg.use_graph && (r = build_graph(b)) != PULSE_OK) return r;
        double best = 1e300;
        for (uint32_t i = 0; i < reps; ++i) {
            HIP_TRY(hipEventRecord(e0, b->ctx->stream));
            if ((r = render(b, 0)) != PULSE_OK) return r;
            HIP_TRY(hipEventRecord(e1, b->ctx->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float f = 0;
            HIP_TRY(hipEventElapsedTime(&f, e0, e1));
            best = std::min(best, static_cast<double>(f));
        }
        ms = best;
        seen.push_back({c, best});
        return PULSE_OK;
    };
    pulse_batch_config best = b->cfg;
    double best_ms = 0;
    // coordinate descent over the knobs, starting from the current config
    // (one untimed frame first: first-touch of the queues and scenes)
    if ((rc = render(b, 0)) == PULSE_OK && (rc = time_config(best, best_ms)) == PULSE_OK) {
        // the work distribution and the chunk size interact: search them jointly first
        for (uint32_t w : {0u, 16u, 24u, 32u}) {
            for (uint32_t ch : {64u, 256u, 1024u}) {
                pulse_batch_config c = best;
                c.persistent_waves_per_cu = w;
                c.chunk_rays = ch;
                double ms = 0;
                if (check_config(b, c) != PULSE_OK) continue;
                if ((rc = time_config(c, ms)) != PULSE_OK) break;
                if (ms < best_ms * 0.98) {
                    best_ms = ms;
                    best = c;
                }
            }
            if (rc) break;
        }
    }
    if (rc == PULSE_OK) {
        std::vector<uint32_t> stack_opts = {0};
        for (uint32_t s : kStacks)
            if (s > pick_stack(std::max(b->stack_need, b->wide_need), 1) && stack_opts.size() < 2) stack_opts.push_back(s);
        for (int pass = 0; pass < 2 && rc == PULSE_OK; ++pass) {
            for (int knob = 0; knob < 9 && rc == PULSE_OK; ++knob) {
                std::vector<pulse_batch_config> cands;
                auto add = [&](auto set) {
                    pulse_batch_config c = best;
                    set(c);
                    if (check_config(b, c) == PULSE_OK) cands.push_back(c);
                };
                switch (knob) {
                    case 0:
                        for (uint32_t w : {0u, 16u, 24u, 32u})
                            add([&](pulse_batch_config &c) { c.persistent_waves_per_cu = w; });
                        break;
                    case 1:
                        for (uint32_t ch : {64u, 256u, 1024u}) add([&](pulse_batch_config &c) { c.chunk_rays = ch; });
                        break;
                    case 6:
                        for (uint32_t l : {16u, 24u, 32u}) add([&](pulse_batch_config &c) { c.lds_stack_entries = l; });
                        break;
                    case 7:
                        for (uint32_t r : {8u, 16u, 32u}) add([&](pulse_batch_config &c) { c.refill_lanes = r; });
                        break;
                    case 2:
                        for (uint32_t r : {1u, 2u}) add([&](pulse_batch_config &c) { c.regroup_every = r; });
                        break;
                    case 3:
                        for (uint32_t l : {0u, 1u, 2u})
                            add([&](pulse_batch_config &c) {
                                c.traversal_loop = l;
                                c.stack_entries = 0;
                            });
                        break;
                    case 4:
                        for (uint32_t s : stack_opts) add([&](pulse_batch_config &c) { c.stack_entries = s; });
                        break;
                    case 5:
                        for (uint32_t t : {8u, 16u, 64u}) add([&](pulse_batch_config &c) { c.tile_width = t; });
                        break;
                    case 8:
                        for (uint32_t g : {0u, 1u}) add([&](pulse_batch_config &c) { c.use_graph = g; });
                        break;
                }
                for (const auto &c : cands) {
                    double ms = 0;
                    if ((rc = time_config(c, ms)) != PULSE_OK) break;
                    if (ms < best_ms * 0.98) {
                        best_ms = ms;
                        best = c;
                    }
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = apply_config(b, best)) != PULSE_OK) return rc;
    if (best_out) pulse_batch_get_config(b, best_out);
    if (best_ms_out) *best_ms_out = best_ms;
    return PULSE_OK;
}

int pulse_trace_batch(pulse_ctx *ctx, const pulse_view *views, uint32_t n_views, const pulse_params *prm) {
    pulse_batch *b = nullptr;
    int rc = pulse_batch_create(ctx, views, n_views, prm, &b);
    if (rc) return rc;
    rc = render(b, prm->accumulation_count);
    int rc2 = pulse_batch_destroy(b);
    return rc ? rc : rc2;
}

}  // extern "C"
