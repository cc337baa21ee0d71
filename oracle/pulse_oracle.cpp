// pulse_oracle.cpp -- TEST INFRASTRUCTURE ONLY: C entry points of the CPU
// oracle (pulse_oracle.hpp) for the Python tests and the CPU baseline.
// Multithreaded drivers use std::thread with dynamic row scheduling.

#include "pulse_oracle.hpp"

#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdlib>
#include <thread>
#include <vector>

using namespace oracle;

namespace {

struct OwnedScene {
    Scene s;
    std::vector<char> storage[13];
};

template <class F>
void parallel_for(uint32_t n, int threads, F &&fn) {
    if (threads <= 1 || n < 2) {
        for (uint32_t i = 0; i < n; ++i) fn(i, 0);
        return;
    }
    std::atomic<uint32_t> next{0};
    auto worker = [&](int tid) {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i, tid);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &t : pool) t.join();
}

void export_counters(const std::vector<Counters> &per_thread, uint64_t *out) {
    if (!out) return;
    Counters c;
    for (auto &x : per_thread) c.add(x);
    uint64_t v[11] = {c.tlas_interior, c.tlas_leaf, c.blas_interior, c.blas_leaf, c.instances_entered,
                      c.tri_tests, c.rays, c.max_stack, c.iteration_cap_hits, c.shadow_rays, c.primary_rays};
    std::memcpy(out, v, sizeof(v));
    std::memcpy(out + 11, c.cat, sizeof(c.cat));
}

}  // namespace

extern "C" {

typedef struct oracle_scene oracle_scene;

// Copies the 13 host buffers so the caller may free its arrays.
oracle_scene *oracle_scene_create(const pulse_scene_desc *d) {
    if (!d) return nullptr;
    auto *o = new OwnedScene();
    const pulse_buffer_view *views[13] = {nullptr,
                                          &d->primitives,
                                          &d->triangle_data,
                                          &d->triangle_indices,
                                          &d->blas_nodes,
                                          &d->tlas_nodes,
                                          &d->instance_indices,
                                          &d->instances,
                                          &d->materials,
                                          &d->light_emission_strength_cdf,
                                          &d->light_triangle_area_cdfs,
                                          &d->light_mesh_areas,
                                          &d->light_indices};
    for (int i = 1; i < 13; ++i) {
        o->storage[i].resize(views[i]->bytes + 16);
        if (views[i]->bytes) std::memcpy(o->storage[i].data(), views[i]->data, views[i]->bytes);
    }
    Scene &s = o->s;
    s.uniform = d->uniform;
    s.primitives = reinterpret_cast<const pulse_primitive *>(o->storage[1].data());
    s.triangle_data = reinterpret_cast<const pulse_triangle_data *>(o->storage[2].data());
    s.triangle_indices = reinterpret_cast<const uint32_t *>(o->storage[3].data());
    s.blas_nodes = reinterpret_cast<const pulse_blas_node *>(o->storage[4].data());
    s.tlas_nodes = reinterpret_cast<const pulse_tlas_node *>(o->storage[5].data());
    s.instance_indices = reinterpret_cast<const uint32_t *>(o->storage[6].data());
    s.instances = reinterpret_cast<const pulse_mesh_instance *>(o->storage[7].data());
    s.materials = reinterpret_cast<const pulse_material *>(o->storage[8].data());
    s.light_triangle_area_cdfs = reinterpret_cast<const float *>(o->storage[10].data());
    s.light_mesh_areas = reinterpret_cast<const float *>(o->storage[11].data());
    s.light_emission_strength_cdf = reinterpret_cast<const float *>(o->storage[9].data());
    s.n_strength_cdf = d->light_emission_strength_cdf.bytes / 4;
    s.light_indices = reinterpret_cast<const pulse_light_data_index *>(o->storage[12].data());
    s.n_primitives = d->primitives.bytes / sizeof(pulse_primitive);
    s.n_instances = d->instances.bytes / sizeof(pulse_mesh_instance);
    s.n_tlas = d->tlas_nodes.bytes / sizeof(pulse_tlas_node);
    s.n_blas = d->blas_nodes.bytes / sizeof(pulse_blas_node);
    return reinterpret_cast<oracle_scene *>(o);
}

void oracle_scene_destroy(oracle_scene *sc) { delete reinterpret_cast<OwnedScene *>(sc); }

void oracle_scene_set_cdf_linear_scan(oracle_scene *sc, int on) {
    reinterpret_cast<OwnedScene *>(sc)->s.cdf_linear_scan = on != 0;
}

void oracle_scene_set_cull_world(oracle_scene *sc, int on) { reinterpret_cast<OwnedScene *>(sc)->s.cull_world = on != 0; }

static const Scene &S(oracle_scene *sc) { return reinterpret_cast<OwnedScene *>(sc)->s; }

// Closest hit per ray, f32.  counters: 11 u64 (see export_counters).
int oracle_trace_rays_f32(oracle_scene *sc, const pulse_ray *rays, pulse_hit *hits, uint32_t n, int threads,
                          int bruteforce, uint64_t *counters) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    parallel_for(n, threads, [&](uint32_t i, int tid) {
        Ray<float> r;
        r.origin = load3<float>(rays[i].origin);
        r.dir = load3<float>(rays[i].dir);
        r.t = rays[i].t_max;
        if (bruteforce) {
            pc[tid].rays++;
            trace_bruteforce(s, r, pc[tid]);
        } else {
            traverse_tlas(s, r, pc[tid]);
        }
        hits[i] = pulse_hit{r.t, r.instance_index, r.triangle_index, r.u, r.v};
    });
    export_counters(pc, counters);
    return 0;
}

// Closest hit per ray in f64 (t, u, v as doubles; ids as u32 pairs).
int oracle_trace_rays_f64(oracle_scene *sc, const pulse_ray *rays, double *tuv, uint32_t *ids, uint32_t n,
                          int threads) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    parallel_for(n, threads, [&](uint32_t i, int tid) {
        Ray<double> r;
        r.origin = load3<double>(rays[i].origin);
        r.dir = load3<double>(rays[i].dir);
        r.t = rays[i].t_max;
        traverse_tlas(s, r, pc[tid]);
        tuv[3 * i] = r.t;
        tuv[3 * i + 1] = r.u;
        tuv[3 * i + 2] = r.v;
        ids[2 * i] = r.instance_index;
        ids[2 * i + 1] = r.triangle_index;
    });
    return 0;
}

// One given candidate per ray: the reference's triangle test of triangle
// slot ids[2i+1] of instance ids[2i] (object space, record t starting at
// 1e30, utilities.wgsl:430-460) and the world t through the hit point
// (318-320).  hits[i].t = 1e30 if the triangle is not hit.
int oracle_intersect_candidates_f32(oracle_scene *sc, const pulse_ray *rays, const uint32_t *ids, pulse_hit *hits,
                                    uint32_t n) {
    const Scene &s = S(sc);
    Counters c;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t inst = ids[2 * i], tri = ids[2 * i + 1];
        hits[i] = pulse_hit{kTFar, 0u, 0u, 0.0f, 0.0f};
        if (inst >= s.n_instances || tri >= s.instances[inst].triangle_count) return -1;
        const pulse_mesh_instance &mi = s.instances[inst];
        Ray<float> ro;
        V3<float> o = load3<float>(rays[i].origin);
        ro.origin = transform_position<float>(mi.world_object, o);
        ro.dir = normalize(transform_direction<float>(mi.world_object, load3<float>(rays[i].dir)));
        ro.t = kTFar;
        ray_triangle_intersect(s, ro, tri, inst, c);
        if (ro.t == kTFar) continue;
        V3<float> hw = transform_position<float>(mi.object_world, ro.origin + ro.t * ro.dir);
        hits[i] = pulse_hit{length(hw - o), inst, tri, ro.u, ro.v};
    }
    return 0;
}

// occluded[i] = closest-hit t < rays[i].t_max (t_max carries the threshold;
// the traversal itself starts from 1e30 as the reference's shadow ray does).
int oracle_shadow_rays_f32(oracle_scene *sc, const pulse_ray *rays, uint32_t *occluded, uint32_t n, int threads) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    parallel_for(n, threads, [&](uint32_t i, int tid) {
        Ray<float> r;
        r.origin = load3<float>(rays[i].origin);
        r.dir = load3<float>(rays[i].dir);
        r.t = kTFar;
        if (s.cull_world) {
            occluded[i] = traverse_tlas_anyhit(s, r, rays[i].t_max, pc[tid]) ? 1u : 0u;
            return;
        }
        traverse_tlas(s, r, pc[tid]);
        occluded[i] = r.t < rays[i].t_max ? 1u : 0u;
    });
    return 0;
}

// Renders params->spp samples per pixel into accum (W*H*4 f32, running mean).
// counters: 11 u64.  Returns wall-clock milliseconds of the render loop.
double oracle_render_f32(oracle_scene *sc, const pulse_camera *cam, const pulse_params *prm, float *accum,
                         int threads, uint64_t *counters, uint32_t row_begin, uint32_t row_end) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    if (row_end == 0 || row_end > cam->height) row_end = cam->height;
    auto t0 = std::chrono::steady_clock::now();
    parallel_for(row_end - row_begin, threads, [&](uint32_t rr, int tid) {
        uint32_t y = row_begin + rr;
        DecisionTape tape;
        for (uint32_t x = 0; x < cam->width; ++x) {
            float *px = accum + 4 * (static_cast<size_t>(y) * cam->width + x);
            V3<float> old{px[0], px[1], px[2]};
            for (uint32_t k = 0; k < prm->spp; ++k) {
                uint32_t n = prm->accumulation_count + k;
                V3<float> c = trace_path<float>(s, *cam, *prm, x, y, n, pc[tid], tape);
                old = accumulate(old, c, n);
            }
            px[0] = old.x;
            px[1] = old.y;
            px[2] = old.z;
            px[3] = 1.0f;
        }
    });
    auto t1 = std::chrono::steady_clock::now();
    export_counters(pc, counters);
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// The realtime GI pass over a G-buffer (host arrays of W*H float4: world
// position, world normal, base colour, emissive); out: W*H float4 = the
// averaged colour, alpha 1.  Returns ms.
double oracle_render_gi_f32(oracle_scene *sc, const pulse_params *prm, const float *pos, const float *nrm,
                            const float *base, const float *emis, uint32_t width, uint32_t height, float *out,
                            int threads, uint64_t *counters) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    auto t0 = std::chrono::steady_clock::now();
    parallel_for(height, threads, [&](uint32_t y, int tid) {
        DecisionTape tape;
        for (uint32_t x = 0; x < width; ++x) {
            const size_t i = static_cast<size_t>(y) * width + x;
            V3<float> c = trace_gi_pixel<float>(s, *prm, static_cast<uint32_t>(i), load3<float>(pos + 4 * i),
                                                load3<float>(nrm + 4 * i), load3<float>(base + 4 * i),
                                                load3<float>(emis + 4 * i), pc[tid], tape);
            out[4 * i] = c.x;
            out[4 * i + 1] = c.y;
            out[4 * i + 2] = c.z;
            out[4 * i + 3] = 1.0f;
        }
    });
    auto t1 = std::chrono::steady_clock::now();
    if (counters) export_counters(pc, counters);
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Renders the same samples in f32 and in f64.  replay != 0: the f64 path takes
// every branch decision from the f32 path (decision replay); flip_stats[0]
// counts decisions the f64 arithmetic would have taken differently and
// flip_stats[1] the pixels with at least one such decision.  replay == 0: the
// f64 path decides for itself; flip_stats then counts pixels whose decision
// sequence differs.
int oracle_render_pair(oracle_scene *sc, const pulse_camera *cam, const pulse_params *prm, float *accum32,
                       double *accum64, int threads, int replay, uint64_t *flip_stats) {
    const Scene &s = S(sc);
    std::vector<Counters> pc(std::max(1, threads));
    std::atomic<uint64_t> flips{0}, flip_px{0};
    parallel_for(cam->height, threads, [&](uint32_t y, int tid) {
        std::vector<uint64_t> log32, log64;
        for (uint32_t x = 0; x < cam->width; ++x) {
            size_t pix = static_cast<size_t>(y) * cam->width + x;
            V3<float> o32{accum32[4 * pix], accum32[4 * pix + 1], accum32[4 * pix + 2]};
            V3<double> o64{accum64[4 * pix], accum64[4 * pix + 1], accum64[4 * pix + 2]};
            bool pixel_flipped = false;
            for (uint32_t k = 0; k < prm->spp; ++k) {
                uint32_t n = prm->accumulation_count + k;
                log32.clear();
                DecisionTape rec;
                rec.mode = DecisionTape::RECORD;
                rec.log = &log32;
                V3<float> c32 = trace_path<float>(s, *cam, *prm, x, y, n, pc[tid], rec);
                DecisionTape t64;
                if (replay) {
                    t64.mode = DecisionTape::REPLAY;
                    t64.log = &log32;
                } else {
                    log64.clear();
                    t64.mode = DecisionTape::RECORD;
                    t64.log = &log64;
                }
                V3<double> c64 = trace_path<double>(s, *cam, *prm, x, y, n, pc[tid], t64);
                if (replay) {
                    flips += t64.flips;
                    if (t64.flips) pixel_flipped = true;
                } else if (log32 != log64) {
                    pixel_flipped = true;
                }
                o32 = accumulate(o32, c32, n);
                o64 = accumulate(o64, c64, n);
            }
            if (pixel_flipped) flip_px++;
            accum32[4 * pix] = o32.x;
            accum32[4 * pix + 1] = o32.y;
            accum32[4 * pix + 2] = o32.z;
            accum32[4 * pix + 3] = 1.0f;
            accum64[4 * pix] = o64.x;
            accum64[4 * pix + 1] = o64.y;
            accum64[4 * pix + 2] = o64.z;
            accum64[4 * pix + 3] = 1.0;
        }
    });
    if (flip_stats) {
        flip_stats[0] = flips;
        flip_stats[1] = flip_px;
    }
    return 0;
}

// Gradient of L = mean over pixels and rgb of (I - target)^2 by central
// differences in f64 along the f32 decisions (decision replay, no probes):
// with every decision fixed the f64 image is a smooth function of the
// material parameters.  params: n_materials*7 f64 values {base.rgb,
// roughness, emissive.rgb}; ids: parameter indices (7*m + j); h: step.
// Along replayed decisions a random roulette survival (p <= 1) keeps its
// weight 1/p and carries the likelihood ratio of the decision to later
// roulettes; a p above 1 (the path survives whatever the draw) follows the
// perturbed throughput.
// frozen_dirs = 0: the sampled directions follow the roughness (the
// reparameterised derivative); 1: the specular half vectors sampled at the
// unperturbed parameters (f64) are held fixed (the detached derivative);
// 2: no replay at all -- f64 renders whose decisions follow the perturbed
// parameters (the derivative of the estimator itself, visibility changes
// included; use a larger h).
// Returns 0; grad_out[i] = dL/dparam[ids[i]].  loss_out (optional) = L at
// the unperturbed parameters.
// Rows [row0, row1) only: the loss is the mean over those rows' pixels (the
// target holds the whole image); pixels whose target is NaN are left out of
// the loss and of its pixel count.  eval32 = 1: the perturbed renders are
// evaluated in f32 along the same decisions, with the half vectors frozen as
// the f32 sampling drew them (the f32 estimator itself; FD noise of f32
// arithmetic: use h of about 1e-3) instead of f64.  pixel_grad (optional, n_ids x W x
// (row1 - row0) doubles): each pixel's share of each derivative.
int oracle_grad_fd_rows(oracle_scene *sc, const pulse_camera *cam, const pulse_params *prm, const float *target,
                        const double *params, uint32_t n_materials, const uint32_t *ids, uint32_t n_ids, double h,
                        double *grad_out, double *loss_out, int threads, int frozen_dirs, uint32_t row0,
                        uint32_t row1, int eval32, double *pixel_grad) {
    OwnedScene *os = reinterpret_cast<OwnedScene *>(sc);
    Scene &s = os->s;
    const uint32_t W = cam->width, spp = std::max(prm->spp, 1u);
    row1 = std::min(row1, cam->height);
    if (row0 >= row1) return -1;
    const uint32_t H = row1 - row0, npix = W * H;
    // record the f32 decisions of every path once (index: row - row0)
    std::vector<std::vector<uint64_t>> logs(static_cast<size_t>(npix) * spp);
    std::vector<Counters> pc(std::max(1, threads));
    parallel_for(H, threads, [&](uint32_t yy, int tid) {
        const uint32_t y = row0 + yy;
        for (uint32_t x = 0; x < W; ++x)
            for (uint32_t k = 0; k < spp; ++k) {
                DecisionTape rec;
                rec.mode = DecisionTape::RECORD;
                rec.log = &logs[(static_cast<size_t>(yy) * W + x) * spp + k];
                trace_path<float>(s, *cam, *prm, x, y, prm->accumulation_count + k, pc[tid], rec);
            }
    });
    std::vector<double> p(params, params + 7 * static_cast<size_t>(n_materials));
    std::vector<std::vector<double>> dirs(frozen_dirs == 1 ? logs.size() : 0);
    DecisionTape::DirMode dmode = frozen_dirs == 1 ? DecisionTape::RECORD_DIRS : DecisionTape::DIRS_OFF;
    // roulette probabilities of the unperturbed paths, replayed along the
    // decisions
    std::vector<std::vector<double>> rrp(frozen_dirs == 2 ? 0 : logs.size());
    DecisionTape::RrMode rmode = frozen_dirs == 2 ? DecisionTape::RR_OFF : DecisionTape::RECORD_RR;
    uint64_t used = 0;
    for (uint32_t yy = 0; yy < H; ++yy)
        for (uint32_t x = 0; x < W; ++x) used += std::isnan(target[4 * (static_cast<size_t>(row0 + yy) * W + x)]) ? 0 : 1;
    if (used == 0) return -1;
    // one pixel's mean over its samples at precision R
    auto pixel = [&](auto zero, uint32_t yy, uint32_t x, int tid) {
        using R = decltype(zero);
        V3<R> I{R(0), R(0), R(0)};
        for (uint32_t k = 0; k < spp; ++k) {
            DecisionTape rep;
            rep.mode = frozen_dirs == 2 ? DecisionTape::OFF : DecisionTape::REPLAY;
            rep.probe = false;
            rep.log = &logs[(static_cast<size_t>(yy) * W + x) * spp + k];
            if (frozen_dirs == 1) {
                rep.dir_mode = dmode;
                rep.dirs = &dirs[(static_cast<size_t>(yy) * W + x) * spp + k];
            }
            if (frozen_dirs != 2) {
                rep.rr_mode = rmode;
                rep.rr = &rrp[(static_cast<size_t>(yy) * W + x) * spp + k];
            }
            V3<R> c = trace_path<R>(s, *cam, *prm, x, row0 + yy, prm->accumulation_count + k, pc[tid], rep);
            I = accumulate(I, c, k);
        }
        return V3<double>{static_cast<double>(I.x), static_cast<double>(I.y), static_cast<double>(I.z)};
    };
    std::vector<double> terms(static_cast<size_t>(npix), 0.0);  // per-pixel loss terms of the last call
    auto loss = [&](const std::vector<double> &pp, bool f32) {
        s.mat_override = pp.data();
        std::vector<double> part(H, 0.0);
        parallel_for(H, threads, [&](uint32_t yy, int tid) {
            const uint32_t y = row0 + yy;
            double acc = 0.0;
            for (uint32_t x = 0; x < W; ++x) {
                const float *t = target + 4 * (static_cast<size_t>(y) * W + x);
                terms[static_cast<size_t>(yy) * W + x] = 0.0;
                if (std::isnan(t[0])) continue;
                V3<double> I = f32 ? pixel(0.0f, yy, x, tid) : pixel(0.0, yy, x, tid);
                double dx = I.x - t[0], dy = I.y - t[1], dz = I.z - t[2];
                double e = dx * dx + dy * dy + dz * dz;
                terms[static_cast<size_t>(yy) * W + x] = e;
                acc += e;
            }
            part[yy] = acc;
        });
        s.mat_override = nullptr;
        double tot = 0.0;
        for (double v : part) tot += v;
        return tot / (3.0 * static_cast<double>(used));
    };
    // the unperturbed loss, which also records the frozen directions, in
    // the precision of the perturbed evaluations
    double l0 = loss(p, eval32 != 0);
    if (loss_out) *loss_out = l0;
    if (frozen_dirs == 1) dmode = DecisionTape::FREEZE_DIRS;
    if (frozen_dirs != 2) rmode = DecisionTape::FREEZE_RR;
    for (uint32_t i = 0; i < n_ids; ++i) {
        std::vector<double> pp = p, pm = p;
        pp[ids[i]] += h;
        pm[ids[i]] -= h;
        const double lp = loss(pp, eval32 != 0);
        std::vector<double> tp = terms;
        const double lm = loss(pm, eval32 != 0);
        grad_out[i] = (lp - lm) / (2.0 * h);
        if (pixel_grad)
            for (size_t k = 0; k < terms.size(); ++k)
                pixel_grad[i * terms.size() + k] = (tp[k] - terms[k]) / (2.0 * h * 3.0 * static_cast<double>(used));
    }
    return 0;
}

int oracle_grad_fd_ex(oracle_scene *sc, const pulse_camera *cam, const pulse_params *prm, const float *target,
                      const double *params, uint32_t n_materials, const uint32_t *ids, uint32_t n_ids, double h,
                      double *grad_out, double *loss_out, int threads, int frozen_dirs) {
    return oracle_grad_fd_rows(sc, cam, prm, target, params, n_materials, ids, n_ids, h, grad_out, loss_out, threads,
                               frozen_dirs, 0, cam->height, 0, nullptr);
}

int oracle_grad_fd(oracle_scene *sc, const pulse_camera *cam, const pulse_params *prm, const float *target,
                   const double *params, uint32_t n_materials, const uint32_t *ids, uint32_t n_ids, double h,
                   double *grad_out, double *loss_out, int threads) {
    return oracle_grad_fd_ex(sc, cam, prm, target, params, n_materials, ids, n_ids, h, grad_out, loss_out, threads, 0);
}

// ----- unit hooks for the sampler / RNG tests -----
uint32_t oracle_rand_u(uint32_t *state) { return rand_u(*state); }
float oracle_rand_f(uint32_t *state) { return rand_f(*state); }
uint32_t oracle_pixel_seed(uint32_t pixel_index, uint32_t acc) { return pixel_seed(pixel_index, acc); }

// importance_sample_ggx_d for one (n, wo, material, rng): out = wi[3], refl[3].
void oracle_sample_ggx_f32(const float *n, const float *wo, const pulse_material *m, uint32_t *rng, float *out) {
    DecisionTape tape;
    MaterialR<float> mat = load_material<float>(*m);
    SampleResult<float> r = importance_sample_ggx_d<float>(load3<float>(n), load3<float>(wo), mat, *rng, tape);
    out[0] = r.wi.x;
    out[1] = r.wi.y;
    out[2] = r.wi.z;
    out[3] = r.reflectance.x;
    out[4] = r.reflectance.y;
    out[5] = r.reflectance.z;
}
void oracle_sample_ggx_f64(const float *n, const float *wo, const pulse_material *m, uint32_t *rng, double *out) {
    DecisionTape tape;
    MaterialR<double> mat = load_material<double>(*m);
    SampleResult<double> r = importance_sample_ggx_d<double>(load3<double>(n), load3<double>(wo), mat, *rng, tape);
    out[0] = r.wi.x;
    out[1] = r.wi.y;
    out[2] = r.wi.z;
    out[3] = r.reflectance.x;
    out[4] = r.reflectance.y;
    out[5] = r.reflectance.z;
}
void oracle_sample_cosine_f32(const float *n, float e0, float e1, float *out) {
    V3<float> w = sample_cosine_hemisphere<float>(load3<float>(n), e0, e1);
    out[0] = w.x;
    out[1] = w.y;
    out[2] = w.z;
}
// GGX half-vector in the normal's frame for (e1, e2): out = wm[3].
void oracle_ggx_half_vector_f32(const float *n, float roughness, float e1, float e2, float *out) {
    DecisionTape tape;
    float a = roughness * roughness;
    float a2 = a * a;
    float theta = M<float>::acos(M<float>::sqrt((1.0f - e1) / (e1 * (a2 - 1.0f) + 1.0f)));
    float phi = e2 * kTWO_PI;
    ON<float> on = orthonormal_from_normal(load3<float>(n), tape);
    V3<float> wm = spherical_to_cartesian_in_on(theta, phi, on);
    out[0] = wm.x;
    out[1] = wm.y;
    out[2] = wm.z;
}
float oracle_d_ggx_f32(float NdotM, float roughness) { return D_GGX(NdotM, roughness); }
float oracle_g_smith_f32(float NdotO, float NdotI, float roughness) { return G_smith(NdotO, NdotI, roughness); }
void oracle_fresnel_f32(float c, const float *f0, float *out) {
    V3<float> f = fresnel_schlick(c, load3<float>(f0));
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
}
void oracle_sample_triangle_f32(float e0, float e1, const float *p0, const float *p1, const float *p2, float *out) {
    DecisionTape tape;
    V3<float> p = sample_triangle_uniformly<float>(e0, e1, load3<float>(p0), load3<float>(p1), load3<float>(p2), tape);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}
uint32_t oracle_cdf_pick(oracle_scene *sc, float e, uint32_t cdf_offset, uint32_t count, int linear) {
    const Scene &s = S(sc);
    return linear ? sample_light_triangle_area_cdf(s, e, cdf_offset, count)
                  : sample_light_triangle_area_cdf_bsearch(s, e, cdf_offset, count);
}
// The scan of utilities.wgsl:677-691 over a caller-provided CDF array.
uint32_t oracle_cdf_scan_array(const float *cdf, uint32_t count, float e, int linear) {
    Scene s;
    s.light_triangle_area_cdfs = cdf;
    return linear ? sample_light_triangle_area_cdf(s, e, 0, count) : sample_light_triangle_area_cdf_bsearch(s, e, 0, count);
}
// The emission-strength pick (utilities.wgsl:659-673) over a caller's CDF of
// `count` lights.
uint32_t oracle_strength_pick_array(const float *cdf, uint32_t count, float e) {
    Scene s;
    s.light_emission_strength_cdf = cdf;
    s.uniform.light_count = count;
    return sample_light_emission_strength_cdf(s, e);
}
// Lambert NEE at (p0, n0) with the scene's lights: out = direct[3].
void oracle_direct_light_f32(oracle_scene *sc, const float *p0, const float *n0, const float *base, uint32_t *rng,
                             float *out) {
    const Scene &s = S(sc);
    Counters c;
    DecisionTape tape;
    V3<float> d = sample_direct_light<float>(s, load3<float>(p0), load3<float>(n0), load3<float>(base), *rng, c, tape);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}
// n NEE samples at one receiver (p0, n0, material, wo), sample i drawn from
// rng state seeds[i]: ggx = 0 -> sample_direct_light (Lambert), 1 ->
// sample_direct_light_ggx.  out = n x direct[3].
void oracle_direct_light_batch_f32(oracle_scene *sc, const float *p0, const float *n0, const pulse_material *m,
                                   const float *wo, int ggx, const uint32_t *seeds, uint32_t n, float *out,
                                   int threads) {
    const Scene &s = S(sc);
    const MaterialR<float> mat = load_material<float>(*m);
    parallel_for(n, threads, [&](uint32_t i, int) {
        Counters c;
        DecisionTape tape;
        uint32_t rng = seeds[i];
        V3<float> d = ggx ? sample_direct_light_ggx<float>(s, load3<float>(p0), load3<float>(n0), mat, load3<float>(wo),
                                                           rng, c, tape)
                          : sample_direct_light<float>(s, load3<float>(p0), load3<float>(n0), mat.base, rng, c, tape);
        out[3 * i] = d.x;
        out[3 * i + 1] = d.y;
        out[3 * i + 2] = d.z;
    });
}
// importance_sample_ggx_d for n draws at one (n, wo, material), draw i from
// rng state seeds[i]: out = n x (wi[3], reflectance[3]).
void oracle_sample_ggx_batch_f32(const float *nrm, const float *wo, const pulse_material *m, const uint32_t *seeds,
                                 uint32_t n, float *out) {
    const MaterialR<float> mat = load_material<float>(*m);
    for (uint32_t i = 0; i < n; ++i) {
        DecisionTape tape;
        uint32_t rng = seeds[i];
        SampleResult<float> r = importance_sample_ggx_d<float>(load3<float>(nrm), load3<float>(wo), mat, rng, tape);
        float *o = out + 6 * i;
        o[0] = r.wi.x;
        o[1] = r.wi.y;
        o[2] = r.wi.z;
        o[3] = r.reflectance.x;
        o[4] = r.reflectance.y;
        o[5] = r.reflectance.z;
    }
}
// One Russian-roulette step (path_tracer.wgsl:104-108) on throughput t[3]:
// returns 1 if the path survives (t then divided by p), 0 if it stops.
int oracle_russian_roulette_f32(float *t, uint32_t *rng) {
    DecisionTape tape;
    V3<float> v = load3<float>(t);
    bool alive = russian_roulette<float>(v, *rng, tape);
    t[0] = v.x;
    t[1] = v.y;
    t[2] = v.z;
    return alive ? 1 : 0;
}
void oracle_pt_math(const float *x, uint32_t n, float *sin_out, float *cos_out, float *acos_out) {
    for (uint32_t i = 0; i < n; ++i) {
        sin_out[i] = pt_sinf(x[i]);
        cos_out[i] = pt_cosf(x[i]);
        acos_out[i] = pt_acosf(x[i]);
    }
}

}  // extern "C"
