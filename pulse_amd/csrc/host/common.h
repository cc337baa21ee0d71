// Shared host-side helpers of libpulse_hip: error reporting and layout checks.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "pulse_hip.h"
#include "pulse_scene.h"

static_assert(sizeof(pulse_scene_uniform) == 8, "SceneUniform is 8 B (types.wgsl:3-6)");
static_assert(sizeof(pulse_primitive) == 48, "Primitive is 48 B (types.wgsl:23-27)");
static_assert(sizeof(pulse_triangle_data) == 80, "TriangleData is 80 B (types.wgsl:29-36)");
static_assert(sizeof(pulse_blas_node) == 32, "BLASNode is 32 B (types.wgsl:38-43)");
static_assert(sizeof(pulse_tlas_node) == 32, "TLASNode is 32 B (types.wgsl:45-50)");
static_assert(sizeof(pulse_mesh_instance) == 160, "MeshInstance is 160 B (types.wgsl:52-60)");
static_assert(sizeof(pulse_material) == 48, "Material is 48 B (types.wgsl:62-68)");
static_assert(sizeof(pulse_light_data_index) == 8, "LightDataIndex is 8 B (types.wgsl:70-74)");
static_assert(sizeof(pulse_ray) == 32, "pulse_ray is 32 B");
static_assert(sizeof(pulse_hit) == 20, "pulse_hit is 20 B");

namespace pulse {

// Thread-local message behind pulse_last_error().
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *last_error();

}  // namespace pulse

#define PULSE_FAIL(code, ...)               \
    do {                                    \
        ::pulse::set_error(__VA_ARGS__);    \
        return (code);                      \
    } while (0)
