/*
 * The CPU oracle of the render hot path: TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load liboracle.so, and only to check or time against it; the product
 * (libkodanuki_hip.so) never calls it.
 *
 * Two restatements live here:
 *   ref64.c  — the reference algorithm in FP64 as the reference computes it:
 *              tape point / interval / dual evaluation (source/sdf/program.cpp,
 *              source/sdf/numbers.cpp), the cell cut (plugins/geometry/tree/
 *              cut.cpp), the CSG field straight from the scene tree
 *              (source/sdf/scene.cpp:36-75), and the FP64 sphere trace
 *              (plugins/geometry/sphere_trace/sphere_trace.cpp).
 *   twin32.c — the same algorithms in the FP32 arithmetic the GPU path
 *              uses (intervals rounded outward exactly, transcendentals from
 *              double polynomials), operation for operation, so that octree
 *              cells, cut programs and hit leaves can be compared to the bit.
 *   path64.c — the plain spectral path tracer in FP64 (transport/paths).
 */
#ifndef KODANUKI_ORACLE_H
#define KODANUKI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "kodanuki.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FP64 reference semantics ------------------------------------------- */
/* A program with double constants: steps[i].value is ignored, values[i]
 * holds the double. */
double or_eval64(const kdn_step* steps, const double* values, uint32_t count, uint32_t result, const double p[3], double* regs);
/* Interval registers: lo[i], hi[i] over the box. */
void or_interval64(const kdn_step* steps, const double* values, uint32_t count, const double box[6], double* lo, double* hi);
/* Value and gradient at p. */
double or_dual64(const kdn_step* steps, const double* values, uint32_t count, uint32_t result, const double p[3], double g[3],
                 double* regs4);
/* The field straight from the scene tree: min over top nodes of the CSG
 * recursion with closed-form primitives (scene.cpp:36-75). */
double or_scene_distance(const kdn_leaf* leaves, const kdn_node* nodes, const uint32_t* children, const uint32_t* tops,
                         uint32_t top_count, const double p[3]);

/* FP64 sphere trace over every object program (sphereTraced). */
typedef struct or_scene64 {
    const kdn_step* steps;
    const double* values;
    const kdn_object* objects;
    uint32_t object_count;
    const uint32_t* leaf_material;
    double lower[3], size[3];
} or_scene64;
void or_sphere_trace64(const or_scene64* scene, const kdn_ray* rays, uint32_t count, kdn_hit* hits, double* distances64);

/* ---- FP32 twin ------------------------------------------------------------- */
/* Interval registers over the box, FP32 with outward rounding. */
void or_interval32(const kdn_step* steps, uint32_t count, const float box[6], float* lo, float* hi);
float or_eval32(const kdn_step* steps, uint32_t count, uint32_t result, const float p[3], float* regs);

/* Build the octree level by level; outputs laid out exactly as the GPU
 * library's kdn_octree_download. Call with null outputs to get the sizes. */
typedef struct or_octree {
    kdn_node_record* nodes;
    kdn_entry_record* entries;
    kdn_step* pool;
    uint64_t n_nodes, n_entries, n_pool;
    uint64_t cap_nodes, cap_entries, cap_pool;
} or_octree;
int or_build_octree32(const kdn_step* steps, const kdn_object* objects, uint32_t object_count, const float lower[3],
                      const float size[3], int depth, or_octree* out);
void or_free_octree(or_octree* tree);

/* March rays through an octree built by or_build_octree32 (FP32 twin of the
 * GPU extend kernel). */
void or_trace32(const or_octree* tree, const float lower[3], const float size[3], int depth, const uint32_t* leaf_material,
                const kdn_ray* rays, uint32_t count, kdn_hit* hits);

/* ---- the plain spectral path tracer (FP64) --------------------------------- */
typedef struct or_render_scene {
    or_scene64 geometry;
    const kdn_material* materials;
    uint32_t material_count;
    const kdn_light* lights;
    uint32_t light_count;
    kdn_camera camera;
    float sky_zenith[3], sky_horizon[3], sky_ground[3];
} or_render_scene;

/* Render pixels [x0,x1) x [y0,y1) of a width x height image, samples
 * [first, first + spp), into rgb sums (3 doubles per pixel of the window). */
void or_render64(const or_render_scene* scene, int width, int height, int x0, int y0, int x1, int y1, uint32_t first,
                 uint32_t spp, uint32_t seed, uint32_t surfaces, uint32_t min_depth, double* rgb, int threads);

/* The same with the Sobol member drawing an eye path's first 28
 * coordinates where sobol is set. */
void or_render64_sequenced(const or_render_scene* scene, int width, int height, int x0, int y0, int x1, int y1, uint32_t first,
                           uint32_t spp, uint32_t seed, uint32_t surfaces, uint32_t min_depth, int sobol, double* rgb, int threads);

/* Reference helpers checked on their own. */
double or_sobol_number(uint64_t stream, uint64_t index, uint64_t coordinate);
uint64_t or_mixed_bits(uint64_t v);
double or_hashed_number(uint64_t stream, uint64_t index, uint64_t coordinate);
void or_sampled_wavelengths(double u, double nanometres[4], double density[4]);
void or_colour_of(const double nanometres[4], const double density[4], const double spectrum[4], double rgb[3]);
void or_spectrum_of(const double rgb[3], const double nanometres[4], double out[4]);
/* The GGX escape table of a width (VIEWS = 32 entries) and its average. */
void or_escape_of(double width, double table[32], double* average);

#ifdef __cplusplus
}
#endif

#endif
