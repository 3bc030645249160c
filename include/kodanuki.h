/*
 * kodanuki.h — the C ABI of the MI355X render hot path.
 *
 * One shared library, libkodanuki_hip.so, exports every function below.
 * Signatures hold plain C types only: pointers, sizes, PODs. Every call
 * returns 0 on success or a nonzero KDN_E* code; the message of the last
 * failure is read back with kdn_last_error().
 *
 * Host-owned buffers are passed as plain pointers. Entry points whose
 * names end in _device take device pointers (HBM) and a HIP stream handle
 * (NULL = the null stream) so that a caller that keeps its inputs resident
 * hands them over without a copy.
 *
 * The reference is a single process C++26 renderer whose hot path is
 * reached through the plugin hooks of source/core/plugin.h. Each entry
 * point names the reference interface it stands in for.
 */
#ifndef KODANUKI_H
#define KODANUKI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define KDN_OK 0
#define KDN_E_ARGUMENT 1   /* a bad argument or a size beyond a stated limit */
#define KDN_E_SCENE 2      /* the .sdf text does not load */
#define KDN_E_DEVICE 3     /* a HIP call failed */
#define KDN_E_STATE 4      /* a call out of order (no scene, no octree) */
#define KDN_E_MEMORY 5     /* a device allocation failed */

/* ---- the flat instruction tape ------------------------------------------ */
/*
 * One step of a program: the reference Step (source/sdf/program.h:56-63)
 * with u16 register indices (the reference's own per-cell limit, KeptSteps
 * entries are u16, program.h:104-128) and an FP32 value. 16 bytes.
 *
 * flags, for a Constant: 0 when value is the scene's double exactly,
 * 1 when the double lies just above value (its enclosure is
 * [value, nextup(value)]), 2 when it lies just below.
 */
typedef struct kdn_step {
    uint8_t op;
    uint8_t flags;
    uint16_t a, b, c;
    float value;
    uint32_t leaf;
} kdn_step;

/* Opcodes, numbered as the reference Opcode enum (program.h:20-54). */
enum {
    KDN_CONSTANT = 0, KDN_ARGUMENT, KDN_POINT_X, KDN_POINT_Y, KDN_POINT_Z,
    KDN_NORMAL_X, KDN_NORMAL_Y, KDN_NORMAL_Z, KDN_ADD, KDN_SUBTRACT,
    KDN_MULTIPLY, KDN_DIVIDE, KDN_NEGATE, KDN_SQUARE_ROOT, KDN_ABSOLUTE,
    KDN_FLOOR, KDN_SINE, KDN_COSINE, KDN_SIGN, KDN_ARC_TANGENT, KDN_POWER,
    KDN_MINIMUM, KDN_MAXIMUM, KDN_SMOOTH_MINIMUM, KDN_POSITIVE_SQUARE,
    KDN_LESS, KDN_LESS_EQUAL, KDN_EQUAL, KDN_NOT_EQUAL, KDN_AND, KDN_OR,
    KDN_NOT, KDN_SELECT, KDN_OPCODE_COUNT
};

/* An object program inside a tape: steps [first, first + count), result
 * relative to first. */
typedef struct kdn_object {
    uint32_t first;
    uint32_t count;
    uint32_t result;
    uint32_t material;   /* the material of its first leaf (informational) */
} kdn_object;

#define KDN_NO_LEAF 0xFFFFFFFFu
#define KDN_NO_MATERIAL 0xFFFFFFFFu

/* ---- rays and hits -------------------------------------------------------- */
/* The reference Ray (source/core/batch.h:125-155), 36 bytes. */
typedef struct kdn_ray {
    float origin[3];
    float direction[3];
    float far_;
    uint32_t inside;
    uint32_t within;
} kdn_ray;

#define KDN_HIT_MISSED 0
#define KDN_HIT_SURFACE 1
#define KDN_HIT_SOLID 2

/* The reference Hit (source/core/batch.h:157-205) less its tree-only
 * fields, 48 bytes. */
typedef struct kdn_hit {
    uint32_t kind;
    float distance;
    float point[3];
    float normal[3];
    uint32_t leaf;
    uint32_t other;
    float share;
    uint32_t cell;
} kdn_hit;

/* ---- the loaded scene (host) -------------------------------------------- */
typedef struct kdn_scene kdn_scene;

#define KDN_MATTE 0
#define KDN_PLASTIC 1
#define KDN_METAL 2
#define KDN_GLASS 3

/* A material as the scene says it (source/sdf/scene.h:337-402, the
 * fields the render hot path reads). */
typedef struct kdn_material {
    uint32_t kind;
    float albedo[3];
    float roughness;
    float refraction;
    float dispersion;
    float glow[3];
    float strength;
} kdn_material;

/* Light kinds (source/sdf/scene.h:155-181). A moon is aimed at and met as
 * a sun is (light.cpp treats the two alike). */
#define KDN_LIGHT_POINT 0
#define KDN_LIGHT_QUAD 1
#define KDN_LIGHT_DIRECTIONAL 2
#define KDN_LIGHT_SPHERE 3
#define KDN_LIGHT_SUN 4
#define KDN_LIGHT_MOON 5

/* A lamp (source/sdf/scene.h:481-528, quad and point kinds). */
typedef struct kdn_light {
    uint32_t kind;
    float at[3];
    float along[3];
    float across[3];
    float power[3];
    float direction[3];   /* toward a sun or a directional light */
    float radius;         /* a sphere's */
    float angle;          /* a sun's full angle across, radians */
} kdn_light;

/* A leaf as placed (source/sdf/scene.h:237-288): its primitive, the
 * primitive's numbers, the world→local map and its material. */
#define KDN_PRIM_SPHERE 0
#define KDN_PRIM_BOX 1
#define KDN_PRIM_TORUS 2
#define KDN_PRIM_PLANE 3
#define KDN_PRIM_CYLINDER 4
#define KDN_PRIM_CAPSULE 5
#define KDN_PRIM_ELLIPSOID 6
typedef struct kdn_leaf {
    uint32_t primitive;
    uint32_t material;
    double args[8];
    double linear[9];
    double origin[3];
    double stretch;
} kdn_leaf;

/* A CSG node (source/sdf/scene.h:290-310): join 0 Leaf, 1 Merge, 2 Cut,
 * 3 Clip; its children are the child_count entries of the children list
 * from first_child on, as kdn_scene_nodes returns them. */
typedef struct kdn_node {
    uint32_t join;
    double radius;
    uint32_t leaf;
    uint32_t first_child;
    uint32_t child_count;
} kdn_node;

typedef struct kdn_camera {
    float eye[3];
    float target[3];
    float up[3];
    float fov;
} kdn_camera;

typedef struct kdn_render_settings {
    int32_t width, height, seed, samples, surfaces;
    float far_;
    float sky_zenith[3];
    float sky_horizon[3];
    float sky_ground[3];
    int32_t batch;      /* render { batch }: samples a batch of a pixel's estimate holds (0: unset) */
    float exposure;     /* grade { exposure } (1) */
    int32_t curve;      /* grade { curve }: KDN_CURVE_FILMIC (the default) or KDN_CURVE_CLAMPED */
    float tolerance;    /* render { tolerance }: the noise a pixel may keep, in steps of a byte (0: unset) */
    int32_t split;      /* render { split }: samples that share a first vertex (0: unset) */
} kdn_render_settings;

#define KDN_CURVE_FILMIC 0
#define KDN_CURVE_CLAMPED 1

typedef struct kdn_scene_info {
    uint32_t steps;       /* steps over every object program */
    uint32_t objects;
    uint32_t leaves;
    uint32_t nodes;
    uint32_t top_nodes;   /* object roots */
    uint32_t children;    /* entries of the children list */
    uint32_t materials;
    uint32_t lights;
} kdn_scene_info;

/*
 * Load .sdf text and compile its object programs: the reference
 * loadScene (source/syntax/evaluator.h:369) followed by objectPrograms
 * (source/sdf/scene.cpp:290-296, scene.h:897 union split). base_dir is where
 * `import` paths are read from (NULL: only the built-in standard.sdf).
 * On failure err holds "line:column: message".
 */
int kdn_scene_load(const char* text, const char* base_dir, kdn_scene** out, char* err, size_t err_size);
void kdn_scene_free(kdn_scene* scene);
int kdn_scene_describe(const kdn_scene* scene, kdn_scene_info* info);
/* Copy out the compiled tape: steps[info.steps], objects[info.objects]. */
int kdn_scene_tape(const kdn_scene* scene, kdn_step* steps, kdn_object* objects);
/* The same tape with its constants in double, for FP64 checks. */
int kdn_scene_tape_values(const kdn_scene* scene, double* values);
int kdn_scene_leaves(const kdn_scene* scene, kdn_leaf* leaves);
/* One leaf alone over world space: leafProgram (source/sdf/scene.cpp:
 * 284-288), the program the survey's chains walk. Writes at most cap steps
 * (their constants in double in values) and the step count and result
 * register; KDN_E_ARGUMENT for a leaf out of range or cap too small. */
int kdn_scene_leaf_program(const kdn_scene* scene, uint32_t leaf, kdn_step* steps, double* values, uint32_t cap, uint32_t* count,
                           uint32_t* result);
int kdn_scene_nodes(const kdn_scene* scene, kdn_node* nodes, uint32_t* top_nodes, uint32_t* children);
int kdn_scene_materials(const kdn_scene* scene, kdn_material* materials);
int kdn_scene_lights(const kdn_scene* scene, kdn_light* lights);
int kdn_scene_settings(const kdn_scene* scene, kdn_camera* camera, kdn_render_settings* render);
/* The box the octree is built over: lower[3] and size[3], each size a
 * power of two and each corner a multiple of the finest cell, so that
 * every cell box is exact in FP32. It stands in for sceneRegion
 * (source/plugins/geometry/tree/build.cpp:405, an interval search for the
 * occupied bounds inside reachRegion, sphere_trace.cpp:39-46). */
int kdn_scene_region(const kdn_scene* scene, int depth_cap, float* lower3, float* size3);

/* ---- the device context -------------------------------------------------- */
typedef struct kdn_ctx kdn_ctx;

/* Open a context on a HIP device: stands in for Backend openWorkers
 * (source/core/plugin.h:333-357). */
int kdn_open(int device, kdn_ctx** out);
void kdn_close(kdn_ctx* ctx);
/* Copy the message of the last failure (NUL-terminated, truncated). */
int kdn_last_error(const kdn_ctx* ctx, char* buf, size_t size);
/* Last failure of a call that has no context (kdn_open, scene calls). */
const char* kdn_global_error(void);

/* Put the scene's tape, materials and lights in HBM: the Backend
 * uploadBlock hook (source/core/plugin.h:356, core/block.h:21-49). */
int kdn_upload_scene(kdn_ctx* ctx, const kdn_scene* scene);

/* ---- the interval-cut octree ---------------------------------------------- */
/*
 * Build the octree level by level on the GPU: per cell, interval
 * evaluation of every kept program over the cell box (sdf/numbers.cpp
 * Interval ops, FP32 with outward ulp padding), the cut rewrite and prune
 * (plugins/geometry/tree/cut.cpp:117-348), the cell decision cutCell
 * (cut.cpp:446-476) and compaction of surviving cells. Stands in for
 * Geometry buildWorld → buildTree (core/plugin.h:168, tree/build.cpp:985)
 * with the split rule "Surface and level < depth_cap". The depth cap is
 * 0..11 (the march keeps a ray's ancestry of 12 levels in LDS).
 */
typedef struct kdn_octree_stats {
    uint32_t levels;                  /* depth_cap + 1 */
    uint32_t nodes[16];               /* nodes per level */
    uint32_t surface[16];             /* Surface nodes per level */
    uint32_t filled[16];
    uint32_t entries;                 /* (object, program) entries over all nodes */
    uint64_t pool_steps;              /* steps in the program pool */
    double build_ms;
} kdn_octree_stats;

int kdn_build_octree(kdn_ctx* ctx, const float* lower3, const float* size3, int depth_cap, kdn_octree_stats* stats);

/* One octree node as downloaded: level, integer coordinates at its level,
 * kind (0 Empty, 1 Filled, 2 Surface leaf, 3 Branch), first child index in
 * the next level (Branch), and its (object, program) entries. */
typedef struct kdn_node_record {
    uint32_t level;
    uint32_t x, y, z;
    uint32_t kind;
    uint32_t first_child;
    uint32_t entry_first;
    uint32_t entry_count;
} kdn_node_record;

typedef struct kdn_entry_record {
    uint32_t object;
    uint32_t program;     /* offset of the program header in the pool */
} kdn_entry_record;

/* Sizes of the built octree: nodes, entries, pool words. */
int kdn_octree_sizes(kdn_ctx* ctx, uint64_t* nodes, uint64_t* entries, uint64_t* pool_steps);
/* Download the whole octree. The pool holds each program as a header step
 * (op = 255, a|b<<16 = length, c = result) followed by its steps. */
int kdn_octree_download(kdn_ctx* ctx, kdn_node_record* nodes, kdn_entry_record* entries, kdn_step* pool);

/* ---- extend: rays → hits -------------------------------------------------- */
/*
 * Trace rays through the octree: sphere tracing over each cell's cut
 * programs, FP32, with hit leaf by the reference source walk. Stands in
 * for Execution extendRays (core/plugin.h:183,
 * plugins/execution/single/plugin.cpp:27-41) over Geometry
 * sphere_trace (plugins/geometry/sphere_trace/sphere_trace.cpp:216-228).
 */
int kdn_extend(kdn_ctx* ctx, const kdn_ray* rays, uint32_t count, kdn_hit* hits);
int kdn_extend_device(kdn_ctx* ctx, const void* rays, uint32_t count, void* hits, void* stream);

/* ---- the wavefront integrator --------------------------------------------- */
/*
 * One pixel pass: `spp` samples of every pixel, from sample index
 * `first_sample`, through the wavefront integrator (ray generation,
 * extend, shading, NEE shadow rays, MIS, hero-wavelength throughput,
 * roulette), folded into FP64 accumulators. On the device the pass
 * refuses scenes painted with glass (KDN_E_STATE): glass scattering is
 * held back there until a device fault of its renders is explained. Stands in for the
 * Backend hook the reference lacks (a whole pixel pass per call) over the
 * Transport steps extendRays → arriveAtHits → connectLights →
 * scatterWalks → fold (core/worker.cpp:31-39,
 * plugins/transport/paths/plugin.cpp:107-155).
 */
typedef struct kdn_pass_desc {
    int32_t width, height;
    uint32_t first_sample;
    uint32_t spp;
    uint32_t seed;
    uint32_t surfaces;      /* the most surfaces a path stands on */
    uint32_t min_depth;     /* roulette starts at this depth (reference 4) */
    uint32_t flags;         /* bit 0: keep the accumulators (progressive);
                               bit 1: time the kernels with HIP events;
                               bit 2: count the march's FP32 operations
                                      (an instrumented, slower pass);
                               bit 3: no finishing launch (every bounce a
                                      wavefront launch to the end);
                               bit 4: the learned guide: sample directions
                                      with it (one-sample MIS with the
                                      BSDF) once it has trained, and train
                                      it on the pass's paths;
                               bit 5: the plain guide: the first vertex
                                      of each path samples with its
                                      pixel's guide (kdn_highlight_guides
                                      made for this image; not with bit 4);
                               bit 6: keep every pixel's estimate as the
                                      reference's frame does (batch means
                                      and squares, kdn_read_estimates),
                                      batches of `batch` samples;
                               bit 7: the Sobol member: an eye path's first
                                      28 coordinates off the Owen-scrambled
                                      Sobol sequence (sampler/sobol), the
                                      rest hashed as without it;
                               bit 8: the count pass of adaptive stopping:
                                      each pixel draws its allotment less its
                                      trial (kdn_stopping_allot), samples
                                      first_sample.. (spp unused); with bits
                                      0 and 6 the estimates restart from the
                                      trial's totals (sumStart);
                               bit 9: keep the frame record's surfaces:
                                      each pixel's first surface of its last
                                      sample (first_sample + spp - 1), for
                                      kdn_temporal_frame (not with bit 8) */
    uint32_t batch;         /* samples a batch holds (bit 6; 0: no batch is ever closed) */
    uint32_t round_paths;   /* the most paths of one round (0: 2^23; at least 4096) */
} kdn_pass_desc;

typedef struct kdn_pass_stats {
    double ms;
    uint64_t segments;       /* extend rays traced */
    uint64_t shadow_rays;
    uint32_t bounces;        /* wavefront iterations */
    uint32_t trace_launches; /* extend launches (shadow launches apart) */
    double trace_ms;         /* extend kernels, HIP events (flags bit 1) */
    double shadow_ms;        /* shadow-ray kernels */
    double shade_ms;         /* shading kernels */
    double other_ms;         /* generate and fold */
    double trace_flops;      /* FP32 operations of the extend march (bit 2) */
    double shadow_flops;
    double finish_ms;        /* the finishing launch (paths to their ends) */
    double finish_flops;     /* its march operations (bit 2) */
    uint32_t shadow_launches;
    uint32_t finish_launches;
    double guide_ms;         /* the guide's inference and training kernels */
    double guide_infer_ms;   /* its inference launches alone */
    double guide_loss;       /* its last training step's loss */
    uint64_t guide_queries;  /* vertices its inference launches took */
    uint32_t guide_samples;  /* training samples the pass gave */
    uint32_t guide_steps;    /* Adam steps taken */
} kdn_pass_stats;

int kdn_render_pass(kdn_ctx* ctx, const kdn_pass_desc* desc, kdn_pass_stats* stats);

/* Adaptive stopping: the reference's stopping/confidence_interval member
 * (confidence_interval.h:36-106, its Hooks trialSamples .. allottedSplit,
 * called by core/frame.cpp:28-62). A frame draws a trial of two batches in
 * every pixel (a pass with flags bit 6), judges each pixel's trial against
 * the tolerance, allots each pixel a count from its eight neighbours'
 * judgements (never its own), and draws the rest (a pass with flags
 * bits 0 and 8). Every sample here draws its own first vertex: split 1. */
typedef struct kdn_stopping {
    int32_t samples;    /* the most samples a pixel draws in a frame */
    int32_t batch;      /* samples a batch holds */
    double tolerance;   /* half the interval's width, in steps of a byte after the grade */
    float exposure;     /* the grade the noise is judged through */
    int32_t curve;      /* KDN_CURVE_FILMIC or KDN_CURVE_CLAMPED */
    int32_t split;      /* samples that share a first vertex: 1 */
} kdn_stopping;
/* effortFor (core/stopping.cpp:23-34): the ceiling and batch a tolerance
 * asks for where the scene names none. */
int kdn_stopping_effort(double tolerance, int32_t* samples, int32_t* batch);
/* confidenceTrial (confidence_interval.cpp:26-29): samples every pixel's
 * trial draws, 0 for none (then the frame is one pass of `samples`). */
int kdn_stopping_trial(const kdn_stopping* stopping, int32_t* samples);
/* Judge every pixel's trial (the estimates a pass with flags bit 6 kept:
 * confidenceWanted and splitSpread) and allot each its count
 * (confidenceAllotted, splitFor) on the device; the counts stay on the
 * device for the next pass with flags bit 8, and *total is that pass's
 * paths (the samples every pixel draws beyond its trial). */
int kdn_stopping_allot(kdn_ctx* ctx, const kdn_stopping* stopping, int32_t trial, uint64_t* total);
/* Read the last allotment: per pixel (by rows) the samples its own trial
 * wants (infinite where no count would do), the spread of its shared part
 * and of the rest (2 floats), and its allotted count; any may be null. */
int kdn_read_stopping(kdn_ctx* ctx, float* wanted, float* spreads, int32_t* allotted);

/* An eye path's numbers as the integrator draws them (numberAt,
 * plugins/transport/paths/sampler.cpp:15-30): for each sample index, the
 * number at `coordinate` of the stream, off the Sobol sequence
 * (sobolNumber, plugins/sampler/sobol/sobol.h:36) where `sobol` is set and
 * the coordinate is below 28, hashed (math/random.cpp hashedNumber)
 * otherwise; computed on the device. At most 2^24 indices. */
int kdn_sequence_numbers(kdn_ctx* ctx, uint64_t stream, const uint32_t* indices, uint32_t count, uint32_t coordinate, int32_t sobol,
                         double* out);
/* Read the FP64 accumulators: per pixel the sum of linear RGB (3 doubles),
 * the sum of its squares per channel (3 doubles) and the sample count. */
int kdn_read_accumulation(kdn_ctx* ctx, double* rgb_sum, double* sq_sum, uint32_t* count);
/* The mean image as linear RGB floats (width*height*3). */
int kdn_read_image(kdn_ctx* ctx, float* rgb);

/* A pixel's estimate as the reference's frame keeps it (Estimate,
 * core/stopping.h:57-100): its samples' colours summed in sample order and
 * counted, the colours of the batch being drawn, the count of whole
 * batches with the sum of their means and of their means' squares, and the
 * shared parts (a sample drawn alone shares its whole colour: its shared
 * part is its channel mean; foldedSample and closedBatch,
 * core/stopping.cpp:36-67). */
typedef struct kdn_estimate {
    double total[3];
    int32_t taken;
    int32_t batches;
    double batch[3];
    double means[3];
    double squares[3];
    double shared;
    double shared_means;
    double shared_squares;
    double own_squares;
} kdn_estimate;
/* The estimates the passes with flags bit 6 kept (width*height, by rows
 * from the top left); zeroed with the accumulators. */
int kdn_read_estimates(kdn_ctx* ctx, kdn_estimate* estimates);
/* Fold sample colours into estimates on the device, as a pass does:
 * colours[(pixel * spp + sample) * 3 + channel], each pixel's samples in
 * order, into estimates[pixel] (host buffers, read and written back), with
 * batches of `batch` samples (the fold of the reference's foldedPixel,
 * core/frame.cpp:221-235). */
int kdn_fold_samples(kdn_ctx* ctx, const double* colours, uint32_t pixels, uint32_t spp, uint32_t batch, kdn_estimate* estimates);

/* ---- images (host) ----------------------------------------------------------- */
/* A linear RGB image (width*height*3 floats by rows from the top left)
 * graded and written as a binary PPM (writePpm, shared/image.cpp:72-79):
 * each value times the exposure, through the filmic curve (or clamped),
 * to the power 1/2.2, rounded to a byte. A file that cannot be written is
 * KDN_E_ARGUMENT with "path:1:1: cannot write" (kdn_global_error). */
int kdn_image_write_ppm(const float* rgb, int32_t width, int32_t height, float exposure, int32_t curve, const char* path);
/* The same bytes into memory (width*height*3), no file. */
int kdn_image_encode(const float* rgb, int32_t width, int32_t height, float exposure, int32_t curve, uint8_t* bytes);
/* The linear values as a little-endian PFM, bottom row first (writePfm,
 * shared/image.cpp:88-96). */
int kdn_image_write_pfm(const float* rgb, int32_t width, int32_t height, const char* path);

/* ---- the denoiser ------------------------------------------------------------ */
/*
 * An edge-avoiding à-trous filter guided by the variance (SVGF-like), the
 * reference's denoisedFrame (plugins/reconstruction/atrous/atrous.h:462,
 * atrous.cpp:18-500) on the device, to its bits: the light divided by the
 * albedo, a variance per pixel (from its history, or from the 7x7 pixels
 * around it for histories under 4 samples), 4 passes of a 5x5 B3 spline
 * at steps 1, 2, 4, 8 with normal, plane and luminance edge stops, the
 * albedo multiplied back and the filtered light faded into the held light
 * by 1 / (1 + length / 8).
 */
/* What a camera ray met first (Surface, core/batch.h:244-279): distance
 * along the ray (infinite on a miss), unit normal facing the ray, leaf
 * (KDN_NO_LEAF on a miss) and albedo (glass white, zero on a miss). */
typedef struct kdn_surface {
    float distance;
    float normal[3];
    uint32_t leaf;
    float albedo[3];
} kdn_surface;

/* A lens (Lens, core/camera.h:14-56, pinhole): the eye, its unit frame,
 * tan(field of view / 2) and width / height. */
typedef struct kdn_lens {
    double eye[3], forward[3], right[3], up[3];
    double reach, aspect;
} kdn_lens;

/* A pixel of the temporal history (TemporalPixel, temporal/temporal.h:18-70). */
typedef struct kdn_denoise_pixel {
    double colour[3];    /* its light (its lighting where demodulated) */
    double moment;       /* the blend of the luminance squared */
    float length;        /* samples the blend holds */
    float distance;      /* from the lens; infinite on the sky */
    double normal[3];    /* unit shading normal; zero on the sky */
    uint32_t leaf;       /* KDN_NO_LEAF on the sky */
    uint32_t pad;
    double albedo[3];
} kdn_denoise_pixel;

/* The lens at the history's size (DenoiseCamera, atrous.h:386-404). */
typedef struct kdn_denoise_frame {
    int32_t width, height;
    double forward[3], right[3], up[3];  /* unit, the lens's frame */
    double reach;                        /* tan(field of view / 2) */
    double aspect;                       /* width / height */
    double centre_x, centre_y;           /* where in its pixel a ray passes (0.5) */
    uint32_t demodulated;                /* the history holds lighting */
    uint32_t pad;
} kdn_denoise_frame;

/* Denoise a frame: pixels[width * height] by rows from the top left ->
 * rgb[width * height * 3] (host buffers). */
int kdn_denoise(kdn_ctx* ctx, const kdn_denoise_frame* frame, const kdn_denoise_pixel* pixels, float* rgb);
/* The same on device buffers, on a stream (NULL: the context's). */
int kdn_denoise_device(kdn_ctx* ctx, const kdn_denoise_frame* frame, const void* pixels, void* rgb, void* stream);
/* kdn_denoise_device that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture or a serving loop's inner steps. The filter's
 * working memory is a static array of every GPU the library runs on (76 MiB,
 * held until the process ends), so frames are limited to 2^20 pixels
 * (KDN_E_ARGUMENT beyond); pixels and rgb are the caller's device buffers.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads) race and must be ordered by the
 * caller. Results are the same bits as kdn_denoise_device's. */
int kdn_denoise_device_static(kdn_ctx* ctx, const kdn_denoise_frame* frame, const void* pixels, void* rgb, void* stream);
typedef struct kdn_denoise_stats {
    double gbuffer_ms;   /* primary rays and the history, HIP events */
    double filter_ms;    /* the filter's kernels */
} kdn_denoise_stats;
/* Denoise the context's accumulated image (kdn_render_pass): its FP64
 * sums are the history (mean, moment, samples), a pass of primary rays
 * through the pixel centres gives distance, normal, leaf and albedo.
 * stats may be NULL. */
int kdn_denoise_accumulation(kdn_ctx* ctx, float* rgb, kdn_denoise_stats* stats);

/* ---- the temporal member ---------------------------------------------------- */
/*
 * temporalFrame (reconstruction/temporal/temporal.h:237) on the device, for
 * a viewer's frames drawn with a G-buffer at factor 1 (every pixel drawn
 * each frame through its centre; the history demodulated): each pixel's
 * history reprojected through the G-buffer into the camera before and read
 * from the pixels on its plane, blended with the frame's sample where that
 * sample's first surface is the G-buffer's leaf, clamped to its 3x3
 * neighbourhood; the shown light is the lighting times the G-buffer's
 * albedo. The history stays resident in HBM between frames. The upscaling
 * mode (shown pixels finer than the drawn ones) is not built.
 */
typedef struct kdn_temporal_desc {
    uint32_t moved;      /* the camera moved since the frame before (ShownFrame.moved) */
    uint32_t restarted;  /* drop every history (ShownFrame.restarted) */
    uint32_t recounted;  /* the scene's leaves were renumbered: leaves match as sky or not (ShownFrame.recounted) */
    uint32_t denoise;    /* run the a-trous passes on the history after it (kdn_denoise_device) */
} kdn_temporal_desc;
typedef struct kdn_temporal_stats {
    double gbuffer_ms;   /* the G-buffer's primary rays */
    double temporal_ms;  /* the temporal kernel */
    double filter_ms;    /* the a-trous passes (desc.denoise) */
} kdn_temporal_stats;
/* Set the resident scene's camera for the passes and frames after it (a
 * viewer's orbit). */
int kdn_set_camera(kdn_ctx* ctx, const kdn_camera* camera);
/* One viewer frame after a fresh pass (kdn_render_pass with flags bit 9,
 * not keeping the accumulators): the G-buffer of the scene's camera, the
 * frame's lights (the pass's means) and record, and the resident history
 * -> the next history, and the shown (and, with desc.denoise, the
 * denoised) image, width * height * 3 floats each (host; either may be
 * NULL). stats may be NULL. */
int kdn_temporal_frame(kdn_ctx* ctx, const kdn_temporal_desc* desc, float* shown, float* denoised, kdn_temporal_stats* stats);
/* Read the resident history (width * height pixels by rows). */
int kdn_temporal_history(kdn_ctx* ctx, kdn_denoise_pixel* pixels);
/* The temporal step on given frames (host buffers): this frame's lens and
 * the one before, its lights (width * height * 3 doubles), record and
 * G-buffer surfaces, and the history before (NULL, or before_width x
 * before_height pixels) -> the history after and the shown image. The
 * resident history is neither read nor changed. */
int kdn_temporal_step(kdn_ctx* ctx, int32_t width, int32_t height, const kdn_lens* lens, const kdn_lens* before,
                      const kdn_temporal_desc* desc, const double* parts, const kdn_surface* record, const kdn_surface* gbuffer,
                      const kdn_denoise_pixel* before_pixels, int32_t before_width, int32_t before_height,
                      kdn_denoise_pixel* pixels, float* shown);

/* ---- the learned guide ------------------------------------------------------ */
/*
 * A network that learns, per scene, where light arrives from: a hash grid
 * of the position (16 levels of 2^19 entries, 2 features each, 16 to 128
 * cells a side by default), real spherical harmonics of the outgoing
 * direction, the normal and a code of the hero wavelength feed three ReLU
 * layers of 64 (BF16 weights, MFMA with FP32 accumulation) whose outputs
 * decode to a mixture of 8 von Mises-Fisher lobes and a selection
 * probability. A guided pass (kdn_render_pass flag bit 4) samples a
 * direction at each of the first four vertices of a path from the mixture
 * with that probability and from the BSDF otherwise (one-sample MIS, so
 * the estimate stays unbiased), and trains the network online, with Adam,
 * on the radiance its own paths bring back. The reference has no learned
 * guide; it joins the direction sampling the reference's Sampler
 * dispatches (plugins/transport/paths/sampler.h:23,357). The calls below
 * expose the
 * network for tests and tools.
 */
typedef struct kdn_guide_query {
    float point[3];    /* world position */
    float out[3];      /* unit direction toward the viewer */
    float normal[3];   /* unit normal facing out */
    float nm;          /* hero wavelength */
} kdn_guide_query;

typedef struct kdn_guide_sample {
    float point[3];
    float out[3];
    float normal[3];
    float nm;
    float in[3];       /* the direction light came back along */
    float weight;      /* radiance over the density it was sampled with */
    float bsdf_pdf;    /* the BSDF's density of in: the loss is then on the
                          one-sample MIS density a q + (1 - a) bsdf_pdf, which
                          trains the selection a too; negative: on q alone */
    float pad;
} kdn_guide_sample;

typedef struct kdn_guide_lobe {
    float axis[3];
    float sharpness;
    float share;
} kdn_guide_lobe;

typedef struct kdn_guide_info {
    uint32_t levels, features, table_size, width, layers, lobes, depths;
    uint32_t train_paths;    /* paths of a round that keep training records, at most */
    uint32_t resolution[16];
    uint32_t dense_levels;   /* bit l: level l indexed densely */
    uint32_t steps;          /* Adam steps taken */
    float loss_scale;
    float share;             /* the guide's share of directions */
    double last_loss;
    uint32_t last_samples;
    uint32_t pad;
    uint64_t grid_params, mlp_params;
    float lower[3], size[3]; /* the region positions are scaled from */
} kdn_guide_info;

/* How a guided pass uses and trains the guide. */
typedef struct kdn_guide_config {
    float learning_rate;     /* Adam (default 1e-2) */
    float share;             /* the guide's share of the sampled directions (0.3);
                                with learn_select, the most of them it may take */
    float epochs;            /* passes over a round's samples (2) */
    uint32_t min_steps;      /* Adam steps a round, at least (16) */
    uint32_t max_steps;      /* and at most (64, the limit) */
    uint32_t max_resolution; /* the finest grid level's cells per side (128) */
    uint32_t train_paths;    /* paths of a round that keep training records, at most (2^20) */
    uint32_t learn_select;   /* 1: the selection probability is share times the
                                sigmoid of a learned output; 0: share itself (0) */
    uint32_t product;        /* 1: the lobes learn the integrand f cos L;
                                0: the incident radiance L (1) */
    uint32_t pad;
} kdn_guide_config;

/* Initialise the network from a seed over a region (NULL: the octree's);
 * the configuration returns to its defaults. A new max_resolution
 * re-spaces the grid's levels (configure before training). */
int kdn_guide_reset(kdn_ctx* ctx, uint32_t seed, const float* lower3, const float* size3);
int kdn_guide_configure(kdn_ctx* ctx, const kdn_guide_config* config);
int kdn_guide_describe(kdn_ctx* ctx, kdn_guide_info* info);
/* FP32 parameters: grid[grid_params] (level, entry, feature) and
 * mlp[mlp_params] (layer, input, output); NULL skips one. */
int kdn_guide_params(kdn_ctx* ctx, float* grid, float* mlp);
int kdn_guide_set_params(kdn_ctx* ctx, const float* grid, const float* mlp);
/* The mixture at each query: lobes[8 * n], and the probability of
 * sampling it rather than the BSDF, select[n] (may be NULL). */
int kdn_guide_infer(kdn_ctx* ctx, const kdn_guide_query* queries, uint32_t n, kdn_guide_lobe* lobes, float* select);
/* Loss sum over samples of -weight * log q(in) and its gradients (no
 * normalisation, no loss scale); grid_grad may be NULL. */
int kdn_guide_gradients(kdn_ctx* ctx, const kdn_guide_sample* samples, uint32_t n, float* grid_grad, float* mlp_grad, double* loss);
/* Adam steps on the samples (batches in turn, weights normalised by their
 * mean); losses[steps] receives each step's loss (NaN: a skipped step). */
int kdn_guide_train(kdn_ctx* ctx, const kdn_guide_sample* samples, uint32_t n, uint32_t steps, uint32_t batch, double* losses);

/* ---- serving: many render jobs on one device ------------------------------ */
/*
 * A server renders jobs of mixed scenes, sizes and sample counts together.
 * Each kdn_serve_step is one round of the integrator: the admitted jobs'
 * next slices (a number of samples of every pixel fixed by the job alone)
 * share the round's wavefront queues, and each slice folds into its own
 * job's FP64 accumulators in sample order, so a job's result is the same,
 * to the bit, alone and in any mix. Jobs are admitted in arrival order
 * while the device memory they hold (resident scenes and octrees,
 * accumulators, the shared path state) stays within the budget; a scene
 * is built once per (scene, depth cap) and kept while jobs use it, then
 * evicted least recently used. All calls on one server come from one
 * thread.
 *
 * The reference renders one frame per process (core/render.h:212-260,
 * core/worker.cpp:31-39): this is the layer a render service puts over
 * kdn_render_pass, and it has no counterpart there.
 */
typedef struct kdn_server kdn_server;

typedef struct kdn_job_desc {
    const kdn_scene* scene;  /* kept alive by the caller until the job ends */
    int32_t width, height;
    uint32_t spp;            /* samples per pixel, from sample index 0 */
    uint32_t seed;
    uint32_t surfaces;       /* the most surfaces a path stands on */
    uint32_t min_depth;      /* roulette starts at this depth */
    int32_t depth;           /* the octree's depth cap, 1..11 */
    uint32_t flags;          /* bit 0: the plain guide (a survey of the job's
                                image at admission; its guides are the job's
                                memory, 736 B a pixel); bit 1: the learned
                                guide (the job's own network over its scene's
                                region from its seed, trained after each of its
                                slices but the last on the job's own paths; one
                                slice of the job a round, up to half a round on
                                servers of 2^23 paths or more; none for a job of
                                one slice; about 420 MB of the job's memory);
                                not both */
    uint32_t pad;
} kdn_job_desc;

#define KDN_JOB_WAITING 0
#define KDN_JOB_RUNNING 1
#define KDN_JOB_DONE 2
#define KDN_JOB_FAILED 3    /* not admitted: kdn_serve_job_error says why */

typedef struct kdn_job_status {
    uint32_t state;
    uint32_t samples_done;   /* samples of every pixel folded */
    double submitted_ms;     /* times since kdn_serve_open */
    double admitted_ms;
    double finished_ms;
} kdn_job_status;

typedef struct kdn_serve_stats {
    kdn_pass_stats pass;     /* the round's integrator statistics */
    uint64_t paths;          /* paths of the round */
    uint64_t resident_bytes; /* device memory the server holds */
    uint32_t rounds;         /* rounds run so far */
    uint32_t jobs_in_round;
    uint32_t slices_in_round;
    uint32_t waiting;        /* jobs not yet admitted */
    uint32_t running;        /* admitted jobs not done */
    uint32_t pad;
    double ms;               /* the step's wall time */
} kdn_serve_stats;

/* A server on the device of a context. It keeps its own path state for
 * max_paths paths (the most of one round, at least 65536); budget_bytes
 * bounds the device memory it holds: that path state, the resident scenes
 * and the admitted jobs' accumulators. */
int kdn_serve_open(kdn_ctx* ctx, uint64_t budget_bytes, uint32_t max_paths, kdn_server** out);
int kdn_serve_submit(kdn_server* server, const kdn_job_desc* job, uint64_t* id);
/* Admit what fits, then run one round; with nothing admitted it only
 * admits. */
int kdn_serve_step(kdn_server* server, kdn_serve_stats* stats);
int kdn_serve_status(kdn_server* server, uint64_t id, kdn_job_status* status);
/* A job's accumulators, as kdn_read_accumulation lays them out. */
int kdn_serve_result(kdn_server* server, uint64_t id, double* rgb_sum, double* sq_sum, uint32_t* count);
/* Forget a job that is not running, freeing its accumulators. */
int kdn_serve_release(kdn_server* server, uint64_t id);
/* The message of the last failure on the server (NULL: of kdn_serve_open). */
const char* kdn_serve_error(const kdn_server* server);
/* Why a job failed ("" for one that did not). */
const char* kdn_serve_job_error(const kdn_server* server, uint64_t id);
void kdn_serve_close(kdn_server* server);

/* ---- The plain (highlight) guide ---------------------------------------
 * Replaces the reference's sampler/highlight_guide plugin: its frame hook
 * guideSurveyPrepared (plugins/sampler/highlight_guide/plugin.cpp:13-28)
 * that surveys the frame's highlights (surveyOf, survey.h:59) and makes
 * every pixel's guide (pixelGuides, survey.h:64), and the mixture that
 * samples the first vertex with it (mixtureDirection / mixtureDensity,
 * mixture.h). A pixel's packed guide is KDN_HG_FLOATS floats: count,
 * material, share, unused, then per lump axis (3), across (3), point (3),
 * mean (2), factor (3) and weight. */
#define KDN_HG_FLOATS 184

typedef struct kdn_hguide_stats {
    double ms;               /* survey and pixel guides, HIP events */
    uint32_t survey_points;  /* every eighth pixel each way, from the fourth */
    uint32_t highlights;     /* highlights the survey kept, over its points */
    uint32_t guided_pixels;  /* pixels with at least one lump */
    uint32_t pad;
    double mean_lumps;       /* lumps per guided pixel */
    double mean_share;       /* the draws' share the guide takes, per guided pixel */
} kdn_hguide_stats;

/* Survey the uploaded scene's highlights as the camera sees them at
 * width x height and make every pixel's guide (kept on the device for
 * kdn_render_pass flag bit 5; the scene's octree must be built). */
int kdn_highlight_guides(kdn_ctx* ctx, int32_t width, int32_t height, kdn_hguide_stats* stats);
/* Pixel (x, y)'s packed guide (KDN_HG_FLOATS floats). */
int kdn_highlight_guide_at(kdn_ctx* ctx, int32_t x, int32_t y, float* packed);
/* A packed guide (1 to 12 lumps of non-negative weight, not all zero) at a
 * vertex: the lumps' density of count directions (3 floats each) as the
 * mixture reads it (lumpDensity, guide.cpp:227-237: zero unless both
 * directions leave the side the normal faces), and count directions
 * drawn from the lumps by (pick, u, v) triples (drawnFromLump,
 * guide.cpp:239-252). */
int kdn_highlight_density(kdn_ctx* ctx, const float* packed, const float place[3], const float normal[3], const float out[3],
                          const float* dirs, uint32_t count, float* density);
int kdn_highlight_draw(kdn_ctx* ctx, const float* packed, const float place[3], const float* numbers, uint32_t count, float* drawn);

/* A chain of the survey: from a start (with two directions across its
 * surface, along which the highlight's motion is measured), by one bounce
 * off a shiny leaf of facet width `width`, to a lamp, walked from a seed
 * point on the leaf (BounceChain, plugins/sampler/highlight_guide/
 * chain.h:15-39). */
typedef struct kdn_chain_query {
    double from[3];
    double sideways[2][3];
    double seed[3];
    double width;
    uint32_t leaf;
    uint32_t light;          /* index into the scene's lights */
} kdn_chain_query;

/* Where the bounce settled and how the highlight spreads and moves there
 * (SettledChain, chain.h:44-74): found 0 leaves the rest zero; spread is
 * the upper triangle (xx, xy, yy) on the plane one unit along the axis,
 * across/up its basis; motion[k] is the point's move per unit the start
 * moves along sideways[k]. */
typedef struct kdn_chain {
    uint32_t found;
    uint32_t pad;
    double point[3];
    double axis[3];
    double across[3];
    double up[3];
    double spread[3];
    double motion[2][3];
} kdn_chain;

/* Walk n chains over the uploaded scene's leaves on the device, in FP64:
 * walkedChain (chain.h:76-97, chain.cpp:507-519), the walk the survey
 * places its highlights with. */
int kdn_highlight_chains(kdn_ctx* ctx, const kdn_chain_query* queries, uint32_t n, kdn_chain* chains);

#ifdef __cplusplus
}
#endif

#endif
