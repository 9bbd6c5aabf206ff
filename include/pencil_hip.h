/*
 * pencil_hip.h - C ABI of the MI355X (gfx950) rendering core for Pencil.
 *
 * The core replaces the data-parallel hot path behind the reference's
 * Renderer::render: SAH-BVH traversal, ray-primitive intersection (sphere,
 * triangle/quad, cube, cylinder, cone, plane), Whitted shading and path tracing
 * with next-event estimation.  Everything above this header (the .pecl parser,
 * Scene/Camera/Material, the CLI) stays on the host side of the caller.
 *
 * Conventions
 *   - Plain C, fixed-size fields, no C++ or torch types.
 *   - Every function returns PCL_OK (0) or a negative PCL_ERR_* code; the message
 *     is available from pcl_last_error(ctx).  Nothing throws across the ABI.
 *   - Calls are synchronous: work is enqueued on the context's HIP stream and the
 *     call returns when the stream is idle.  One host thread per context.
 *   - The caller owns every host buffer it passes; the library owns device
 *     buffers it allocates until pcl_scene_free / pcl_destroy.  No pointer is
 *     kept past a call.
 *   - Colors are linear HDR float RGB, row-major, row 0 = top of the image, the
 *     layout of the reference's std::vector<Color> (src/renderer.hpp:198-201).
 */
#ifndef PENCIL_HIP_H
#define PENCIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCL_ABI_VERSION 8

enum {
    PCL_OK              = 0,
    PCL_ERR_INVALID     = -1,  /* bad argument / handle */
    PCL_ERR_HIP         = -2,  /* a HIP runtime call failed */
    PCL_ERR_OOM         = -3,  /* device allocation failed */
    PCL_ERR_UNSUPPORTED = -4,  /* scene uses a feature outside this core */
    PCL_ERR_LIMIT       = -5   /* a size limit of this core was exceeded */
};

/* ---- scene description (host side, insertion order) -------------------- */

/* Primitive kinds: the six in-scope Primitive subclasses of
 * src/primitives.hpp (quad = two triangles, src/primitives.hpp:390-396). */
enum {
    PCL_PRIM_SPHERE   = 0,  /* p[0..2] center, p[3] radius            (primitives.hpp:92-134)  */
    PCL_PRIM_TRIANGLE = 1,  /* p[0..2] v0, p[3..5] v1, p[6..8] v2     (primitives.hpp:137-192) */
    PCL_PRIM_CUBE     = 2,  /* p[0..2] center, p[3] half extent       (primitives.hpp:195-252) */
    PCL_PRIM_PLANE    = 3,  /* p[0..2] point, p[3..5] normal (raw)    (primitives.hpp:255-284) */
    PCL_PRIM_CYLINDER = 4,  /* p[0..2] center, p[3..5] axis (raw), p[6] radius, p[7] height (primitives.hpp:399-505) */
    PCL_PRIM_CONE     = 5   /* p[0..2] apex, p[3..5] axis (raw), p[6] radius, p[7] height   (primitives.hpp:508-606) */
};

/* Constructor arguments of one primitive, exactly as the reference's
 * constructor receives them (derived values such as normalized axes are
 * computed by the library the way the constructors do).
 * OR-ing PCL_PRIM_STORED into `kind` says p[] instead holds the members a
 * constructed object stores (plane normal / cylinder and cone axis already
 * unit length; cylinder p[7] still the full height): the library then does
 * not normalize again, which could move a unit vector by one ulp. */
#define PCL_PRIM_STORED 0x100
typedef struct pcl_prim {
    int32_t kind;
    int32_t material_id;
    float   p[12];
} pcl_prim;

enum { PCL_MAT_DIFFUSE = 0, PCL_MAT_METAL = 1, PCL_MAT_DIELECTRIC = 2, PCL_MAT_PRINCIPLED = 3 };

/* The shading fields of Material (src/material.hpp:29-80) the core uses.
 * Texture, normal map, displacement, SSS, dispersion and principled fields
 * are outside the core; a scene that needs them is rejected. */
typedef struct pcl_material {
    int32_t kind;
    float   albedo[3];
    float   reflectivity;
    float   ior;
    float   specular;
    float   shininess;
    float   specular_color[3];
    float   emission[3];
} pcl_material;

enum { PCL_LIGHT_POINT = 0, PCL_LIGHT_DIRECTIONAL = 1 };

/* Light (src/light.hpp:20-26). */
typedef struct pcl_light {
    int32_t kind;
    float   pos_or_dir[3];
    float   intensity[3];
    float   radius;
    int32_t is_emissive;
} pcl_light;

typedef struct pcl_scene_desc {
    const pcl_prim*     prims;      /* insertion order (Scene::add_primitive) */
    int32_t             n_prims;
    const pcl_material* materials;
    int32_t             n_materials;
    const pcl_light*    lights;
    int32_t             n_lights;
    float               background[3];
    float               ambient[3];
    int32_t             use_bvh;    /* Scene::use_bvh (src/scene.hpp:53): 0 linear scan, 1 BVH with the
                                       reference's topology (host SAH build, BVH::build src/bvh.hpp:54-197),
                                       2 BVH built on the device (Morton-ordered LBVH, milliseconds for 1M
                                       primitives; another topology: the same hits except where equal-t
                                       ties resolve by DFS order) */
} pcl_scene_desc;

/* Camera constructor arguments (src/camera.hpp:29-36, camera.cpp:48-68). */
typedef struct pcl_camera {
    float   eye[3];
    float   look[3];
    float   up[3];
    float   vfov_deg;
    float   aspect;
    float   aperture;
    float   focus_dist;
    int32_t bokeh_blades;
} pcl_camera;

/* The RenderSettings fields (src/renderer.hpp:34-83) on the core's path. */
typedef struct pcl_settings {
    int32_t width;
    int32_t height;
    int32_t samples;            /* n: n*n stratified jittered samples per pixel */
    int32_t max_depth;          /* Whitted recursion depth */
    float   shadow_bias;        /* default 1e-3 */
    int32_t soft_shadow_samples;/* default 4 */
    int32_t pathtrace;          /* 0 = Whitted trace(), 1 = trace_path(), 2 = trace_path_mis()
                                   (settings.mis, src/renderer.hpp:1205-1335) */
    int32_t pathtrace_bounces;  /* default 4 */
} pcl_settings;

typedef struct pcl_stats {
    uint64_t rays_traced;       /* reference convention: trace()/trace_path() calls */
    uint64_t closest_queries;   /* closest-hit BVH queries cast */
    uint64_t anyhit_queries;    /* any-hit (shadow) BVH queries cast */
    double   kernel_ms;         /* HIP-event time of the launches on the context stream */
    double   total_ms;          /* host wall time of the call */
    uint64_t deferred_glass;    /* path-traced renders: glass vertices whose Whitted subtree a second
                                   launch evaluated (0 when none was deferred) */
} pcl_stats;

/* Ray batches for the traversal entry point. */
typedef struct pcl_ray {
    float origin[3];
    float dir[3];
    float tmin;
    float tmax;
} pcl_ray;

typedef struct pcl_hit {
    float   t;          /* hit distance, or -1 on a miss */
    int32_t prim_id;    /* insertion index of the primitive that Scene::hit
                           returns (any-hit: 1 occluded / -1 clear) */
} pcl_hit;

typedef struct pcl_scene_info {
    int32_t n_prims;
    int32_t n_nodes;        /* internal nodes of the binary BVH (the reference's topology) */
    int32_t depth;          /* longest root-to-leaf path, in edges */
    int32_t root;           /* >=0 internal node, <0 leaf ~slot, INT32_MIN empty */
    uint64_t device_bytes;  /* nodes + primitives + materials + lights */
    double   build_ms;      /* host BVH build time */
    double   upload_ms;     /* host-to-device upload time */
    int32_t n_wide_nodes;   /* 4-wide nodes the traversal reads (128 B each) */
    int32_t stack_need;     /* deepest traversal stack (entries); at most 112 */
} pcl_scene_info;

typedef struct pcl_ctx pcl_ctx;

/* Context: one HIP device, one stream.  Replaces nothing in the reference
 * (which has no device); owned by the caller of Renderer::render. */
int         pcl_init(int device, pcl_ctx** out);
void        pcl_destroy(pcl_ctx* ctx);
const char* pcl_last_error(const pcl_ctx* ctx);
int         pcl_abi_version(void);

/* Scene upload: builds the SAH BVH with the reference's topology
 * (BVH::build, src/bvh.hpp:54-197) on the host and uploads it as 128-byte
 * 4-wide nodes, with the primitives in BVH leaf order, materials and lights.
 * Replaces ParsedScene::finalize -> Scene::build_bvh (src/scene.hpp:48-51)
 * for the device.  *scene_out receives a handle >= 0.  PCL_ERR_LIMIT when
 * the 4-wide node array would exceed 4 GiB (2^25 nodes, about 67 M
 * primitives): the traversal addresses it with 32-bit byte offsets. */
int pcl_scene_upload(pcl_ctx* ctx, const pcl_scene_desc* desc, int32_t* scene_out);
int pcl_scene_free(pcl_ctx* ctx, int32_t scene);
int pcl_scene_info_get(pcl_ctx* ctx, int32_t scene, pcl_scene_info* info);
/* BVH storage order: out[slot] = insertion index of the primitive in leaf
 * `slot` (the order of BVH::storage_ after build, src/bvh.hpp:54-58). */
int pcl_scene_leaf_order(pcl_ctx* ctx, int32_t scene, int32_t* out, int32_t n);
/* The host half of pcl_scene_upload alone, needing no device: validates the
 * description, builds the BVH and reports its shape (info) and leaf order
 * (leaf_order[slot] = insertion id; n >= n_prims).  Lets a caller check the
 * topology against its own BVH::build without a GPU. */
int pcl_bvh_build_host(const pcl_scene_desc* desc, pcl_scene_info* info, int32_t* leaf_order, int32_t n);

/* Ray-batch traversal: Scene::hit (closest, any_hit = 0) or Scene::occluded
 * (any_hit = 1), src/scene.hpp:55-81.  Host buffers.  A ray whose interval
 * [tmin, tmax] is empty or has a NaN bound reports no hit. */
int pcl_trace_rays(pcl_ctx* ctx, int32_t scene, const pcl_ray* rays, int32_t n,
                   int32_t any_hit, pcl_hit* hits, pcl_stats* stats);
/* Same, with rays and hits already in device memory. */
int pcl_trace_rays_device(pcl_ctx* ctx, int32_t scene, const pcl_ray* d_rays, int32_t n,
                          int32_t any_hit, pcl_hit* d_hits, pcl_stats* stats);

/* Work counters of an instrumented launch: internal BVH nodes entered and
 * primitive tests per PCL_PRIM_* kind, summed over the launch.  They size
 * the algorithmic bytes of the work (64 B per node, the primitive's bytes
 * per test).  No reference counterpart: the reference counts neither. */
typedef struct pcl_work_counts {
    uint64_t node_visits;
    uint64_t prim_tests[6];
} pcl_work_counts;

/* pcl_trace_rays_device with the work counters (a separate, instrumented
 * kernel variant; the plain entry points run uninstrumented). */
int pcl_trace_rays_counts(pcl_ctx* ctx, int32_t scene, const pcl_ray* d_rays, int32_t n, int32_t any_hit,
                          pcl_hit* d_hits, pcl_work_counts* counts);

/* Renderer::render (src/renderer.hpp:198-295) for the in-scope settings:
 * per-pixel xorshift stream, n*n jittered samples, trace() (Whitted,
 * :403-822) or trace_path() (NEE, :829-972).  out_rgb: W*H*3 floats (host). */
int pcl_render(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam,
               const pcl_settings* settings, float* out_rgb, pcl_stats* stats);
/* Same, writing into a device buffer of W*H*3 floats. */
int pcl_render_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam,
                      const pcl_settings* settings, float* d_out_rgb, pcl_stats* stats);
/* pcl_render_device through the instrumented kernel variant, with counts. */
int pcl_render_counts(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                      float* d_out_rgb, pcl_stats* stats, pcl_work_counts* counts);

/* Deferred glass vertices of path-traced renders (pcl_render*, and the batched
 * tracer's finish launch): enable = 0 evaluates every glass vertex in place,
 * otherwise they go to a compacted second launch when the buffers fit; the
 * image is the same bit for bit either way.  max_tasks > 0 caps the task array
 * (vertices beyond it are evaluated in place), 0 sizes it by the frame.
 * Starts from the environment (PCL_DEFER_GLASS, PCL_DEFER_TASKS), else on and
 * automatic.  No reference counterpart (a device scheduling choice). */
int pcl_set_glass_deferral(pcl_ctx* ctx, int32_t enable, int64_t max_tasks);

/* ---- batched wavefront tracer --------------------------------------------
 * Renders many (scene, camera, settings) items -- scenes of uneven size and
 * resolution -- in one job on a persistent wavefront tracer: a fixed pool of
 * pixel slots runs each pixel's Renderer::render program as a state machine;
 * closest-hit, shading, any-hit and resolve kernels communicate through
 * ballot-compacted queues and one iteration is replayed as a hipGraph until
 * the queues drain.  Results equal pcl_render's for each item.  Replaces a
 * loop of Renderer::render calls (src/renderer.hpp:198) over the items. */
typedef struct pcl_batch_item {
    int32_t      scene;       /* handle from pcl_scene_upload */
    int32_t      pad;
    pcl_camera   camera;
    pcl_settings settings;
    float*       d_out_rgb;   /* device buffer of width*height*3 floats */
} pcl_batch_item;

/* Tuning knobs; 0 = library default / autotuned value. */
typedef struct pcl_wf_config {
    int32_t slots;            /* pixel slots in flight (queue depth) */
    int32_t blocks_per_cu;    /* persistent traversal blocks per CU */
    int32_t stack_k;          /* LDS traversal-stack entries per lane: 8, 16 or 32 */
    int32_t iters_per_graph;  /* wavefront iterations captured per hipGraph replay */
    int32_t drain_below;      /* once every job pixel is taken and fewer slots than this are
                                 active, each slot stops at its next sample boundary and one
                                 per-pixel launch finishes the remaining samples; -1 = never,
                                 0 = four times the traversal grids' lanes (CUs x blocks_per_cu x 512) */
    int32_t inline_shadows;   /* a vertex with at most this many shadow rays traces them in the
                                 closest-hit kernel instead of the shadow queue; -1 = never */
    int32_t streams;          /* engines the job's 64-pixel tiles are dealt to (interleaved), each
                                 with its own stream, slots and graph, run concurrently so that one
                                 engine's iteration tails and drain overlap the others' work; the
                                 slots are split among them.  1..4; 0 = by the job: 1 when every
                                 item renders one sample per pixel, 3 when an item path-traces
                                 with >= 16 spp, else 2 */
} pcl_wf_config;

typedef struct pcl_batch_stats {
    uint64_t      rays_traced;      /* reference convention, summed over items */
    uint64_t      closest_queries;
    uint64_t      anyhit_queries;
    uint64_t      iterations;       /* wavefront iterations with work */
    double        kernel_ms;        /* HIP-event time of the whole job on the stream */
    double        total_ms;
    pcl_wf_config used;
    uint64_t      finished_pixels;  /* pixels whose last samples ran in the drain's finish launch */
    uint64_t      deferred_glass;   /* (ABI 8) glass vertices of path-traced items whose Whitted subtree
                                       the finish launch left to its compacted second launch */
} pcl_batch_stats;

int pcl_render_batch(pcl_ctx* ctx, const pcl_batch_item* items, int32_t n, const pcl_wf_config* cfg,
                     pcl_batch_stats* stats);
/* Times candidate configurations on this job (coordinate descent over slots,
 * blocks_per_cu, stack_k, iters_per_graph and inline_shadows) and returns the
 * fastest; also remembered as the context default.  drain_below and streams
 * stay 0 (the rules above, applied to the job being rendered): their best
 * values grow with the samples per pixel, which a short tuning job cannot show. */
int pcl_wf_autotune(pcl_ctx* ctx, const pcl_batch_item* items, int32_t n, pcl_wf_config* best,
                    double* best_ms);

/* ---- 8-bit output on the device ------------------------------------------
 * color_to_rgb8 (src/image_io.hpp:46-59) with tonemap_color's ACES curve
 * (src/tonemap.hpp:27-61): scale by 2^exposure, ACES, clamp to [0, 1], gamma
 * 1/2.2, (uint8)(x * 255 + 0.5), per channel.  Bit-identical to the host
 * encode: the gamma step uses a threshold table built with the host's powf. */
/* d_hdr: n_px*3 floats (16-byte aligned), d_rgb8: n_px*3 bytes (4-byte aligned), device. */
int pcl_tonemap_rgb8_device(pcl_ctx* ctx, const float* d_hdr, int64_t n_px, float exposure, uint8_t* d_rgb8);
/* pcl_render followed by the device encode: only the 8-bit image (W*H*3 bytes,
 * host) leaves the device -- what write_image (src/image_io.hpp:61-113) stores. */
int pcl_render_rgb8(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                    float exposure, uint8_t* out_rgb8, pcl_stats* stats);
/* The 256-entry threshold table (t[0] = -inf; t[k] = smallest v with encode(v) >= k).
 * Host only; needs no device. */
int pcl_tonemap_thresholds(float* t256);

/* ---- differentiation: adjoint (path-replay backprop) ----------------------
 * Parameters are every material's albedo (RGB) and emission (RGB), laid out
 * grad[m*6 + 0..2] = d/d albedo_m, grad[m*6 + 3..5] = d/d emission_m, the
 * order in which a caller perturbs Material::albedo / Material::emission
 * (src/material.hpp:29-60) for finite differences.  The adjoint pass replays
 * every pixel's samples on the pixel's own RNG stream (the same paths and the
 * same float operations as the forward render, so it yields the rendered
 * pixel too) and pushes the image adjoint down them; per-pixel partials are
 * accumulated in FP64 and reduced in a fixed order, so the gradient is
 * bit-reproducible run to run.  Russian roulette's 1/p weight is held
 * constant (its branch is discrete).  At most PCL_ADJ_MAX_MATERIALS
 * materials.  No reference counterpart: the reference has no derivatives; its
 * gradient is estimated by central differences of whole renders. */
#define PCL_ADJ_MAX_MATERIALS 64

/* Replaces the scene's material table (same number of materials, same kinds).
 * BVH, primitives and lights are kept.  This is the only state an inverse-
 * rendering step changes. */
int pcl_scene_set_materials(pcl_ctx* ctx, int32_t scene, const pcl_material* materials, int32_t n);

/* Vector-Jacobian product of the rendered image: grad[p] = sum over pixels and
 * channels of d_adj[px*3+c] * dI[px*3+c]/dp.  d_adj: device, W*H*3 floats.
 * grad: host, n_materials*6 doubles. */
int pcl_render_vjp_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                          const float* d_adj, double* grad, pcl_stats* stats);

/* L2 image loss  loss = sum_px,c (I - target)^2  of the rendered image I and
 * its gradient, in one adjoint pass (I is what pcl_render returns, bit for
 * bit).  target_rgb: host W*H*3 floats; grad: host n_materials*6 doubles;
 * loss: host, one double; out_rgb: optional host image; stats: the pass's
 * query counts and time. */
int pcl_adjoint(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                const float* target_rgb, double* grad, double* loss, float* out_rgb, pcl_stats* stats);
/* Same with the target (and the optional image output) in device memory. */
int pcl_adjoint_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                       const float* d_target, double* grad, double* loss, float* d_out_rgb, pcl_stats* stats);

/* pcl_adjoint_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight optimisation loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results (bit for bit)
 * as pcl_adjoint_device; its work buffers (per-8x8-tile FP64 partials, the
 * reduced gradient, counters) live in a fixed array of
 * PCL_NOALLOC_SCRATCH_BYTES in device memory, one per GPU, held until the
 * process ends.  Launches needing more than that -- 8x8 tiles x (6 x
 * materials + 1) doubles, e.g. about 2.7 M pixels with 8 materials -- fail
 * with PCL_ERR_LIMIT before anything is enqueued.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other threads,
 * other contexts) race and must be serialised by the caller. */
#define PCL_NOALLOC_SCRATCH_BYTES (16u << 20)
int pcl_adjoint_device_noalloc(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                               const float* d_target, double* grad, double* loss, float* d_out_rgb,
                               pcl_stats* stats);

/* Device-memory helpers for callers without their own allocator. */
int pcl_device_alloc(pcl_ctx* ctx, size_t bytes, void** d_ptr);
int pcl_device_free(pcl_ctx* ctx, void* d_ptr);
int pcl_memcpy_h2d(pcl_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int pcl_memcpy_d2h(pcl_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* PENCIL_HIP_H */
