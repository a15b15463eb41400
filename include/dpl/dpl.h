/*
 * dpl.h — C ABI of the MI355X BVH build-and-trace engine (libdpl.so).
 *
 * The reference has no C ABI for this path; its device boundary is the C++ namespace API
 * diplodocus::cuda_kernels (src/gpu/renderer/gpu_renderer_api.h:8-11), GpuScene
 * (src/gpu/scene/gpu_scene.h:53-92) and the acceleration-structure plug-in interface
 * AccelerationStructure::{Build, Intersect, IntersectAny}
 * (src/acceleration/acceleration_structure.h:12-19). Each entry point below names the
 * reference interface it replaces.
 *
 * Conventions
 *   - Every function returns dpl_status (0 = OK) and never throws; dpl_last_error() gives a
 *     message for the last failure on a context.
 *   - Pointer arguments are host or device memory as stated by a dpl_mem argument; nothing
 *     is guessed. Device pointers must be valid on the context's device.
 *   - Work is ordered on the context's stream (dpl_context_set_stream). Functions that return
 *     host data (stats, downloads, host-memory outputs) synchronise that stream.
 *   - A context is used by one thread at a time; distinct contexts are independent.
 *   - The caller owns every input and output buffer; the library owns what is behind the
 *     opaque handles.
 */
#ifndef DPL_DPL_H_
#define DPL_DPL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPL_VERSION 1

typedef enum dpl_status {
    DPL_OK = 0,
    DPL_ERR_INVALID_ARG = 1,
    DPL_ERR_OOM = 2,
    DPL_ERR_DEVICE = 3,
    DPL_ERR_UNSUPPORTED = 4,
    DPL_ERR_INTERNAL = 5
} dpl_status;

typedef enum dpl_mem { DPL_MEM_HOST = 0, DPL_MEM_DEVICE = 1 } dpl_mem;

/* Mirrors AccelerationStructureType (src/config/acceleration_structure_config.h:9). */
typedef enum dpl_accel_type {
    DPL_ACCEL_DUMMY = 0,
    DPL_ACCEL_SBVH = 1,
    DPL_ACCEL_PLOC = 2,
    DPL_ACCEL_PLOC_EMC = 3,
    DPL_ACCEL_PLOC_SOBB = 4,
    DPL_ACCEL_PLOC_EMC_SOBB = 5
} dpl_accel_type;

/* Mirrors MortonType / BoundingVolumeType (src/gpu/config/gpu_acceleration_structure_config.h:5-7). */
typedef enum dpl_morton_type { DPL_MORTON30 = 0, DPL_EMC60 = 1 } dpl_morton_type;
typedef enum dpl_bv_type { DPL_BV_AABB = 0, DPL_BV_SOBB = 1 } dpl_bv_type;

/* Build parameters. Mirrors GpuAccelerationStructureConfig
 * (src/gpu/config/gpu_acceleration_structure_config.h:9-15) plus the engine's tuning knobs.
 * Zero in block_size / sort_digit_bits selects the default. */
typedef struct dpl_build_params {
    int32_t morton_type;            /* dpl_morton_type */
    int32_t bv_type;                /* dpl_bv_type */
    int32_t nn_search_radius;       /* PLOC radius, 1..32 */
    int32_t max_triangles_per_leaf; /* 1..15 */
    int32_t kdop_size;              /* SOBB direction-set hint (reference default 32) */
    int32_t block_size;             /* PLOC tile threads: 0, 128, 256 or 512 */
    int32_t sort_digit_bits;        /* radix digit width: 0, 4 or 8 */
    int32_t keep_intermediates;     /* bit 0: keep codes/order/merge list for download;
                                       bit 1: time every stage and PLOC iteration (adds events,
                                       each about 5 us of idle GPU time) */
} dpl_build_params;

/* Fills *p from an acceleration-structure type and the reference's config values.
 * Replaces BridgeAccelerationConfig (src/gpu/config/gpu_config_bridge.cc:24-60).
 * Returns DPL_ERR_UNSUPPORTED for DUMMY and SBVH (CPU-only structures). */
dpl_status dpl_build_params_from_type(int32_t accel_type, int32_t nn_search_radius, int32_t max_triangles_per_leaf,
                                      int32_t kdop_size, dpl_build_params* p);

/* Build statistics. Mirrors AccelerationStats (src/stats/acceleration_stats.h:5-17) and
 * adds the SAH cost and per-stage device times. */
typedef struct dpl_build_stats {
    double build_ms;         /* device time of the whole build (hipEvents), upload excluded */
    double upload_ms;        /* host-to-device copy of the triangles (0 for device input) */
    double stage_ms[8];      /* bounds, codes, sort, ploc, collapse, layout, sobb, stats
                                (keep_intermediates bit 1), else 0 */
    int32_t node_count;      /* final nodes: inner + leaves */
    int32_t inner_node_count;
    int32_t leaf_node_count;
    int32_t ploc_iterations;
    int64_t memory_consumption; /* bytes of the final BVH (nodes + ordered triangles + ids) */
    double sah_cost;         /* (Ct*sum_inner SA + Ci*sum_leaf SA*|leaf|) / SA(root), Ct=Ci=1 */
    int32_t tree_height;     /* longest root-to-leaf path of the final tree */
    int32_t ploc_launches;   /* multi-workgroup PLOC iterations that merged (one fused launch, or a
                                neighbour launch and a merge launch, each) */
    double ploc_launch_ms;   /* their summed device time (keep_intermediates bit 1), else 0 */
    double ploc_bytes;       /* their algorithmic bytes: 32 c_k + 32 c_k+1 + (91 + 24 x) m_k, with
                                x = SOBB directions - 3 (10 for the 26-DOP set), 0 for AABB (bit 1) */
} dpl_build_stats;

typedef struct dpl_context dpl_context;
typedef struct dpl_bvh dpl_bvh;

/* 32-byte ray and 16-byte hit. Replaces Ray / RayHit (src/scene/ray.h:7-13;
 * src/scene/ray_hit.h:8-17) at the boundary; the reference's implicit t_min (kEpsilon) is
 * explicit here. prim = -1 on a miss; prim is the input (original) triangle index. */
typedef struct dpl_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} dpl_ray;

typedef struct dpl_hit {
    float t, b1, b2;
    int32_t prim;
} dpl_hit;

#define DPL_TRACE_ANY 1u       /* any-hit (shadow) query: IntersectAny */
#define DPL_TRACE_CULL_BACK 2u /* backface culling */
#define DPL_TRACE_WIDE 4u      /* traverse the wide tree made by dpl_bvh_widen */
/* dpl_trace: sort the rays into a coherent order before tracing them (by where each ray's
 * span crosses the scene box; hits still land at each ray's own index and equal the unsorted
 * trace's bit for bit). Pays for batches of incoherent rays (about 10% on 4M random rays
 * through a 2M-triangle scene, sort included); camera rays in scanline order lose by it. */
#define DPL_TRACE_COHERENT 8u

/* Context: one device, one stream, scratch pools. The reference selects the device with
 * cudaSetDevice(0) (src/gpu/cuda_utils.h:41). */
dpl_status dpl_context_create(int32_t device, dpl_context** out);
/* BVHs and trace graphs should be destroyed first. Those still alive are detached: afterwards
 * their destroy call frees their device memory, and every other call on them returns
 * DPL_ERR_INVALID_ARG. */
void dpl_context_destroy(dpl_context* ctx);
const char* dpl_last_error(const dpl_context* ctx);
/* stream: a hipStream_t; NULL selects the device's null stream. A new context uses a
 * non-blocking stream of its own. */
dpl_status dpl_context_set_stream(dpl_context* ctx, void* stream);
dpl_status dpl_context_synchronize(dpl_context* ctx);
/* Number of compute units of the context's device (for callers that size work). */
int32_t dpl_context_cu_count(const dpl_context* ctx);

/* Builds a BVH over n triangles (n x 9 floats: v0 xyz, v1 xyz, v2 xyz).
 * Replaces AccelerationStructure::Build (src/acceleration/acceleration_structure.h:15-16)
 * for the PLOC types that the reference's factory rejects
 * (src/acceleration/acceleration_structure_factory.cc:16-17). stats may be NULL. */
dpl_status dpl_bvh_build(dpl_context* ctx, const float* tri_xyz, int64_t n, dpl_mem mem,
                         const dpl_build_params* params, dpl_bvh** out, dpl_build_stats* stats);

/* Rebuilds an existing BVH in place from new triangles (same or smaller capacity reuses
 * the device buffers). */
dpl_status dpl_bvh_rebuild(dpl_bvh* bvh, const float* tri_xyz, int64_t n, dpl_mem mem,
                           const dpl_build_params* params, dpl_build_stats* stats);

/* Builds mesh_count independent BVHs over tri_xyz[offsets[m]..offsets[m+1]) into one handle
 * in ONE persistent launch: each workgroup takes whole meshes from a queue (largest first)
 * and runs every stage for its mesh. Each mesh's tree equals dpl_bvh_build of that mesh alone
 * (codes, order, merges, final nodes, triangle order; primitive ids local to the mesh).
 * mesh_offsets: mesh_count+1 host int64, offsets[0] = 0, strictly increasing. AABB bounds
 * only (SOBB: DPL_ERR_UNSUPPORTED); block_size 256 or 512. There is no per-mesh counterpart
 * in the reference, whose factory builds one structure per scene
 * (src/acceleration/acceleration_structure_factory.cc:9-19). stats: sums over meshes, with
 * sah_cost the triangle-weighted mean of the per-mesh SAH costs. */
dpl_status dpl_bvh_build_batched(dpl_context* ctx, const float* tri_xyz, const int64_t* mesh_offsets,
                                 int32_t mesh_count, dpl_mem mem, const dpl_build_params* params, dpl_bvh** out,
                                 dpl_build_stats* stats);

/* Per-mesh view of a batched BVH (a single-scene BVH reports one mesh). Final nodes of mesh m
 * are nodes[node_offset .. node_offset + node_count) of dpl_bvh_download, its triangles
 * tri_ordered / prim_order[prim_offset .. prim_offset + prim_count). */
typedef struct dpl_mesh_info {
    int64_t node_offset;
    int64_t prim_offset;
    int32_t node_count;
    int32_t prim_count;
    int32_t tree_height;
    int32_t ploc_iterations;
    double sah_cost;
} dpl_mesh_info;

dpl_status dpl_bvh_get_meshes(const dpl_bvh* bvh, dpl_mesh_info* out /* mesh_count entries */);

void dpl_bvh_destroy(dpl_bvh* bvh);
dpl_status dpl_bvh_get_stats(const dpl_bvh* bvh, dpl_build_stats* stats);

/* SAH cost of the tree with every child bounded by its axis-aligned box: for a SOBB tree the
 * boxes are refitted from the triangles below each child (the SOBB-area cost is
 * dpl_build_stats.sah_cost); for an AABB tree it is sah_cost itself. Gives the SOBB tree's
 * cost in the same formula as an AABB tree, e.g. the reference's SBvh (src/acceleration/
 * sbvh.cc:343-367, extended with SA as dpl_build_stats.sah_cost states). Synchronises. */
dpl_status dpl_bvh_sah_aabb(const dpl_bvh* bvh, double* sah);

/* Closest-hit (or any-hit with DPL_TRACE_ANY) for n rays. Replaces
 * AccelerationStructure::Intersect / IntersectAny (acceleration_structure.h:17-18) and the
 * brute-force DummyIntersect (src/gpu/renderer/gpu_renderer_util.h:29-94).
 * Closest-hit returns the smallest t with tmin < t < tmax, ties to the lower prim id. */
dpl_status dpl_trace(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                     dpl_hit* hits, dpl_mem mem);

/* dpl_trace plus traversal counters (inner nodes visited, triangle tests), the GPU
 * counterparts of AccelerationStats::traversal_count / intersection_count
 * (src/stats/acceleration_stats.h:15-16). Synchronises. */
dpl_status dpl_trace_counters(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                              dpl_hit* hits, dpl_mem mem, int64_t* node_visits, int64_t* triangle_tests);

/* dpl_trace without any allocation, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. rays and hits are device pointers; a single-scene BVH,
 * closest hit or DPL_TRACE_ANY, optional DPL_TRACE_CULL_BACK (no DPL_TRACE_WIDE). Its ray
 * counter and traversal stacks live in a fixed 64 MiB scratch array per device (held by each
 * GPU the library runs on until the process ends), so the persistent grid is capped to the
 * workgroups whose spill stacks fit it (every tree fits: a 45-level tree still gets about 1700
 * workgroups); results equal dpl_trace's bit for bit. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other contexts, other threads) race and
 * must be ordered by the caller. Does not synchronise. */
dpl_status dpl_trace_noalloc(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, int64_t n, uint32_t flags,
                             dpl_hit* hits);

/* Batched BVH: ray i is traced against mesh ray_mesh[i] (device or host as `mem`); hit prim
 * ids are local to that mesh; a mesh id out of range gives a miss. */
dpl_status dpl_trace_batched(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, const int32_t* ray_mesh,
                             int64_t n, uint32_t flags, dpl_hit* hits, dpl_mem mem);

/* A trace pass captured once in a hipGraph and replayed with one launch: rays, ray_mesh
 * (NULL for a single-scene BVH) and hits are device buffers whose addresses the graph keeps;
 * their contents may change between launches. The graph owns its ray counter and spill
 * stacks, so the context's other calls may run between launches. The graph also keeps the
 * BVH's device addresses, sizes and height: a dpl_bvh_rebuild, dpl_bvh_widen or
 * dpl_bvh_destroy of that BVH after the capture invalidates it, and dpl_trace_graph_launch
 * then returns DPL_ERR_INVALID_ARG without launching (capture a new graph). Replaces the
 * per-frame LaunchRaytracingKernel call of GpuRenderer::StartRender
 * (src/renderer/gpu_renderer.cc:53). */
typedef struct dpl_trace_graph dpl_trace_graph;
dpl_status dpl_trace_graph_create(dpl_context* ctx, const dpl_bvh* bvh, const dpl_ray* rays, const int32_t* ray_mesh,
                                  int64_t n, uint32_t flags, dpl_hit* hits, dpl_trace_graph** out);
dpl_status dpl_trace_graph_launch(dpl_trace_graph* graph); /* on the context's stream */
void dpl_trace_graph_destroy(dpl_trace_graph* graph);

/* ---- Wide trees ----
 * Converts a built BVH into a tree of `width` (4 or 8) children per node whose child boxes
 * are quantized to 8 bits per bound (64-byte nodes for width 4, 96-byte for 8), kept beside
 * the two-child tree until the next rebuild; a batched BVH converts every mesh's tree in
 * place of its node slots. dpl_trace, dpl_trace_counters, dpl_trace_batched and
 * dpl_trace_graph_create with DPL_TRACE_WIDE traverse it; closest hits equal the two-child
 * tree's (counters then count wide nodes). SOBB trees are refitted to axis-aligned boxes
 * first. No counterpart in the reference, whose GPU intersector is brute force
 * (src/gpu/renderer/gpu_renderer_util.h:29-94) and whose SBvh is two-child
 * (src/acceleration/sbvh.h:25-31); it serves the same AccelerationStructure::Intersect /
 * IntersectAny calls (acceleration_structure.h:17-18) as dpl_trace. Synchronises. */
typedef struct dpl_wide_info {
    int32_t width;       /* 0: no wide tree */
    int32_t node_bytes;  /* 64 or 96 */
    int64_t node_count;  /* the two-child tree's node slots: wide node i converts node i */
    double widen_ms;     /* device time of the conversion (refit included) */
} dpl_wide_info;
dpl_status dpl_bvh_widen(dpl_bvh* bvh, int32_t width, dpl_wide_info* info /* may be NULL */);
dpl_status dpl_bvh_get_wide(const dpl_bvh* bvh, dpl_wide_info* info);
/* Host copy of the wide nodes: node_count x node_bytes. Node layout: origin xyz (f32),
 * meta (u32: biased exponent byte per axis, child count << 24), width child references
 * (as the two-child nodes'), 6 planes of width bytes (lo x, hi x, lo y, hi y, lo z, hi z;
 * byte k = child k), zero tail; child k's bound = fmaf(q, 2^(exponent - 127), origin). */
dpl_status dpl_bvh_download_wide(const dpl_bvh* bvh, void* nodes);

/* Whitted ray tracing: the reference's CpuRaytracer (src/renderer/cpu_raytracer.cc:50-288)
 * over the BVH — area-light next-event estimation with hash-sampled light points (seed 42),
 * Phong, mirror reflection and refraction to max_depth, emissive hits return their emission.
 * Replaces the brute-force RaytracingBounceKernel path of GpuRenderer
 * (src/gpu/renderer/gpu_renderer_api.h:10). Material as src/scene/material.h:7-18
 * (r_ior = 1 / ior as the OBJ loader sets it, obj_scene_loader.cc:118). */
typedef struct dpl_material {
    float diffuse[3];
    float specular[3];
    float transmittance[3];
    float emission[3];
    float shininess;
    float ior;
    float r_ior;
    float pad;
} dpl_material;

typedef struct dpl_whitted_params {
    int32_t width, height;
    int32_t max_depth;              /* 0..8 */
    int32_t area_light_sample_cnt;  /* per light per hit */
    int32_t backface_culling;
    int32_t reserved[3];
    float background[3];
    float camera[11];               /* pos3 dir3 up3 fov far (src/scene/camera.h:7-13) */
} dpl_whitted_params;

/* All arrays are device memory. tri_xyz: the n triangles the BVH was built from, in their
 * original order; vertex_normals: n x 9 floats or NULL (a triangle whose three normals are
 * zero uses its geometric normal); mat_id: n; materials: material_count; light_tris: the
 * emissive triangles in triangle order (the reference's area lights); image: H x W x 3. */
dpl_status dpl_render_whitted(dpl_context* ctx, const dpl_bvh* bvh, const dpl_whitted_params* params,
                              const float* tri_xyz, const float* vertex_normals, const int32_t* mat_id,
                              const dpl_material* materials, int32_t material_count, const int32_t* light_tris,
                              int32_t light_count, float* image);

/* The counters a render keeps, as the reference's CpuRaytracer keeps them in Stats:
 * RaytracingStats (src/stats/raytracing_stats.h:7-11; counted at cpu_raytracer.cc:103 per
 * pixel, :131 / :147 per reflection / refraction ray, :223 per IsShadowed call) and the
 * AccelerationStats query counters (acceleration_stats.h:14-16): one query per closest-hit or
 * any-hit trace; traversal_count = inner nodes visited and intersection_count = triangle
 * tests of those queries (the BVH's own work, not the reference SBvh's). */
typedef struct dpl_whitted_stats {
    int64_t primary_ray_count;
    int64_t secondary_ray_count;
    int64_t shadow_ray_count;
    int64_t query_count;
    int64_t traversal_count;
    int64_t intersection_count;
    double raytracing_ms; /* device time of the render kernel */
} dpl_whitted_stats;

/* dpl_render_whitted that also fills *stats (host memory). Synchronises. Replaces the
 * stats bookkeeping of CpuRaytracer::StartRender / GpuRenderer::StartRender
 * (cpu_raytracer.cc:84-88, src/renderer/gpu_renderer.cc:57-61). */
dpl_status dpl_render_whitted_stats(dpl_context* ctx, const dpl_bvh* bvh, const dpl_whitted_params* params,
                                    const float* tri_xyz, const float* vertex_normals, const int32_t* mat_id,
                                    const dpl_material* materials, int32_t material_count, const int32_t* light_tris,
                                    int32_t light_count, float* image, dpl_whitted_stats* stats);

/* Autotuner: times builds of the given workload over nn_search_radius {8, 16, 32}, block_size
 * {128, 256, 512} ({256, 512} batched) and sort_digit_bits {4, 8}, and returns in *best the
 * fastest whose SAH cost is at most (1 + sah_tolerance) x the lowest SAH among them (the
 * other fields as in *base). Results are cached in the JSON file cache_path (NULL: no cache)
 * under a key of workload kind, log2 triangle and mesh counts, code and bounds types, leaf
 * size and device; a cached entry is returned without building. The reference has no
 * counterpart (its SBvh parameters are fixed in acceleration_structure_config.h:22-29). */
typedef struct dpl_autotune_info {
    double build_ms;   /* device build time of the chosen parameters */
    double sah_cost;
    int32_t candidates; /* configurations timed (0 when served from the cache) */
    int32_t from_cache;
} dpl_autotune_info;

dpl_status dpl_autotune(dpl_context* ctx, const float* tri_xyz, int64_t n, const int64_t* mesh_offsets,
                        int32_t mesh_count, dpl_mem mem, const dpl_build_params* base, double sah_tolerance,
                        const char* cache_path, dpl_build_params* best, dpl_autotune_info* info);

/* Sizes (in elements) of the arrays dpl_bvh_download fills; any count may be zero when the
 * BVH was built without keep_intermediates. */
typedef struct dpl_bvh_layout {
    int64_t prim_count;
    int64_t node_count;       /* final 64-byte nodes */
    int64_t ploc_node_count;  /* 2n-1 PLOC nodes (32 bytes) when kept */
    int64_t iteration_count;  /* PLOC iterations when kept */
    int32_t mesh_count;
    int32_t node_bytes;       /* 64 */
} dpl_bvh_layout;
dpl_status dpl_bvh_get_layout(const dpl_bvh* bvh, dpl_bvh_layout* layout);

/* Host copies of the BVH (any pointer may be NULL):
 *   nodes        node_count x 64 B   (child boxes + child references)
 *   tri_ordered  prim_count x 48 B   (v0 xyz id | v1 xyz 0 | v2 xyz 0, leaf order)
 *   prim_order   prim_count int32    (leaf order -> input triangle index)
 * and, when built with keep_intermediates:
 *   codes        prim_count uint64   (per input triangle)
 *   sorted_idx   prim_count uint32   (stable order by code)
 *   ploc_nodes   (2n-1) x 32 B       (box, left, right; leaves first, merges in creation order)
 *   iter_counts  iteration_count int32 (clusters at the start of each PLOC iteration) */
dpl_status dpl_bvh_download(const dpl_bvh* bvh, void* nodes, void* tri_ordered, int32_t* prim_order, uint64_t* codes,
                            uint32_t* sorted_idx, void* ploc_nodes, int32_t* iter_counts);

/* Stand-alone build stages for parity checks and the autotuner (device memory). */
dpl_status dpl_compute_codes(dpl_context* ctx, const float* tri_xyz, int64_t n, int32_t morton_type, uint64_t* codes,
                             float* bounds12);
dpl_status dpl_radix_sort_pairs(dpl_context* ctx, uint64_t* keys, uint32_t* values, int64_t n, int32_t key_bits,
                                int32_t digit_bits);

/* ---- Differentiable diffuse path tracing ----
 * Replaces the GPU "bounce" kernel launcher LaunchPathtracingKernel
 * (src/gpu/renderer/gpu_renderer_api.h:11) with a diffuse path tracer over the BVH.
 * camera: pos3 dir3 up3 fov far (src/scene/camera.h:7-13). */
typedef struct dpl_render_params {
    int32_t width, height, spp, max_depth;
    uint64_t seed;
    float camera[11];
    int32_t material_count;
    int32_t reserved;
} dpl_render_params;

/* Device memory: mat_id (prim_count int32, input order), albedo/emission (M x 3 floats),
 * image (W*H*3 floats, row-major, src/framebuffer/framebuffer.cc:47-52). */
dpl_status dpl_render_forward(dpl_context* ctx, const dpl_bvh* bvh, const dpl_render_params* rp, const int32_t* mat_id,
                              const float* albedo, const float* emission, float* image);

/* Adjoint by path replay: given dL/dimage (W*H*3), accumulates dL/dalbedo (M x 3) and
 * dL/demission (M x 3) into the output arrays (zeroed by the call). Device memory. */
dpl_status dpl_render_adjoint(dpl_context* ctx, const dpl_bvh* bvh, const dpl_render_params* rp,
                              const int32_t* mat_id, const float* albedo, const float* emission,
                              const float* d_image, float* grad_albedo, float* grad_emission);

#ifdef __cplusplus
}
#endif

#endif /* DPL_DPL_H_ */
