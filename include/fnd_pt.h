/*
 * fnd_pt.h — C ABI of the MI355X path-tracing pass.
 *
 * This is the drop-in boundary between a Foundation-style renderer and the
 * HIP kernels that trace rays on a gfx950 GPU.  Plain C: no C++ types, no
 * exceptions, plain pointers and sizes.  Every call returns an fnd_status and
 * never aborts; the text of the last failure is kept per context
 * (fnd_pt_last_error).
 *
 * Where each entry point sits against the reference (mos9527/Foundation):
 *   - fnd_pt_create / fnd_pt_destroy replace the device-side half of
 *     Renderer::Renderer(RHIApplicationObjectHandle<RHIDevice>, Core::Allocator*)
 *     and Renderer::~Renderer() (src/Renderer/Renderer.hpp:41-42,
 *     src/Renderer/Renderer.cpp:34-37,402-406).  The borrowed allocator is the
 *     C form of the Vulkan allocation callbacks that forward to Core::Allocator
 *     (src/Platform/RHI/Vulkan/Application.hpp:13-25).
 *   - fnd_pt_scene_create replaces the vertex/index upload of the constructor
 *     (src/Renderer/Renderer.cpp:133-198): it copies the caller's arrays and
 *     builds the acceleration structure on the device.
 *   - fnd_pt_render replaces the raster pass recorded by Renderer::Record and
 *     submitted by Renderer::Draw (src/Renderer/Renderer.cpp:312-366,386-393).
 *   - fnd_pt_wait is the fence wait after submit (src/Renderer/Renderer.cpp:394).
 *   - Handles are uint64_t with ~0 as the invalid value, the convention of
 *     src/Platform/RHI/Details/Details.hpp:11-12; the context is the factory that
 *     owns every object it hands out and destroying it destroys them
 *     (src/Platform/RHI/Details/Details.hpp:39-46,125-141).
 *   - fnd_pt_trace / fnd_pt_trace_buffers have no counterpart in the reference
 *     (it has no compute dispatch, src/Platform/RHI/Command.hpp:32-116); they are
 *     the ray-query entry used by the traversal benchmark and the parity tests.
 *
 * Threading: a context is single-threaded, like the reference's one render
 * thread (src/Editor/Editor.cpp:20-23).  One context owns one HIP stream for
 * its work, plus a second one that answers radiance-cache queries during a
 * frame's last bounces (joined before fnd_pt_render returns) and puts a ray
 * query's rays in spatial order while the previous query's traversal runs
 * (the traversal waits for it), and a third
 * that runs a frame's radiance-cache training after fnd_pt_render returns,
 * overlapped with the next call until that call first reads the cache;
 * fnd_pt_wait, every fnd_pt_cache_* call and fnd_pt_destroy wait for it.
 * Calls are synchronous unless FND_ASYNC is set, then fnd_pt_wait is required.
 */
#ifndef FND_PT_H
#define FND_PT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FND_PT_ABI_VERSION 7u

typedef uint64_t fnd_handle;
#define FND_INVALID_HANDLE (~(uint64_t)0)

typedef enum fnd_status {
    FND_OK = 0,
    FND_E_INVALID_ARG = 1,
    FND_E_OOM_HOST = 2,
    FND_E_OOM_DEVICE = 3,
    FND_E_DEVICE = 4,
    FND_E_QUEUE_OVERFLOW = 5,
    FND_E_UNSUPPORTED = 6,
    FND_E_NOT_FOUND = 7
} fnd_status;

/* Host allocator: the C form of Core::Allocator injection
 * (src/Core/Allocator/Allocator.hpp:17-41).  NULL means malloc/free. */
typedef struct fnd_host_allocator {
    void* user;
    void* (*alloc)(void* user, size_t size, size_t align);
    void* (*realloc)(void* user, void* ptr, size_t size, size_t align);
    void (*free)(void* user, void* ptr);
} fnd_host_allocator;

typedef void (*fnd_log_fn)(void* user, int level, const char* message);

typedef struct fnd_pt_create_desc {
    int device;                     /* HIP device ordinal */
    uint32_t flags;                 /* reserved, 0 */
    uint64_t device_memory_budget;  /* bytes; 0 = unlimited.  Allocations past it fail with FND_E_OOM_DEVICE */
    fnd_log_fn log;                 /* optional */
    void* log_user;
} fnd_pt_create_desc;

typedef struct fnd_pt_ctx fnd_pt_ctx;

/* ---------------------------------------------------------------- rays */

/* 32 bytes.  A hit is accepted iff tmin < t < tmax.  dir need not be unit. */
typedef struct fnd_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} fnd_ray;

/* 16 bytes.  prim = FND_NO_HIT on a miss.  For a closest-hit query the
 * record is the lexicographic minimum of (t, prim) over all accepted hits,
 * (u, v) the barycentric weights of vertices 1 and 2.  For an any-hit query
 * prim is some accepted primitive (which one is unspecified). */
typedef struct fnd_hit {
    float t;
    uint32_t prim;
    float u;
    float v;
} fnd_hit;
#define FND_NO_HIT 0xFFFFFFFFu

/* ---------------------------------------------------------------- scenes */

typedef enum fnd_material_type {
    FND_MAT_LAMBERT = 0,  /* albedo / pi */
    FND_MAT_GGX = 1,      /* conductor: GGX D, Smith height-correlated G2, Schlick F with F0 = albedo, alpha = roughness^2 */
    FND_MAT_EMISSIVE = 2  /* emits `emission`, reflects nothing */
} fnd_material_type;

typedef struct fnd_material {
    uint32_t type;        /* fnd_material_type */
    float albedo[3];
    float roughness;
    float emission[3];    /* a Lambert surface may also emit (furnace scenes) */
} fnd_material;

/* Parallelogram light: points corner + a*edge0 + b*edge1, a,b in [0,1],
 * one-sided, emitting toward cross(edge0, edge1).  Sampled by next-event
 * estimation; the caller must also put its two triangles in a mesh with an
 * FND_MAT_EMISSIVE material and set their fnd_mesh_desc.light_ids to it. */
typedef struct fnd_quad_light {
    float corner[3];
    float edge0[3];
    float edge1[3];
    float radiance[3];
} fnd_quad_light;

/* Indexed triangle mesh, the layout of src/Renderer/Renderer.cpp:23-27 with
 * 32-bit indices (src/Platform/RHI/Vulkan/Command.cpp:289-305). */
typedef struct fnd_mesh_desc {
    const float* positions;        /* 3 floats per vertex */
    uint32_t num_vertices;
    const uint32_t* indices;       /* 3 per triangle */
    uint32_t num_triangles;
    const uint32_t* material_ids;  /* per triangle, or NULL for material 0 */
    const int32_t* light_ids;      /* per triangle: index into lights or -1; NULL = all -1 */
} fnd_mesh_desc;

/* Row-major 3x4 affine maps.  The primitive id of a hit in instance i is
 * id_base(i) + triangle index in its mesh, id_base = running sum over the
 * instances before i of their mesh's triangle count. */
typedef struct fnd_instance_desc {
    uint32_t mesh;
    uint32_t flags;                /* reserved, 0 */
    float object_to_world[12];
    float world_to_object[12];     /* must be given; used bit-for-bit to move rays into object space */
} fnd_instance_desc;

typedef struct fnd_scene_desc {
    const fnd_mesh_desc* meshes;
    uint32_t num_meshes;
    const fnd_instance_desc* instances;  /* NULL/0: single level, mesh 0 in world space */
    uint32_t num_instances;
    const fnd_material* materials;
    uint32_t num_materials;
    const fnd_quad_light* lights;
    uint32_t num_lights;
    float background[3];                 /* radiance of rays that leave the scene */
    uint32_t build_flags;                /* reserved, 0 */
} fnd_scene_desc;

typedef struct fnd_scene_info {
    uint64_t num_triangles;        /* triangles stored (sum over meshes) */
    uint64_t num_prims;            /* primitive ids addressable (sum over instances) */
    uint64_t num_nodes;            /* 64-byte wide nodes */
    uint32_t max_depth;            /* deepest root-to-leaf chain of wide nodes (TLAS + BLAS) */
    uint32_t stack_entries;        /* worst-case traversal stack depth */
    uint64_t device_bytes;         /* device memory held by the scene */
    double build_ms;               /* host time to build (BVH + encode + upload) */
    float bounds_lo[3];
    float bounds_hi[3];
} fnd_scene_info;

/* ---------------------------------------------------------------- cameras and jobs */

/* Pinhole camera, glm::lookAt conventions with a Z-up world as in
 * src/Renderer/Renderer.cpp:373-380: forward = normalize(at - eye),
 * right = normalize(cross(forward, up)), up' = cross(right, forward). */
typedef struct fnd_camera {
    float eye[3];
    float at[3];
    float up[3];
    float fov_y_deg;
} fnd_camera;

#define FND_JOB_USE_CACHE 1u        /* terminate at the first diffuse vertex into the radiance cache */
#define FND_JOB_TRAIN_CACHE 2u      /* also trace training paths and take one optimiser step */

typedef struct fnd_render_job {
    fnd_handle scene;
    fnd_handle framebuffer;
    fnd_camera camera;
    uint32_t spp;                   /* samples per pixel added by this call */
    uint32_t max_depth;             /* bounces; pbrt convention: 1 = direct light only */
    uint32_t rr_depth;              /* Russian roulette from this bounce on; >= max_depth disables it */
    uint32_t flags;                 /* FND_JOB_* */
    uint64_t seed;
    uint32_t sample_offset;         /* index of the first sample (progressive rendering) */
    uint32_t reserved;
} fnd_render_job;

#define FND_FB_FP64 1u  /* also accumulate in double (oracle images only) */

/* ---------------------------------------------------------------- stats and tuning */

typedef struct fnd_pt_stats {
    double last_kernel_ms;          /* device time of the last trace launch (HIP events on the context stream) */
    double last_call_ms;            /* device time of the whole last trace/render call */
    double trace_kernel_ms_total;   /* cumulative device time of trace launches, resolved at each wait */
    uint64_t trace_launches;        /* trace launches whose time is in trace_kernel_ms_total */
    uint64_t rays_traced;           /* cumulative */
    uint64_t node_visits;           /* cumulative, only with FND_TRACE_COUNT */
    uint64_t tri_tests;             /* cumulative, only with FND_TRACE_COUNT */
    uint64_t instance_visits;       /* cumulative, only with FND_TRACE_COUNT */
    uint64_t paths;                 /* cumulative paths rendered */
    uint64_t render_rays;           /* cumulative rays of fnd_pt_render (extension + shadow) */
    uint64_t device_bytes;          /* current device allocation */
    uint32_t num_cus;
    uint32_t last_grid;             /* blocks of the last trace launch */
} fnd_pt_stats;

/* Knobs the autotuner searches.  0 means "library default". */
typedef struct fnd_tuning {
    uint32_t block_size;            /* threads; (block, stack_lds) pairs built: 64x16, 64x32, 128x8, */
    uint32_t stack_lds;             /* 128x12, 128x16, 128x32, 256x8, 256x12, 256x16 (LDS stack entries) */
    uint32_t waves_per_cu;          /* persistent grid = CUs * waves_per_cu waves, at most what fits */
    uint32_t refill;                /* refetch rays once this many lanes of a wave are idle (1..64) */
    uint32_t queue_capacity;        /* path slots of the wavefront (render); a window of paths holds up to
                                       8 per slot, and slots whose path ended take the window's next paths */
    uint32_t tile;                  /* pixel tile edge for the pixel order of the path space (render) */
} fnd_tuning;

#define FND_TRACE_CLOSEST 0u
#define FND_TRACE_ANYHIT 1u
#define FND_TRACE_COUNT 2u          /* count node visits / triangle tests into fnd_pt_stats */
#define FND_ASYNC 0x100u            /* do not wait for the device before returning */

/* ---------------------------------------------------------------- context */

fnd_status fnd_pt_create(const fnd_pt_create_desc* desc, const fnd_host_allocator* allocator, fnd_pt_ctx** out_ctx);
void fnd_pt_destroy(fnd_pt_ctx* ctx);
const char* fnd_pt_last_error(const fnd_pt_ctx* ctx);
const char* fnd_pt_status_string(fnd_status status);
uint32_t fnd_pt_abi_version(void);
fnd_status fnd_pt_wait(fnd_pt_ctx* ctx);
fnd_status fnd_pt_get_stats(fnd_pt_ctx* ctx, fnd_pt_stats* out);
fnd_status fnd_pt_reset_stats(fnd_pt_ctx* ctx);
fnd_status fnd_pt_set_tuning(fnd_pt_ctx* ctx, const fnd_tuning* tuning);
fnd_status fnd_pt_get_tuning(fnd_pt_ctx* ctx, fnd_tuning* out);

/* ---------------------------------------------------------------- scenes */

fnd_status fnd_pt_scene_create(fnd_pt_ctx* ctx, const fnd_scene_desc* desc, fnd_handle* out_scene);
fnd_status fnd_pt_scene_destroy(fnd_pt_ctx* ctx, fnd_handle scene);
fnd_status fnd_pt_scene_info(fnd_pt_ctx* ctx, fnd_handle scene, fnd_scene_info* out);
/* Copies the encoded 64-byte wide nodes and the 48-byte leaf triangles
 * (host copies of what the kernels read), for host-side checks.  Pass NULL
 * pointers to query the byte sizes. */
fnd_status fnd_pt_scene_export(fnd_pt_ctx* ctx, fnd_handle scene, void* nodes, uint64_t* nodes_bytes,
                               void* tris, uint64_t* tris_bytes);

/* ---------------------------------------------------------------- device buffers */

fnd_status fnd_pt_buffer_create(fnd_pt_ctx* ctx, uint64_t bytes, fnd_handle* out_buffer);
fnd_status fnd_pt_buffer_destroy(fnd_pt_ctx* ctx, fnd_handle buffer);
fnd_status fnd_pt_buffer_write(fnd_pt_ctx* ctx, fnd_handle buffer, uint64_t offset, const void* src, uint64_t bytes);
fnd_status fnd_pt_buffer_read(fnd_pt_ctx* ctx, fnd_handle buffer, uint64_t offset, void* dst, uint64_t bytes);

/* ---------------------------------------------------------------- ray queries */

/* At most 2^31 - 1 rays per call (FND_E_INVALID_ARG beyond).
 * Host arrays in and out (uploads, traces, downloads). */
fnd_status fnd_pt_trace(fnd_pt_ctx* ctx, fnd_handle scene, const fnd_ray* rays, uint32_t num_rays,
                        fnd_hit* hits, uint32_t flags);
/* Device-resident rays and hits (fnd_ray / fnd_hit arrays in buffers).
 * Both calls take a batch of 131072 rays or more in the order of its rays'
 * origin cells on a coarse grid over the scene, for locality; each ray's
 * record is its own whatever the order.  With FND_ASYNC the caller keeps the
 * rays unchanged until fnd_pt_wait. */
fnd_status fnd_pt_trace_buffers(fnd_pt_ctx* ctx, fnd_handle scene, fnd_handle rays, fnd_handle hits,
                                uint32_t num_rays, uint32_t flags);

/* Allocation-free variant of fnd_pt_trace_buffers.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * rays and hits are device pointers of the caller's (16-byte aligned, n
 * fnd_ray / fnd_hit).  The call allocates nothing: the traversal's scratch
 * (its work counter and stack spill area) is a fixed 16 MiB device array of
 * the library, one per GPU, held until the process ends.  Every call on a
 * GPU shares that array, so calls that overlap on one GPU (on different
 * streams or contexts) race; calls on one context's stream are ordered.  When
 * the scene's stack spill for the full persistent grid does not fit, fewer
 * persistent blocks run (same records); a scene whose spill for one block
 * does not fit is refused with FND_E_INVALID_ARG.  Flags: FND_TRACE_ANYHIT,
 * FND_ASYNC (no FND_TRACE_COUNT, no per-launch timing). */
fnd_status fnd_pt_trace_noalloc(fnd_pt_ctx* ctx, fnd_handle scene, const fnd_ray* rays, fnd_hit* hits,
                                uint32_t num_rays, uint32_t flags);

/* ---------------------------------------------------------------- rendering */

fnd_status fnd_pt_framebuffer_create(fnd_pt_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags,
                                     fnd_handle* out_fb);
fnd_status fnd_pt_framebuffer_destroy(fnd_pt_ctx* ctx, fnd_handle fb);
fnd_status fnd_pt_framebuffer_clear(fnd_pt_ctx* ctx, fnd_handle fb);
/* Mean radiance per pixel, RGB, row 0 at the top: width*height*3 values. */
fnd_status fnd_pt_framebuffer_read(fnd_pt_ctx* ctx, fnd_handle fb, float* rgb, uint32_t* out_samples);
fnd_status fnd_pt_framebuffer_read_f64(fnd_pt_ctx* ctx, fnd_handle fb, double* rgb, uint32_t* out_samples);
/* The reference's presentation format, R8G8B8A8_UNORM (src/Renderer/Renderer.cpp:39-44),
 * converted on the device: width*height*4 bytes, row 0 at the top, alpha 255.
 * x = mean * exposure; byte = number of thresholds ((k - 0.5) / 255)^2.2,
 * k = 1..255, at or below x (that is round(255 * clamp(x, 0, 1)^(1/2.2)), with
 * the thresholds rounded from double to float once); NaN and x <= 0 give 0. */
fnd_status fnd_pt_framebuffer_read_rgba8(fnd_pt_ctx* ctx, fnd_handle fb, float exposure, uint8_t* rgba,
                                         uint32_t* out_samples);
/* The same bytes written into device memory of the caller's, for instance
 * the swapchain image's memory imported into HIP, so the frame reaches the
 * presentation engine without a round trip through host memory (the
 * reference stages through a mapped buffer and CopyBufferToImage,
 * src/Platform/RHI/Resource.hpp:62-68, src/Platform/RHI/Command.hpp:94).
 * Row y starts at device_rgba + y * row_pitch; row_pitch is a multiple of 4,
 * at least 4 * width; device_rgba is 4-byte aligned.  Bytes between rows are
 * not written.  Returns after the device has finished. */
fnd_status fnd_pt_framebuffer_present(fnd_pt_ctx* ctx, fnd_handle fb, float exposure, void* device_rgba,
                                      uint64_t row_pitch, uint32_t* out_samples);
/* Relative mean squared error of framebuffer fb's mean radiance against
 * reference ref's, evaluated on the device: the mean over pixels and
 * channels of (x - y)^2 / (y^2 + eps) in double precision, with x as
 * fnd_pt_framebuffer_read forms it (FP32 sum times 1/samples in FP32) and y
 * as fnd_pt_framebuffer_read_f64 forms it (ref needs FND_FB_FP64).  For
 * progressive renders that check convergence after every frame: only the
 * 8-byte result comes back, so the device does not wait for an image
 * readback between frames.  No counterpart in the reference (it has no
 * progressive renderer).  FND_E_INVALID_ARG if the sizes differ, ref has no
 * FP64 sums, or eps is not positive.  Returns after the device has finished. */
fnd_status fnd_pt_framebuffer_rel_mse(fnd_pt_ctx* ctx, fnd_handle fb, fnd_handle ref, double eps, double* out);
/* Renders all jobs in one wavefront launch sequence: their paths form one path
 * space, processed in windows (see fnd_tuning.queue_capacity). */
fnd_status fnd_pt_render(fnd_pt_ctx* ctx, const fnd_render_job* jobs, uint32_t num_jobs);

/* ---------------------------------------------------------------- camera and generator checks */

/* The camera basis the renderer forms for a width x height image (host only,
 * no device): out14 = eye[3], forward[3], right[3], up'[3], tan(fov_y / 2),
 * aspect, from the glm::lookAt convention of fnd_camera, formed in double and
 * rounded to float once.  FND_E_INVALID_ARG (also from fnd_pt_render and
 * fnd_pt_camera_rays) for a camera without a basis: non-finite values,
 * eye == at, fov_y outside (0, 180) degrees, or an up vector of zero length or
 * within 1e-6 radians of the view direction. */
fnd_status fnd_camera_basis(const fnd_camera* camera, uint32_t width, uint32_t height, float* out14);

/* Device building blocks exposed for host-side checks (no counterpart in the
 * reference; the parity tests compare them with the host oracle bit for bit).
 * fnd_pt_camera_rays: the camera ray each pixel of a width x height image
 * starts sample `sample` with (jitter from the generator of (seed, sample,
 * pixel, bounce 0), as fnd_pt_render draws it): width * height fnd_ray,
 * row-major, row 0 at the top, tmin 0, tmax 3e38, unit directions.
 * fnd_pt_rng_draws: the first `draws` 32-bit outputs of the PCG32 generator
 * the renderer keys by (seed, sample, pixel, bounce), for pixels pixel0 ..
 * pixel0 + count - 1: out[(p - pixel0) * draws + k]; at most 2^28 outputs.
 * Both return after the device has finished. */
fnd_status fnd_pt_camera_rays(fnd_pt_ctx* ctx, const fnd_camera* camera, uint32_t width, uint32_t height,
                              uint64_t seed, uint32_t sample, fnd_ray* out);
fnd_status fnd_pt_rng_draws(fnd_pt_ctx* ctx, uint64_t seed, uint32_t sample, uint32_t pixel0, uint32_t count,
                            uint32_t bounce, uint32_t draws, uint32_t* out);

/* ---------------------------------------------------------------- compute command lists */

/* The compute half the reference's RHI lacks.  Its RHICommandList records
 * transitions, pipeline and descriptor binds and draws, and a queue submits
 * lists with a fence the host waits on (src/Platform/RHI/Command.hpp:32-116,
 * src/Platform/RHI/Device.hpp:15-37, src/Renderer/Renderer.cpp:312-394), but
 * there is no Dispatch, the pipeline state is graphics-only
 * (src/Platform/RHI/Vulkan/PipelineState.cpp:103-110) and no descriptor type
 * names a storage image (src/Platform/RHI/Common.hpp:67-74).  Here a command
 * list records, in the same order of calls, compute pipelines, descriptor
 * sets (scenes as acceleration structures, storage buffers, storage images),
 * push constants and dispatches; fnd_pt_submit runs the recorded lists in
 * order on the context's stream and signals a fence, and fnd_pt_fence_wait is
 * the host's fence wait.  Lists and fences belong to the context that made
 * them (destroying it destroys them, the factory rule of
 * src/Platform/RHI/Details/Details.hpp:39-46).
 *
 * Pipelines and what a dispatch of each needs (bindings by number):
 *   FND_PIPELINE_PATH_TRACE    0 scene, 1 accumulation image (framebuffer);
 *                              constants fnd_path_trace_constants; dispatch
 *                              (width, height, spp) of that framebuffer; adds
 *                              spp samples per pixel, as one fnd_pt_render job.
 *                              Consecutive path-trace dispatches into
 *                              different images run as one batched launch.
 *   FND_PIPELINE_TRACE_CLOSEST 0 scene, 1 rays (storage buffer of fnd_ray),
 *   FND_PIPELINE_TRACE_ANYHIT  2 hits (storage buffer of fnd_hit); dispatch
 *                              (rays, 1, 1), both buffers large enough.
 *   FND_PIPELINE_PRESENT       0 accumulation image, 1 R8G8B8A8 storage image
 *                              in device memory (device_ptr, row pitch in
 *                              `range`); constants: float exposure; dispatch
 *                              (width, height, 1); the bytes of
 *                              fnd_pt_framebuffer_present.
 *   FND_PIPELINE_CLEAR         0 accumulation image; dispatch (width, height,
 *                              1); the image's sums and sample count to zero.
 * Recording checks each call against the bound pipeline and returns
 * FND_E_INVALID_ARG for a mismatch; a list with a refused call cannot be
 * submitted until the next fnd_pt_cmd_begin.  Resources must stay alive until
 * the fence of their submission has signalled. */
typedef struct fnd_pt_cmd fnd_pt_cmd;
typedef struct fnd_pt_fence fnd_pt_fence;

typedef enum fnd_pipeline {
    FND_PIPELINE_PATH_TRACE = 1,
    FND_PIPELINE_TRACE_CLOSEST = 2,
    FND_PIPELINE_TRACE_ANYHIT = 3,
    FND_PIPELINE_PRESENT = 4,
    FND_PIPELINE_CLEAR = 5
} fnd_pipeline;

typedef enum fnd_descriptor_type {
    FND_DESCRIPTOR_ACCELERATION_STRUCTURE = 1,  /* resource: a scene handle */
    FND_DESCRIPTOR_STORAGE_BUFFER = 2,          /* resource: a buffer handle; offset, range (0: to the end) */
    FND_DESCRIPTOR_STORAGE_IMAGE = 3            /* resource: a framebuffer handle, or FND_INVALID_HANDLE with
                                                   device_ptr = R8G8B8A8 device memory and range = row pitch */
} fnd_descriptor_type;

typedef struct fnd_descriptor {
    uint32_t binding;
    uint32_t type;                 /* fnd_descriptor_type */
    fnd_handle resource;
    uint64_t offset;
    uint64_t range;
    void* device_ptr;
} fnd_descriptor;

typedef struct fnd_path_trace_constants {
    fnd_camera camera;
    uint32_t max_depth;
    uint32_t rr_depth;
    uint32_t flags;                /* FND_JOB_* */
    uint32_t sample_offset;
    uint64_t seed;
} fnd_path_trace_constants;

fnd_status fnd_pt_cmd_create(fnd_pt_ctx* ctx, fnd_pt_cmd** out);
void fnd_pt_cmd_destroy(fnd_pt_cmd* cmd);
/* Begin() of RHICommandList: forgets what was recorded (and a refused call). */
fnd_status fnd_pt_cmd_begin(fnd_pt_cmd* cmd);
fnd_status fnd_pt_cmd_set_pipeline(fnd_pt_cmd* cmd, uint32_t pipeline);
/* Replaces the bound set; bindings must be distinct. */
fnd_status fnd_pt_cmd_bind_descriptors(fnd_pt_cmd* cmd, const fnd_descriptor* set, uint32_t count);
fnd_status fnd_pt_cmd_push_constants(fnd_pt_cmd* cmd, const void* data, uint32_t bytes);
fnd_status fnd_pt_cmd_dispatch(fnd_pt_cmd* cmd, uint32_t x, uint32_t y, uint32_t z);
fnd_status fnd_pt_cmd_end(fnd_pt_cmd* cmd);
/* Commands recorded since the last begin (for tools and tests). */
uint32_t fnd_pt_cmd_size(const fnd_pt_cmd* cmd);

fnd_status fnd_pt_fence_create(fnd_pt_ctx* ctx, fnd_pt_fence** out);
void fnd_pt_fence_destroy(fnd_pt_fence* fence);
/* Runs `count` ended lists in order; `fence` (may be NULL) signals when the
 * device has finished them. */
fnd_status fnd_pt_submit(fnd_pt_ctx* ctx, fnd_pt_cmd* const* cmds, uint32_t count, fnd_pt_fence* fence);
/* The host waits until the fence has signalled (a fence never submitted
 * counts as signalled). */
fnd_status fnd_pt_fence_wait(fnd_pt_ctx* ctx, fnd_pt_fence* fence);

/* ---------------------------------------------------------------- frame arena */

/* Bump allocation over one block of memory, with the semantics of the
 * reference's Core::StackAllocator (src/Core/Allocator/StackAllocator.hpp:9-63,
 * src/Core/Allocator/StackAllocator.cpp:5-29):
 *   - fnd_arena_reset(a, memory, size) is Reset(Arena{memory, size}); memory
 *     NULL is Reset(), an arena that gives out nothing.
 *   - fnd_arena_alloc(a, size) takes the next `size` bytes, unaligned;
 *     fnd_arena_alloc_aligned(a, size, alignment) rounds the start AND the
 *     end of the block up to `alignment`.
 *   - Both return NULL for size 0, for an arena without memory and when the
 *     (padded) block would pass the end; nothing is taken then.
 *   - fnd_arena_used(a) is GetUsedMemory(): bytes from the start of the
 *     memory to the next free byte, padding included.
 *   - Freeing is a no-op in the reference and there is no reallocation (it
 *     throws, StackAllocator.hpp:51-53), so neither has an entry point here.
 * Differences: `alignment` must be a power of two (NULL otherwise; the
 * reference's mask arithmetic assumes one, Allocator.hpp:9-11), a block that
 * would wrap the 64-bit address space returns NULL, and concurrent allocations
 * from one arena are safe (one compare-and-swap on `current`, where the
 * reference's multi-threaded variant loads, compares and stores separately,
 * StackAllocator.cpp:8-15).  fnd_arena_reset must not overlap an allocation.
 * The memory is never read or written, so it may be device memory. */
typedef struct fnd_arena {
    uint64_t memory;               /* address of the block; 0: none */
    uint64_t end;                  /* memory + size */
    uint64_t current;              /* next free address */
} fnd_arena;

void fnd_arena_reset(fnd_arena* arena, void* memory, uint64_t size);
void* fnd_arena_alloc(fnd_arena* arena, uint64_t size);
void* fnd_arena_alloc_aligned(fnd_arena* arena, uint64_t size, uint64_t alignment);
uint64_t fnd_arena_used(const fnd_arena* arena);

/* Per-frame ray queues.  Every fnd_pt_render resets the context's frame arena
 * and bump-allocates from it, 256-byte aligned, the buffers of the wavefront's
 * path slots (path state, both extension queues, the shadow queue, hits, cache
 * queries, the window's radiance and the slot lists):
 * fnd_pt_queue_bytes(capacity) bytes for `capacity` slots.  This is the
 * reference's command list, which resets a StackAllocator over its arena at
 * Begin() (src/Platform/RHI/Vulkan/Command.cpp:38-50, Command.hpp:44-48).
 * By default the block is the library's own: one device allocation, grown
 * when a call needs more slots.  fnd_pt_set_frame_arena gives it the caller's
 * memory instead (device memory of the context's device, 256-byte aligned,
 * kept alive by the caller until it is replaced or the context is destroyed;
 * the call waits for the context's stream first).  The library then allocates
 * nothing for the queues, and a call whose slots (tuning.queue_capacity, at
 * most the call's paths) do not fit runs with the most slots that do; images
 * do not depend on the slot count.  A block that holds fewer than
 * min(1024, paths of the call) slots fails the call with FND_E_OOM_DEVICE.
 * memory NULL goes back to the library's own block. */
fnd_status fnd_pt_queue_bytes(uint32_t capacity, uint64_t* bytes);
fnd_status fnd_pt_set_frame_arena(fnd_pt_ctx* ctx, void* memory, uint64_t bytes);

typedef struct fnd_frame_arena_info {
    uint64_t size;                 /* bytes of the block (0 before the first render with the library's block) */
    uint64_t used;                 /* bytes the last render call's queues took */
    uint32_t batch_paths;          /* path slots of the last render call */
    uint32_t caller_owned;         /* 1: the block came from fnd_pt_set_frame_arena */
} fnd_frame_arena_info;

fnd_status fnd_pt_frame_arena_info(fnd_pt_ctx* ctx, fnd_frame_arena_info* out);

/* ---------------------------------------------------------------- autotuner */

/* Launch-configuration search for a render workload (no counterpart in the
 * reference).  Times renders of `jobs` with spp = 1 and cache training off,
 * into scratch framebuffers (the callers' images are not touched), over
 * fnd_tuning candidates: the library default plus a fixed pseudo-random
 * sample of {block, LDS stack} x waves/CU x refill x queue capacity x tile,
 * narrowed by successive halving (a third survives each round, timed with
 * twice the renders).  The winner is applied to the context.  With a
 * cache_path, results are kept as text lines keyed by (GPU arch, CUs,
 * log2 triangles, log2 pixels, max depth, log2 jobs); a key found there is
 * applied without a search unless FND_AUTOTUNE_FORCE is set. */
typedef struct fnd_autotune_result {
    fnd_tuning tuning;          /* the configuration now applied */
    fnd_tuning median;          /* the first round's median configuration (for held-out checks) */
    double best_ms;             /* its mean device ms per timing render, last round */
    double default_ms;          /* the library default, first round */
    double median_ms;           /* median over the first round */
    uint32_t candidates;        /* configurations timed */
    uint32_t renders;           /* timing renders run */
    uint32_t from_cache;        /* 1: taken from the cache file */
    uint32_t reserved;
    char key[128];              /* cache key of the workload */
} fnd_autotune_result;

#define FND_AUTOTUNE_FORCE 1u

fnd_status fnd_pt_autotune(fnd_pt_ctx* ctx, const fnd_render_job* jobs, uint32_t num_jobs, const char* cache_path,
                           uint32_t flags, fnd_autotune_result* out);

/* ---------------------------------------------------------------- radiance cache */

/* Online-trained radiance cache (no counterpart in the reference).  A
 * multiresolution hash grid (Mueller et al. 2022) feeds a ReLU MLP without
 * biases (Mueller et al. 2021) that predicts the albedo-demodulated indirect
 * radiance leaving a Lambert point, or a GGX point toward the outgoing
 * direction.  The kernels are compiled for exactly: 16 levels x 2 features,
 * 2^19 entries per level, resolutions 16..2048, 5 hidden layers of 64; any
 * other shape returns FND_E_UNSUPPORTED.  Inputs per query: world position,
 * oriented unit normal, outgoing direction and roughness.  Loss: relative
 * L2, sum over samples and RGB of (p - y)^2 / (p^2 + 0.01) with p held
 * constant in the denominator.  Optimiser: Adam (beta1 0.9, beta2 0.99,
 * eps 1e-15) on FP32 master weights; the kernels read FP16 copies. */
typedef struct fnd_cache_desc {
    uint32_t enable;            /* 0 releases the cache */
    uint32_t levels;            /* 16 */
    uint32_t features;          /* 2 */
    uint32_t log2_table;        /* 19 */
    uint32_t base_resolution;   /* 16 */
    uint32_t max_resolution;    /* 2048 */
    uint32_t hidden;            /* 64 */
    uint32_t hidden_layers;     /* 5 */
    float learning_rate;        /* 1e-2 */
    uint32_t train_paths;       /* training paths per frame (FND_JOB_TRAIN_CACHE), <= 2^16: a subset of the
                                   frame's own paths (at most one in four) that goes on to full depth */
    uint32_t train_steps;       /* optimiser steps per frame, 1..64 */
    uint32_t train_depth;       /* cache vertices a training path records, 1..16; it then ends in the
                                   cache at its next cache vertex (self-training) */
    uint32_t train_batch;       /* samples per optimiser step, 256..65536; each step takes an evenly
                                   strided subset of the frame's samples */
    float glossy_min_roughness; /* GGX vertices at least this rough also terminate into (and train) the
                                   cache, with their outgoing direction as input; > 1 disables */
    uint32_t glossy_any_bounce; /* from this bounce on (0 = camera hit) every GGX vertex does so */
    uint64_t seed;              /* initialisation */
    float bounds_lo[3];         /* region the grid covers; lo == hi takes the bounds of the */
    float bounds_hi[3];         /* first scene rendered with the cache */
    uint32_t grid_gradient;     /* how the hash grid's gradient is summed over a batch:
                                   FND_CACHE_GRAD_BF16 (packed BF16 atomics, the default) or
                                   FND_CACHE_GRAD_FP32 (FP32 atomics: twice the atomics, about
                                   0.26 ms more per 1080p training frame on MI355X, and no
                                   measurable change in the gradient's error or in convergence) */
} fnd_cache_desc;

#define FND_CACHE_GRAD_BF16 0u
#define FND_CACHE_GRAD_FP32 1u

fnd_status fnd_pt_cache_default_desc(fnd_cache_desc* out);
fnd_status fnd_pt_cache_configure(fnd_pt_ctx* ctx, const fnd_cache_desc* desc);
/* Parameter counts: MLP weights (plain layout: W0 [48][64], W1..W4
 * [64][64], W5 [64][16], row = input) and grid features ([level][entry][2]). */
fnd_status fnd_pt_cache_param_count(fnd_pt_ctx* ctx, uint64_t* mlp, uint64_t* grid);
fnd_status fnd_pt_cache_set_params(fnd_pt_ctx* ctx, const float* mlp, const float* grid);
fnd_status fnd_pt_cache_get_params(fnd_pt_ctx* ctx, float* mlp, float* grid);
/* queries: n * 10 floats (position, oriented normal, outgoing direction,
 * GGX roughness; direction and roughness 0 for a Lambert point); out: n * 3 */
fnd_status fnd_pt_cache_infer(fnd_pt_ctx* ctx, const float* queries, uint32_t n, float* out);
/* Loss and gradients of the loss for one batch, without an optimiser step.
 * dmlp: fnd_pt_cache_param_count(mlp) floats, dgrid: (grid) floats or NULL. */
fnd_status fnd_pt_cache_gradients(fnd_pt_ctx* ctx, const float* queries, const float* targets, uint32_t n,
                                  float* loss, float* dmlp, float* dgrid);
/* `steps` optimiser steps on one batch; loss of the last step. */
fnd_status fnd_pt_cache_train(fnd_pt_ctx* ctx, const float* queries, const float* targets, uint32_t n,
                              uint32_t steps, float* loss);

/* ---------------------------------------------------------------- host encoding */

/* The host half of fnd_pt_scene_create without a device: validates the
 * description and encodes the acceleration structure exactly as
 * fnd_pt_scene_create uploads it (64-byte wide nodes, 48-byte leaf
 * triangles, 64-byte instance records).  For host-side checks and tools. */
typedef struct fnd_host_scene fnd_host_scene;
typedef struct fnd_host_scene_view {
    const void* nodes;
    uint64_t num_nodes;
    const void* tris;
    uint64_t num_tris;
    const void* instances;
    uint32_t num_instances;
    uint32_t root;
    uint32_t two_level;
    uint32_t reserved;
    fnd_scene_info info;
} fnd_host_scene_view;

fnd_status fnd_scene_encode(const fnd_scene_desc* desc, fnd_host_scene** out, char* err, size_t err_len);
fnd_status fnd_host_scene_get(const fnd_host_scene* scene, fnd_host_scene_view* out);
void fnd_host_scene_free(fnd_host_scene* scene);

/* ---------------------------------------------------------------- scene ingestion (host only) */

/* Geometry as the reference hands it to the GPU: an interleaved vertex
 * buffer whose position attribute is described like a Vulkan vertex input
 * (stride, byte offset, 2 or 3 floats; z = 0 for 2, as in the reference's
 * {vec2 pos, vec3 color, vec2 texCoord}, src/Renderer/Renderer.cpp:23-27),
 * and an index buffer of R16_UINT or R32_UINT (src/Platform/RHI/Vulkan/
 * Command.cpp:289-305). */
typedef struct fnd_vertex_stream {
    const void* data;
    uint64_t bytes;                /* size of data */
    uint32_t stride;               /* bytes from one vertex to the next */
    uint32_t position_offset;      /* byte offset of the position in a vertex */
    uint32_t position_components;  /* 2 or 3 floats */
    uint32_t num_vertices;
} fnd_vertex_stream;

typedef struct fnd_index_stream {
    const void* data;
    uint32_t index_bits;           /* 16 or 32 */
    uint32_t num_indices;          /* a multiple of 3 (triangle list) */
} fnd_index_stream;

/* Writes num_vertices * 3 floats and num_indices u32 indices, the arrays an
 * fnd_mesh_desc points at.  Fails (no partial success) on an attribute past
 * the stride, a buffer too short for num_vertices, or an index out of range. */
fnd_status fnd_mesh_from_streams(const fnd_vertex_stream* vertices, const fnd_index_stream* indices,
                                 float* positions, uint32_t* indices_out, char* err, size_t err_len);

/* Wavefront OBJ: `v x y z` positions and `f` faces (1-based or negative
 * references, `a/b/c` forms accepted, polygons split into fans around their
 * first vertex); other statements are ignored.  The mesh owns its arrays
 * until fnd_obj_free; fnd_obj_mesh_get fills positions, indices and counts
 * of an fnd_mesh_desc (material and light ids NULL). */
typedef struct fnd_obj_mesh fnd_obj_mesh;
fnd_status fnd_obj_load(const char* path, fnd_obj_mesh** out, char* err, size_t err_len);
fnd_status fnd_obj_mesh_get(const fnd_obj_mesh* mesh, fnd_mesh_desc* out);
void fnd_obj_free(fnd_obj_mesh* mesh);

/* ---------------------------------------------------------------- scene generators */

/* Procedural scenes of the benchmark configurations.  Owned host objects;
 * the pointers in the filled fnd_scene_desc stay valid until
 * fnd_gen_scene_free. */
typedef enum fnd_gen_kind {
    FND_GEN_CORNELL_SOUP = 1,        /* Cornell box (5 quads + 2 boxes) + random soup, one quad light */
    FND_GEN_DISPLACED_SPHERE = 2,    /* one displaced icosphere + random soup, single level */
    FND_GEN_INSTANCED_SPHERES = 3,   /* instances of distinct displaced spheres + 4 quad lights, two level */
    FND_GEN_FURNACE = 4              /* closed sphere seen from inside: emitting + reflecting Lambert */
} fnd_gen_kind;

typedef struct fnd_gen_params {
    uint32_t kind;                   /* fnd_gen_kind */
    uint64_t seed;
    uint32_t soup_triangles;         /* random soup triangle count */
    uint32_t subdivision;            /* icosphere face subdivision n: 20*n*n triangles per sphere */
    uint32_t instances;              /* FND_GEN_INSTANCED_SPHERES */
    uint32_t octaves;                /* value-noise octaves */
    float amplitude;                 /* radial displacement amplitude */
    float albedo;                    /* FND_GEN_FURNACE */
    float emission;                  /* FND_GEN_FURNACE */
    uint32_t ggx_fraction_pct;       /* share of GGX materials, percent */
    uint32_t lights;                 /* quad lights */
} fnd_gen_params;

typedef struct fnd_generated_scene fnd_generated_scene;

fnd_status fnd_gen_default_params(uint32_t kind, fnd_gen_params* out);
/* Scene `index` of the multi-scene batch workload: a displaced sphere plus
 * soup whose triangle count is log-uniform in [10k, 2M] (one in eight soup
 * triangles), seed = index, and a square image side drawn from {256, 512,
 * 1024}; both drawn from PCG32 stream `index`. */
fnd_status fnd_gen_batch_params(uint32_t index, fnd_gen_params* out, uint32_t* image_side);
fnd_status fnd_gen_scene(const fnd_gen_params* params, fnd_generated_scene** out);
fnd_status fnd_gen_scene_get(const fnd_generated_scene* scene, fnd_scene_desc* out_desc, fnd_camera* out_camera);
void fnd_gen_scene_free(fnd_generated_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* FND_PT_H */
