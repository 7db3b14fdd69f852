/*
 * crayhip.h - C-ABI of the MI355X rendering hot path (libcrayhip.so).
 *
 * Plain C, plain pointers and sizes.  Every entry point returns an int status
 * (CRH_OK = 0); on failure crh_last_error() returns a thread-local message in
 * the reference's shape ("HIP call <expr> failed with error <code> (<name>).").
 *
 * Ownership: opaque handles are created and destroyed by the caller through
 * the *_create / *_destroy pairs.  Host arrays passed in are only read during
 * the call.  Device memory is owned by the library.  Every call is
 * synchronous at return.  One context per host thread.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the reference's Source/CRaystal/).
 */
#ifndef CRAYHIP_H
#define CRAYHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define CRH_OK 0
#define CRH_ERR_INVALID_ARG 1 /* CRAYSTAL_CHECK failures, Core/Error.h:121-125 */
#define CRH_ERR_HIP 2         /* CRAYSTAL_CUDA_CHECK, Utils/CudaUtils.h:3-11 */
#define CRH_ERR_OOM 3
#define CRH_ERR_UNSUPPORTED 4
#define CRH_ERR_INTERNAL 5

#define CRH_INVALID_PRIM 0xFFFFFFFFu

/* ---- opaque handles --------------------------------------------------- */
typedef struct crh_context crh_context;
typedef struct crh_bvh crh_bvh;
typedef struct crh_scene crh_scene;
typedef struct crh_film crh_film;
typedef struct crh_buffer crh_buffer;
typedef struct crh_guide crh_guide;

/* ---- value types (layouts fixed; static-asserted in the library) ------ */

/* Ray, 32 B.  Same layout as the reference Ray (Math/Ray.h:8-19). */
typedef struct crh_ray {
    float origin[3];
    float tmin;
    float dir[3];
    float tmax;
} crh_ray;

/* Closest hit, 16 B.  prim = mesh primitive ID (HitInfo::primitiveIndex,
 * Core/Hit.h:15-24), t = RayHit::hitT, (u, v) = HitInfo::barycentric.
 * A miss is prim = CRH_INVALID_PRIM, t = +inf.  A sphere hit
 * (HitType::Sphere) is prim = sphere index | CRH_HIT_SPHERE with u = v = 0;
 * its (u, v) parameterisation is needed only for texture lookups, where the
 * integrator derives it from the hit point. */
#define CRH_HIT_SPHERE 0x40000000u
typedef struct crh_hit {
    uint32_t prim;
    float t;
    float u;
    float v;
} crh_hit;

/* BVH build parameters.  Reference constants NUM_BINS = 32,
 * MIN_TRIS_PER_LEAF = 4 (used as count <= 4), MAX_DEPTH = 20
 * (Core/BVH.h:104-106).  Zero fields take the reference value. */
typedef struct crh_bvh_params {
    uint32_t bins;
    uint32_t max_leaf;
    uint32_t max_depth;
    /* 1: build on the GPU (crh_scene_create; the same nodes and primitive
     * order bit for bit, for bins <= 64 and max_depth <= 64); 0: on the host. */
    uint32_t device;
} crh_bvh_params;

/* Canonical export of a reference BVHNode (Core/BVH.h:17-30): tag 0 = leaf
 * {a = leafOffset, b = leafCount}, tag 1 = internal {a = left, b = right};
 * nodes are in the reference's post-order, root last. 36 B. */
typedef struct crh_bvh_node_ref {
    uint32_t tag;
    uint32_t a;
    uint32_t b;
    float pmin[3];
    float pmax[3];
} crh_bvh_node_ref;

/* BVH build statistics (Core/BVH.h:64-69 BuildStats plus root data). */
typedef struct crh_bvh_stats {
    uint32_t node_count;
    uint32_t max_depth;      /* deepest internal node, as the reference logs */
    uint32_t max_leaf;       /* largest leaf primitive count */
    uint32_t root;           /* root index in the post-order node array */
    uint32_t prim_count;
    uint32_t internal_count;
    float root_min[3];
    float root_max[3];
    double build_ms;
} crh_bvh_stats;

/* Material types (Core/Material.h:12-16). */
#define CRH_MAT_DIFFUSE 0u
#define CRH_MAT_PRINCIPLED 1u
#define CRH_MAT_PHONG 2u
/* Material flags (Core/Material.h:26-30). */
#define CRH_MATF_EMISSIVE 0x1u
#define CRH_MATF_DIFFUSE 0x2u
#define CRH_MATF_DIFFUSE_TEXTURE 0x4u

/* Material, 68 B: the reference MaterialData with the RGB Spectrum
 * (Core/Material.h:34-66).  flags are taken as given; MaterialData::finalize
 * semantics are provided by crh_material_finalize. */
typedef struct crh_material {
    float diffuse[3];
    uint32_t diffuse_tex;
    float specular[3];
    float transmittance[3];
    float emission[3];
    float shininess;
    float ior;
    uint32_t type;
    uint32_t flags;
} crh_material;

/* A texture image (Image as Texture(const Image&) reads it,
 * Core/Texture.cu:46-62): width x height pixels of `channels` (1..4) floats,
 * pixel (x, y) at data[(y * width + x) * channels]; missing channels read 0
 * (a one-channel image is red only, as in the reference). */
typedef struct crh_texture {
    uint32_t width;
    uint32_t height;
    uint32_t channels;
    const float* data;
} crh_texture;

/* One mesh (MeshData, Scene/Triangle.h:36-42).  Indices are local to the
 * mesh; meshes are concatenated in order and the primitive ID is the global
 * triangle index in that order (TriangleMeshManager, Scene/Triangle.cu:98-165).
 * normal must hold nvert*3 floats; texcrd may be NULL (zeros). */
typedef struct crh_mesh {
    const float* position;
    const float* normal;
    const float* texcrd;
    uint32_t nvert;
    const uint32_t* index;
    uint32_t nindex;
    uint32_t material_id;
} crh_mesh;

/* SceneData (Scene/Scene.h:40-47), plus a constant environment radiance (an
 * addition; 0 = the reference's black miss). */
typedef struct crh_scene_desc {
    const crh_mesh* meshes;
    uint32_t nmesh;
    const crh_material* materials;
    uint32_t nmaterial;
    const uint32_t* emissive_index;
    uint32_t nemissive;
    float env_radiance[3];
    crh_bvh_params bvh;
    /* Spectral transport (optional, no counterpart in the reference's RGB
     * integrator): nmaterial x 3 x 95 floats (diffuse, specular, emission
     * spectra of each material, bins of 5 nm from 360 nm to 830 nm) and the
     * environment's 95-bin spectrum (may be NULL = 0).  NULL spectra: the
     * materials' and the environment's RGB values are uplifted to spectra by
     * crh_rgb_to_spectrum, so every scene can also render spectrally. */
    const float* spectra;
    const float* env_spectrum;
    /* Spheres (SceneData::spheres, Scene/Scene.h:40-47; SphereData,
     * Scene/Sphere.h): centre (3 floats) and radius of each, tested by every
     * ray before the triangle BVH as SceneView::intersect does
     * (Scene/Scene.cu:5-29).  The reference gives a sphere no material: it
     * takes the material of the mesh holding the triangle with the same
     * index (Core/Integrator.cu:172-178 reads getMeshDesc(primitiveIndex)),
     * so with sphere_material = NULL every sphere index must be below the
     * triangle count, and a sphere whose index equals the sampled emissive
     * triangle's counts as that light in the MIS BSDF test
     * (Core/Integrator.cu:108).  sphere_material (optional, nsphere ids)
     * gives each sphere its own material instead, and MIS then matches
     * triangles only. */
    const float* sphere_center;
    const float* sphere_radius;
    uint32_t nsphere;
    const uint32_t* sphere_material;
    /* Diffuse textures (SceneData::textureImages), indexed by
     * crh_material.diffuse_tex when its flags hold CRH_MATF_DIFFUSE_TEXTURE;
     * sampled bilinearly at the hit's texture coordinate
     * (TextureView::sample, Core/Texture.cu:15-35; getBSDF,
     * Material/BSDF.cu:11-15). */
    const crh_texture* textures;
    uint32_t ntexture;
} crh_scene_desc;

/* Camera: the reference CameraProxy layout, 96 B (Core/CameraProxy.h:10-24).
 * u/v/w are filled by crh_camera_update (Camera::calculateCameraData,
 * Core/Camera.cu:35-39). */
typedef struct crh_camera {
    float pos[3];
    float fov_y;
    float target[3];
    float near_z;
    float up[3];
    float far_z;
    float u[3];
    uint32_t pad0;
    float v[3];
    uint32_t pad1;
    float w[3];
    uint32_t pad2;
} crh_camera;

/* Integrator modes. */
#define CRH_MODE_MEGAKERNEL 0u
#define CRH_MODE_WAVEFRONT 1u

/* PathTraceIntegrator::Configs (Core/Integrator.h:14-18), plumbed to the
 * device.  Russian roulette applies when depth >= rr_start_depth; the
 * reference's hard-wired behaviour is max_depth 10, rr_threshold 0.05,
 * rr_start_depth 6 (Core/Integrator.cu:209-223 tests depth > 5). */
typedef struct crh_integrator_cfg {
    uint32_t max_depth;
    float rr_threshold;
    uint32_t rr_start_depth;
    uint32_t use_mis;
    uint32_t mode;
    uint32_t count_work; /* != 0: run the counting variant (fills
                            node_visits / tri_tests; slower) */
    uint32_t spectral;   /* != 0: spectral transport (megakernel or wavefront;
                            scenes with spectra, no counting variant): each
                            path carries three of the 95 wavelength bins and
                            the film receives CIE XYZ */
    uint32_t reserved;
} crh_integrator_cfg;

/* Timing/counters returned by the render and trace entry points. */
typedef struct crh_stats {
    double kernel_ms;      /* GPU time of the whole call (HIP events) */
    double main_kernel_ms; /* summed GPU time of the dominant kernel's
                              launches (per-launch HIP events) */
    double total_ms;       /* host wall time of the call */
    uint64_t rays;      /* rays traced, if counted (0 otherwise) */
    uint64_t node_visits;
    uint64_t tri_tests;
    uint64_t launches;
} crh_stats;

/* Launch choices that never change results (picked by the autotuner,
 * craystal_amd/autotune.py).  Zero fields keep the current value. */
typedef struct crh_tuning {
    uint32_t stack;          /* ray batches: traversal stack 1 = LDS, 2 = private (scratch),
                                3 = 8 levels in LDS + the rest private */
    uint32_t trace_block;    /* ray batches: threads per block, 64..256 */
    uint32_t persistent;     /* ray batches: 1 = lanes refill from a work counter, 2 = off */
    uint32_t refill;         /* ray batches: idle lanes per wave that trigger a refill, 1..64 */
    uint32_t pt_grid;        /* ray batches: blocks of the persistent kernels */
    uint32_t render_stack;   /* megakernel: stack as above */
    uint32_t render_pool;    /* megakernel: paths per launch; sample indices per launch =
                                max(1, render_pool / pixels), capped by the spp asked */
    uint32_t render_waves;   /* megakernels: 1 = compiler's register budget, 4 = at most
                                128 VGPRs so 4 waves fit per SIMD, 5 = at most 96 VGPRs,
                                6 = at most 80, 8 = at most 64 (more spills) */
    uint32_t render_regen;   /* megakernel: 1 = persistent blocks whose lanes start a new
                                path as soon as theirs ends, 2 = one path per thread */
    uint32_t render_grid;    /* megakernel with regeneration: persistent blocks */
    uint32_t wf_finish;      /* wavefront: 1..64 = after that many bounces the paths still in
                                flight finish in one persistent launch (a lane per path,
                                refilled as paths end); 65 = every bounce as wavefront
                                launches.  Same film either way.  Guided rendering never
                                finishes before crh_guided_cfg.max_depth (off when every
                                depth is guided) */
    uint32_t wf_stack;       /* wavefront traversal kernels: stack as above */
    uint32_t wf_trace_grid;  /* wavefront traversal kernels: blocks */
    uint32_t wf_trace_block; /* wavefront traversal kernels: threads (64..256) */
    uint32_t wf_shade_grid;  /* wavefront shading kernels: blocks of 256 */
    uint32_t wf_persistent;  /* wavefront traversal: 1 = lanes refill, 2 = off */
    uint32_t wf_refill;      /* wavefront traversal: idle lanes per wave that trigger a refill */
    uint32_t wf_pool;        /* wavefront: paths kept in flight; samples per pixel rendered
                                together = max(1, wf_pool / pixels), capped by the spp asked */
    uint32_t batch_stack;    /* batched rendering: stack as above */
    uint32_t batch_tile_w;   /* batched rendering: tile width 8, 16, 32 or 64 (256 px tiles) */
    uint32_t batch_grid;     /* batched rendering: persistent blocks per pass */
    uint32_t batch_chunks;   /* batched rendering: work ranges with own counters, 1..8 */
    uint32_t batch_graph;    /* batched rendering: 1 = passes replayed from a hipGraph, 2 = off */
    uint32_t batch_waves;    /* batch megakernel register budget: 1, 4, 5, 6 or 8 waves per SIMD */
    uint32_t render_band;    /* megakernel (one path per thread): 1 = the blocks of each XCD
                                take a contiguous band of tiles (its own L2), 2 = tiles in
                                block order */
    uint32_t batch_slots;    /* batched rendering: sample indices per launch, 1..64 (while the
                                set of scenes still sampling stays the same) */
    uint32_t trace_tris;     /* ray batches, persistent: triangles of a leaf tested per
                                traversal step, 1..4 */
    uint32_t guided_fused;   /* guided rendering: 1 = the fused megakernel (the guide's
                                inference inside the path kernel), 2 = the wavefront.  Same
                                film for the same guide */
    uint32_t guided_stack;   /* fused guided megakernel: stack as above */
    uint32_t guided_waves;   /* fused guided megakernel: register budget as render_waves */
    uint32_t guided_grid;    /* fused guided megakernel: persistent blocks */
    uint32_t guided_pool;    /* fused guided megakernel: paths per launch, as render_pool */
} crh_tuning;

/* ---- errors / device ------------------------------------------------- */
const char* crh_last_error(void);
int crh_device_count(int* count);
int crh_version(uint32_t* version);

/* ---- context (replaces the implicit CUDA runtime context) ------------- */
int crh_context_create(int device, crh_context** out);
/* Drops the caller's reference.  Buffers, scenes, films and guides created
 * on the context hold references of their own, so objects may be destroyed
 * in any order; the stream is released with the last of them. */
int crh_context_destroy(crh_context* ctx);
/* Launch tuning for every later call on this context (no counterpart in the
 * reference, whose launch shape is fixed, Core/Integrator.cu:248-253). */
int crh_context_set_tuning(crh_context* ctx, const crh_tuning* tuning);
int crh_context_get_tuning(const crh_context* ctx, crh_tuning* tuning);

/* ---- device buffers (DeviceBuffer, Core/Buffer.h:13-44; size_t sizes) - */
int crh_buffer_create(crh_context* ctx, size_t bytes, crh_buffer** out);
int crh_buffer_upload(crh_buffer* buf, const void* src, size_t bytes);
int crh_buffer_download(const crh_buffer* buf, void* dst, size_t bytes);
int crh_buffer_memset(crh_buffer* buf, int value);
/* Device-to-device copy of bytes from src to dst (no reference counterpart;
 * also the known-size copy used to calibrate HBM traffic counters). */
int crh_buffer_copy(crh_buffer* dst, const crh_buffer* src, size_t bytes);
int crh_buffer_size(const crh_buffer* buf, size_t* bytes);
int crh_buffer_destroy(crh_buffer* buf);

/* ---- BVH (BVH::build, Core/BVH.cu:88-338) ------------------------------ */
/* Host binned-SAH build over one triangle soup; bit-exact with the
 * reference's node array and primitive order for the same parameters. */
int crh_bvh_build(const float* pos, uint32_t nvert, const uint32_t* idx,
                  uint32_t ntri, const crh_bvh_params* params, crh_bvh** out);
/* BVH::build on the GPU of ctx's device (no counterpart in the reference,
 * whose builder is host code): the result of crh_bvh_build bit for bit (node
 * array, primitive order, statistics) for bins <= 64 and max_depth <= 64.
 * Host arrays in; build_ms is the call's wall time (uploads included). */
int crh_bvh_build_device(crh_context* ctx, const float* pos, uint32_t nvert,
                         const uint32_t* idx, uint32_t ntri,
                         const crh_bvh_params* params, crh_bvh** out);
int crh_bvh_get_stats(const crh_bvh* bvh, crh_bvh_stats* stats);
/* nodes: node_count entries; prim_order: prim_count entries
 * (bvhToMeshPrimitiveIndex, Core/BVH.cu:322-338). */
int crh_bvh_export(const crh_bvh* bvh, crh_bvh_node_ref* nodes,
                   uint32_t* prim_order);
int crh_bvh_destroy(crh_bvh* bvh);

/* ---- materials --------------------------------------------------------- */
/* MaterialData::finalize (Core/Material.h:57-65). */
int crh_material_finalize(crh_material* mat);

/* ---- scene (Scene::Scene, Scene/Scene.cu:61-78) ----------------------- */
int crh_scene_create(crh_context* ctx, const crh_scene_desc* desc,
                     crh_scene** out);
int crh_scene_get_bvh_stats(const crh_scene* scene, crh_bvh_stats* stats);
int crh_scene_export_bvh(const crh_scene* scene, crh_bvh_node_ref* nodes,
                         uint32_t* prim_order);
int crh_scene_destroy(crh_scene* scene);

/* ---- camera ------------------------------------------------------------ */
/* Camera::calculateCameraData (Core/Camera.cu:35-39). */
int crh_camera_update(crh_camera* cam);

/* ---- ray queries (SceneView::intersect / intersectOcclusion,
 *      Scene/Scene.cu:5-29; BVHData::intersect, Core/BVH.cu:13-84) ------ */
/* Host arrays in and out (includes PCIe copies). */
int crh_trace_closest(crh_scene* scene, const crh_ray* rays, uint64_t n,
                      crh_hit* hits);
int crh_trace_any(crh_scene* scene, const crh_ray* rays, uint64_t n,
                  uint8_t* occluded);
/* Device-resident variants: rays/hits are crh_buffers already in HBM.
 * stats (optional) receives the kernel time; count_work != 0 runs the
 * counting variant and fills node_visits / tri_tests. */
int crh_trace_closest_device(crh_scene* scene, const crh_buffer* rays,
                             uint64_t n, crh_buffer* hits, int count_work,
                             crh_stats* stats);
int crh_trace_any_device(crh_scene* scene, const crh_buffer* rays, uint64_t n,
                         crh_buffer* occluded, crh_stats* stats);
/* No counterpart in the reference.  crh_trace_closest_device for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments and results, but it calls
 * no hipMalloc / hipFree and no allocating library routine.  Its scratch (the
 * persistent kernel's counter, the per-ray work counts and their totals) is
 * one fixed device array of 8 MiB + 256 B per GPU, allocated when the library
 * loads and held until the process ends; the work counts of longer batches
 * are taken in chunks of 2^20 rays, so n is limited only as in
 * crh_trace_closest_device (n < 2^40).  Every call on a GPU shares that
 * array: calls that overlap on one GPU (two host threads, or two contexts on
 * one device) race. */
int crh_trace_closest_device_noalloc(crh_scene* scene, const crh_buffer* rays,
                                     uint64_t n, crh_buffer* hits, int count_work,
                                     crh_stats* stats);

/* ---- film (Sensor / SensorData, Core/Sensor.{h,cu}) -------------------- */
int crh_film_create(crh_context* ctx, uint32_t width, uint32_t height,
                    crh_film** out);
int crh_film_clear(crh_film* film);
/* sum: width*height*3 floats (Spectrum sums), weight: width*height floats.
 * Either may be NULL. */
int crh_film_readback(const crh_film* film, float* sum, float* weight);
int crh_film_destroy(crh_film* film);

/* ---- render (PathTraceIntegrator::dispatch, Core/Integrator.cu:233-262) */
/* Renders sample indices [first_sample, first_sample + spp) into film. */
int crh_render(crh_scene* scene, const crh_camera* cam,
               const crh_integrator_cfg* cfg, uint32_t first_sample,
               uint32_t spp, crh_film* film, crh_stats* stats);

/* ---- learned path guiding (no counterpart in the reference) ------------
 * An MLP of hidden width 64 (or 128; 4 hidden ReLU layers, BF16 weights on
 * MFMA, FP32 accumulation) maps (position, shading normal, outgoing
 * direction) to an 8-lobe von Mises-Fisher mixture; trained online with
 * Adam on the KL divergence to the renderer's radiance samples.  Parameter
 * layout: for each of the 5 layers, W (K x N row-major, K = 64, w, w, w, w,
 * N = w, w, w, w, 48 for width w) then b (N); CRH_GUIDE_PARAMS floats at
 * width 64, CRH_GUIDE_PARAMS_128 at width 128. */
#define CRH_GUIDE_PARAMS 19760u
#define CRH_GUIDE_PARAMS_128 64048u
#define CRH_GUIDE_OUTPUTS 48u
/* scene bounds normalise the position encoding; seed picks the He init;
 * width 64 */
int crh_guide_create(crh_context* ctx, const float* bmin, const float* bmax,
                     uint32_t seed, crh_guide** out);
/* the same with hidden width 64 or 128 */
int crh_guide_create_width(crh_context* ctx, const float* bmin, const float* bmax,
                           uint32_t seed, uint32_t width, crh_guide** out);
int crh_guide_set_params(crh_guide* g, const float* theta, uint32_t n);
int crh_guide_get_params(const crh_guide* g, float* theta, uint32_t n);
/* n queries (pos, nrm, wo: 3 floats each).  raw: n*48 network outputs;
 * mixture: n*40 = weights[8], means[8][3], kappas[8].  Either may be NULL. */
int crh_guide_eval(crh_guide* g, const float* pos, const float* nrm,
                   const float* wo, uint32_t n, float* raw, float* mixture);
/* One training step on n samples of 16 floats (pos, _, nrm, _, wo, _, wi,
 * target/pdf).  apply_adam = 0 leaves the parameters and returns the
 * gradient in grad (CRH_GUIDE_PARAMS floats, may be NULL); loss may be NULL. */
int crh_guide_train(crh_guide* g, const float* samples, uint32_t n, float lr,
                    int apply_adam, float* loss, float* grad);
/* Training progress: Adam steps taken, training samples the last trained
 * render pass produced, loss of the last step. */
typedef struct crh_guide_info {
    uint32_t steps;
    uint32_t last_samples;
    float last_loss;
    uint32_t width;  /* hidden width of the network: 64 or 128 */
} crh_guide_info;
int crh_guide_get_info(const crh_guide* g, crh_guide_info* info);
/* Losses of the last n Adam steps, oldest first; n <= min(steps, 1024). */
int crh_guide_get_losses(const crh_guide* g, float* losses, uint32_t n);
int crh_guide_destroy(crh_guide* g);

/* Guided rendering: the path tracer with one-sample MIS between BSDF
 * sampling (fraction alpha) and the guide's mixture at every Lambertian /
 * Principled vertex; the estimator's expectation equals the plain
 * integrator's.  It runs on the fused guided megakernel (the guide evaluated
 * on MFMA inside the path kernel; crh_tuning.guided_*) or, with
 * guided_fused = 2, on the wavefront (crh_tuning.wf_*); both give the same
 * film for the same guide.  Passes run in groups of samples per pixel kept
 * in flight together (guided_pool or wf_pool).  With train != 0, the vertices with
 * non-zero radiance become training samples; they accumulate on the film over
 * groups (and calls) and each group of k passes is followed by up to
 * k * train_steps Adam steps, one per full batch of train_batch samples. */
typedef struct crh_guided_cfg {
    float alpha;
    uint32_t train;
    uint32_t train_batch;
    uint32_t train_steps;
    float lr;
    uint32_t max_depth;  /* guide only vertices at depth < max_depth (0 = every depth) */
    uint32_t materials;  /* guided BSDFs: bit 0 Lambertian, bit 1 Principled (0 = both);
                            the others sample their BSDF as the plain integrator does */
    uint32_t target;     /* what the guide learns: 0 = BSDF x cos x incident radiance,
                            1 = incident radiance */
} crh_guided_cfg;
/* stats->node_visits receives the training samples the last pass group
 * produced; stats->launches the number of pass groups. */
int crh_render_guided(crh_scene* scene, const crh_camera* cam,
                      const crh_integrator_cfg* cfg, crh_guide* guide,
                      const crh_guided_cfg* gcfg, uint32_t first_sample,
                      uint32_t spp, crh_film* film, crh_stats* stats);
/* Guided rendering of nviews camera views of one scene with one guide, all
 * views in each launch of the fused guided megakernel (views[i] into
 * films[i], spp samples each); training, when on, uses every view's
 * vertices.  Each film is bit-identical to crh_render_guided of that view
 * alone with the same guide when the guide does not train.  No counterpart
 * in the reference, which renders one camera per dispatch
 * (Core/Integrator.cu:233-262). */
int crh_render_guided_views(crh_scene* scene, const crh_camera* cams, crh_film* const* films,
                            uint32_t nviews, const crh_integrator_cfg* cfg, crh_guide* guide,
                            const crh_guided_cfg* gcfg, uint32_t first_sample, uint32_t spp,
                            crh_stats* stats);

/* ---- spectra (the reference's integrator transports RGB; its 95-bin
 * BroadSpectrum and CIE table, Core/Spectrum.h:278-324, are the analogue) -- */
/* The colour matching functions of the 95 bins: the reference's kXYZTable
 * (Core/SpectrumTabData.h:7-55) as x, y, z-bar per bin, 285 floats. */
int crh_cie_cmf(float* xyz);
/* XYZ of a 95-bin spectrum: BroadSpectrum::toXYZ (Core/Spectrum.cpp:37-45),
 * sum of table * s times kXYZStep * kCIENormalizationY, with z-bar for Z
 * (the reference reads x-bar there, Core/Spectrum.cpp:40). */
int crh_spectrum_to_xyz(const float* spectrum, float* xyz);
/* RGB -> 95-bin spectrum, replacing BroadSpectrum::fromRGB
 * (Core/Spectrum.cpp:49, which returns zeros; its Smits tables are
 * Core/SpectrumTabData.h:67-257): Smits's decomposition into the flat white
 * spectrum plus one secondary and one primary basis spectrum, the basis
 * recomputed for this library's colour matching functions so that the
 * white-balanced linear sRGB of the result (flat 1 -> (1, 1, 1)) equals the
 * input.  Reflectances in [0, 1] map into [0, 1]; homogeneous of degree 1, so
 * emission uses it too.  Negative channels count as 0.  rgb: 3 floats;
 * spectrum: 95 floats out. */
int crh_rgb_to_spectrum(const float* rgb, float* spectrum);

/* ---- batched rendering (no counterpart in the reference, which renders
 * one scene per PathTraceIntegrator::dispatch, Core/Integrator.cu:233-262) - */
typedef struct crh_batch crh_batch;
typedef struct crh_batch_info {
    uint32_t scenes;
    uint32_t max_stack;   /* deepest BVH, levels */
    uint64_t pool_bytes;  /* the concatenated scene data */
    uint64_t nodes;       /* internal BVH nodes over all scenes */
    uint64_t triangles;
} crh_batch_info;
/* Concatenates the scenes' device data (BVH nodes, triangles, vertices,
 * materials, emitter lists) into one pool; each scene's BVH keeps its own
 * indices, relative to its slice.  A scene listed more than once (several
 * camera views of one scene) is stored once and its entries share it; the
 * info counts it once.  The scenes may be destroyed afterwards. */
int crh_batch_create(crh_context* ctx, crh_scene* const* scenes, uint32_t n, crh_batch** out);
int crh_batch_get_info(const crh_batch* b, crh_batch_info* info);
/* Renders spp[i] samples (indices first_sample, first_sample + 1, ...) of
 * scene i through cams[i] into films[i] (distinct films), for every i, with
 * the megakernel integrator: one persistent launch over the 256-pixel tiles of
 * every scene that still takes samples, for up to crh_tuning.batch_slots
 * sample indices at a time (while that set of scenes stays the same), then
 * one film update per sample index over all of them.  With crh_tuning.batch_graph the pass sequence is
 * captured once into a hipGraph and replayed while pass counts, image sizes
 * and tuning stay the same.  Every film ends bit-identical to crh_render of
 * its scene alone with the same arguments. */
int crh_batch_render(crh_batch* b, const crh_camera* cams, crh_film* const* films,
                     const uint32_t* spp, const crh_integrator_cfg* cfg,
                     uint32_t first_sample, crh_stats* stats);
int crh_batch_destroy(crh_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* CRAYHIP_H */
