/*
 * omv.h -- C-ABI of the MI355X BDPT render core.
 *
 * The entry points replace the three C++ class interfaces the reference
 * pathtracer's render loop sits behind (the reference has no plugin or FFI
 * API of its own):
 *
 *   RenderScene   pathtracer/src/scene/RenderScene.h:109-204
 *                 (allocate, setMesh, setInstance, setMaterial, setLight,
 *                  setTexture, setBackground, setFalseColorSubject, commitScene)
 *   BDPTRenderer  pathtracer/src/renderer/BDPTRenderer.h:147-155
 *                 (setParameters, setScene, render)
 *   PinholeCamera pathtracer/src/sensors/PinholeCamera.h + RenderSensor.h:63-96
 *                 (ctor(w,h), setViewPoint, setYFoV, adjustToResolution,
 *                  setRealHeight, getImpression, init/stop)
 *
 * Conventions (all entry points):
 *   - plain C types and pointers only; host buffers are copied at "set" time
 *     (the reference copies mesh arrays and textures in too,
 *     RenderMesh.cpp:109-142, BitmapTexture.cpp:31-37);
 *   - every function returns an int status: OMV_OK (0) or a negative
 *     OMV_E_* code; omv_last_error() returns a thread-local message.
 *     Out-of-range ids are reported as OMV_E_RANGE instead of being silently
 *     ignored as the reference setters do (RenderScene.cpp:88,103,113,119,125);
 *     the scene is left unchanged in that case, which is the reference's
 *     observable behaviour;
 *   - calls into one scene must not overlap; distinct scenes/renderers may be
 *     used from distinct threads.
 */
#ifndef OMV_H
#define OMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define OMV_OK          0
#define OMV_E_ARG      -1   /* null pointer / bad argument                 */
#define OMV_E_RANGE    -2   /* id out of the allocated range               */
#define OMV_E_STATE    -3   /* call not valid in the current state         */
#define OMV_E_DEVICE   -4   /* HIP runtime error                           */
#define OMV_E_SCENE    -5   /* commit failed: nothing renderable           */
#define OMV_E_NOMEM    -6
#define OMV_E_INTERNAL -7
#define OMV_E_IO       -8   /* file could not be opened or written         */

/* ------------------------------------------------------- parameter POD */
/* Estimator extensions.  0 = the reference estimator exactly
 * (BDPTRenderer.cpp:237-301).  Every bit below is a separately named mode
 * that changes the estimator; none is enabled implicitly. */
#define OMV_EXT_NONE        0u
#define OMV_EXT_OM_ROUGH    1u   /* OpenMATERIAL complex-IOR Fresnel used as the
                                    F term of a GGX rough conductor with
                                    roughness params.om_roughness, instead of the
                                    reference's ideal mirror whose connection
                                    gate makes point-lit OM scenes black
                                    (OpenMaterial.cpp:105-109)               */
#define OMV_EXT_HDRI_NEE    2u   /* HDR background importance sampling: at every
                                    opaque PBR camera vertex one direction drawn
                                    from the environment's alias tables, with a
                                    visibility ray, combined with the escape
                                    term by the balance heuristic against the
                                    reference's own BSDF sampling density; same
                                    expected image as the reference estimator  */
#define OMV_EXT_LIGHT_SPLAT 4u   /* light-tracer (t = 1) splats: every stored
                                    light-subpath vertex is connected to the
                                    pinhole and its particle contribution
                                    added, with a visibility ray, to the pixel
                                    it projects to (float atomics); the film is
                                    the reference estimate plus the splat image
                                    over all light paths (W*H*spp); scenes
                                    without point lights are unaffected       */

/* Mirrors RendererParameters (pathtracer/src/renderer/RendererParameters.h:24-96).
 * omv_renderer_params_default() fills the reference defaults. */
typedef struct omv_renderer_params {
    int   camera_bounces;              /* cameraBouncesN   (-B)  default 10   */
    int   light_bounces;               /* lightBouncesN    (-b)  default 10   */
    int   max_path_length;             /* maxPathLength    (-q)  default 8    */
    int   samples_per_pixel;           /* samplesPerPixel  (-S)  default 20   */
    int   light_distance_attenuation;  /* (-A) 0,1,2             default 1    */
    float light_scale_value;           /* default 100                         */
    float light_min_distance;          /* default 0.01                        */
    float ray_cut_pix_value;           /* default 0.002                       */
    float gamma;                       /* post pass only; default 0.5         */
    int   number_of_cores;             /* CPU threads (oracle only) def. 16   */
    int   false_color_mode;            /* falseColorMode                      */
    uint32_t seed;                     /* Philox4x32-10 key word 0            */
    uint32_t extensions;               /* OMV_EXT_* bit set                   */
    float om_roughness;                /* OMV_EXT_OM_ROUGH roughness (alpha=r^2)*/
} omv_renderer_params;

/* Pinhole camera state (PinholeCamera.h members + ViewPoint, RenderSensor.cpp:19-25).
 * rotation is the row-major 3x3 whose columns are (left, up, forward). */
typedef struct omv_camera_desc {
    int   width, height;
    float position[3];
    float rotation[9];
    float half_w, half_h, focus;
} omv_camera_desc;

/* glTF metallic-roughness material: every PbrMaterial getter the render loop
 * reads per vertex (pathtracer/src/assets/materials/PbrMaterial.cpp:224-321). */
#define OMV_PBR_BASE_COLOR          0
#define OMV_PBR_METALLIC_ROUGHNESS  1
#define OMV_PBR_NORMAL              2
#define OMV_PBR_EMISSIVITY          3
#define OMV_PBR_OCCLUSION           4
#define OMV_PBR_TEXTURE_SLOTS       5

#define OMV_ALPHA_BLEND   0
#define OMV_ALPHA_MASK    1
#define OMV_ALPHA_OPAQUE  2

typedef struct omv_pbr_desc {
    float base_color_factor[4];      /* clamped to [0,1] by the loader        */
    float emissive_factor[3];
    float metallic_factor;
    float roughness_factor;
    float normal_scale;              /* default 1                             */
    float occlusion_strength;        /* default 1                             */
    float ior;                       /* default 1.5                           */
    float alpha_cutoff;              /* default 0.5                           */
    int   alpha_mode;                /* OMV_ALPHA_*, default OPAQUE           */
    int   double_sided;
    int   texture_map[OMV_PBR_TEXTURE_SLOTS];     /* texture id or -1          */
    int   texture_channel[OMV_PBR_TEXTURE_SLOTS]; /* uv channel, default 0     */
    float texture_transform[OMV_PBR_TEXTURE_SLOTS][9]; /* KHR_texture_transform
                                        3x3 row-major, identity by default    */
} omv_pbr_desc;

/* One OpenMATERIAL IOR table at one temperature
 * (AssetMaterialIor::loadPropertiesFromJson, AssetMaterialIor.cpp:56-112). */
typedef struct omv_ior_table {
    float temperature;
    int   n_count;
    const float* n_wavelength;   /* metres */
    const float* n_value;
    int   k_count;
    const float* k_wavelength;
    const float* k_value;
} omv_ior_table;

/* False-colour subjects (FCSubject.h:189-482, names as in main.cpp:177). */
#define OMV_FC_NONE   0
#define OMV_FC_PID    1   /* "pid" primitive id          */
#define OMV_FC_GID    2   /* "gid" geometry (mesh) id    */
#define OMV_FC_MID    3   /* "mid" material id           */
#define OMV_FC_RMP    4   /* "rmp" roughness map         */
#define OMV_FC_MMP    5   /* "mmp" metallic map          */
#define OMV_FC_MD     6   /* "md"  mesh density          */
#define OMV_FC_SG     7   /* "sg"  surface gradient      */
#define OMV_FC_MN     8   /* "mn"  material name         */
#define OMV_FC_IN     9   /* "in"  inverted normals      */
/* Instance-aware subjects (no reference counterpart: the reference's ids
 * are per mesh, so the instances of one mesh share them,
 * BDPTRenderer.cpp:64): */
#define OMV_FC_IID   10   /* "iid"  instance id                            */
#define OMV_FC_IPID  11   /* "ipid" primitive id counted over instances:
                             the instance's first face + the face index    */

/* Render statistics. */
typedef struct omv_render_stats {
    uint64_t samples;          /* W*H*spp paths                                */
    uint64_t rays;             /* extension + connection-visibility + mask
                                  re-trace rays actually cast                   */
    uint64_t camera_rays;
    uint64_t light_rays;
    uint64_t shadow_rays;
    uint64_t nodes_visited;    /* BVH nodes fetched (all rays)                 */
    uint64_t tris_tested;      /* triangle records tested (all rays)           */
    uint64_t stack_overflows;  /* must be 0; nonzero flags a traversal error   */
    double   render_ms;        /* device time of the render call (HIP events)  */
    double   kernel_ms;        /* device time of the path kernel(s) only       */
    double   trace_ms;         /* wavefront schedule: summed device time of the
                                  traversal-kernel launches (HIP events)       */
    uint64_t trace_launches;   /* traversal-kernel launches timed in trace_ms  */
} omv_render_stats;

/* ------------------------------------------------------------- errors */
const char* omv_last_error(void);
/* Device helpers (no reference counterpart): wait for all work on the
 * current HIP device; report its name, compute-unit count and memory. */
int omv_device_synchronize(void);
int omv_device_info(char* name, size_t name_cap, int* compute_units, size_t* total_mem);
const char* omv_version(void);
void omv_renderer_params_default(omv_renderer_params* p);

/* ---------------------------------------------------------- scene (RenderScene) */
typedef struct omv_scene omv_scene;

int omv_scene_create(omv_scene** out);
int omv_scene_destroy(omv_scene* s);
/* RenderScene::allocate (RenderScene.cpp:51-80); one extra material slot
 * (index materialsN) holds the "missing material" diffuse (1000,0,1000). */
int omv_scene_allocate(omv_scene* s, int meshes, int instances, int materials,
                       int textures, int lights);
/* RenderScene::setMesh (RenderScene.cpp:82-99).  Strides are in floats.
 * tex[i] may be NULL; texN <= 8 channels.  faces: 3*facesN uint32. */
int omv_scene_set_mesh(omv_scene* s, int mesh_id, int vert_n, int faces_n, int tex_n,
                       int mat_id, const uint32_t* faces,
                       const float* vert, int vstride,
                       const float* norm, int nstride,
                       const float* tans, int tanstride,
                       const float* const* tex, int texstride);
/* RenderScene::setInstance (RenderScene.cpp:101-109); trans is a row-major 4x4,
 * the top 3x4 is used. */
int omv_scene_set_instance(omv_scene* s, int inst_id, const float trans[16], int mesh_id);
/* RenderScene::setMaterial with an OpenMaterial (OpenMaterial.h); the IOR is
 * looked up per RGB wavelength exactly as AssetMaterialIor::getIor
 * (AssetMaterialIor.cpp:115-178) and baked at commit (temperature and the
 * three wavelengths are constants). */
int omv_scene_set_material_om(omv_scene* s, int mat_id, float asset_temperature,
                              int n_tables, const omv_ior_table* tables);
/* OpenMaterial with per-RGB complex IOR given directly (n + i k). */
int omv_scene_set_material_om_nk(omv_scene* s, int mat_id, const float n_rgb[3],
                                 const float k_rgb[3]);
/* RenderScene::setMaterial with a PhysicallyBasedMaterial. */
int omv_scene_set_material_pbr(omv_scene* s, int mat_id, const omv_pbr_desc* d);
/* RenderScene::setTexture (RenderScene.cpp:123-127): w*h RGBA8 texels. */
int omv_scene_set_texture(omv_scene* s, int tex_id, int w, int h, const uint8_t* rgba);
/* fp32 RGBA texture; only read by the fp32-texture extension path. */
int omv_scene_set_texture_f32(omv_scene* s, int tex_id, int w, int h, const float* rgba);
/* RenderLightPoint::set (RenderLightPoint.cpp:20-26).  range<0 = infinite. */
int omv_scene_set_light_point(omv_scene* s, int light_id, const float pos[3],
                              const float intensity[3], float range);
/* HDRBackground (HdrBackground.cpp:38-97): w*h RGB fp32, rotation row-major,
 * radiance scale (main.cpp:400-405 uses the z-up->y-up matrix and pi). */
int omv_scene_set_background_hdr(omv_scene* s, int w, int h, const float* rgb,
                                 const float rotation[9], float scale);
/* Background base class: constant emitted radiance (Background.h). */
int omv_scene_set_background_constant(omv_scene* s, const float rgb[3]);
/* RenderScene::setFalseColorSubject: "pid","gid","mid","rmp","mmp","md","sg","mn","in"
 * (+ "iid", "ipid", the instance-aware subjects). */
int omv_scene_set_false_color_subject(omv_scene* s, const char* name);
/* RenderScene::commitScene (RenderScene.cpp:155-223): mesh normals/tangents,
 * instance validation, light CDF, BVH build and upload.  log receives the
 * reference's log lines, '\n'-separated (may be NULL).  Returns OMV_OK or
 * OMV_E_SCENE when every instance is invalid. */
int omv_scene_commit(omv_scene* s, char* log, size_t log_cap);
/* How commit builds the per-mesh BVHs (no reference counterpart: the
 * reference hands the meshes to Embree, RenderScene.cpp:225-272):
 *   OMV_BUILD_HOST_SAH      (default) binned SAH on the host: fewest node
 *                           visits per ray;
 *   OMV_BUILD_DEVICE_LBVH   linear BVH built on the GPU (Morton codes, radix
 *                           sort, Karras radix tree, collapsed to 4-wide
 *                           nodes): commit time of large meshes drops by an
 *                           order of magnitude, traversal visits more nodes.
 * The film does not depend on the choice (closest hits are independent of
 * the tree).  Takes effect at the next commit. */
#define OMV_BUILD_HOST_SAH     0
#define OMV_BUILD_DEVICE_LBVH  1
int omv_scene_set_build_mode(omv_scene* s, int mode);
/* In-place parameter updates of a committed scene (no reference counterpart:
 * the optimiser loop of the adjoint form changes material parameters between
 * renders without a new commit).  texels: same size and format as set, host
 * or device pointer (w*h RGBA8 or w*h*4 floats); the OM (n, k) replace the
 * baked per-RGB IOR of an OpenMATERIAL material. */
int omv_scene_update_texture(omv_scene* s, int tex_id, const void* texels);
int omv_scene_update_material_om_nk(omv_scene* s, int mat_id, const float n_rgb[3], const float k_rgb[3]);
/* Inspection: the committed interleaved vertex buffer of a mesh (stride
 * 10+2*channels floats: pos3 normal3 tangent4 uv2*ch, RenderMesh.cpp:49-72). */
int omv_scene_mesh_info(omv_scene* s, int mesh_id, int* vert_n, int* faces_n,
                        int* stride, int* mat_id);
int omv_scene_get_vertices(omv_scene* s, int mesh_id, float* out, size_t cap_floats);
/* Inspection (OMV_EXT_HDRI_NEE): the environment sampling tables built on
 * the device at commit -- row table (H entries), per-row column tables and
 * texel probabilities (W*H entries each); host buffers. */
int omv_scene_env_tables(omv_scene* s, float* row_prob, int32_t* row_alias, float* col_prob, int32_t* col_alias,
                         float* tex_pdf);
/* BVH statistics: node count, max depth, leaf triangle records. */
int omv_scene_bvh_info(omv_scene* s, int* nodes, int* max_depth, int* tri_records,
                       size_t* device_bytes);

/* ----------------------------------------------------- camera (PinholeCamera) */
typedef struct omv_camera omv_camera;
int omv_camera_create(omv_camera** out, int width, int height);   /* PinholeCamera(w,h) */
int omv_camera_destroy(omv_camera* c);
/* ViewPoint(loc, xdir, ydir, zdir) + setViewPoint. */
int omv_camera_set_viewpoint(omv_camera* c, const float pos[3], const float xdir[3],
                             const float ydir[3], const float zdir[3]);
int omv_camera_adjust_to_resolution(omv_camera* c);               /* :156-160 */
int omv_camera_set_yfov(omv_camera* c, float yfov_deg);           /* :167-174 */
int omv_camera_set_real_height(omv_camera* c, float height, float aspect);
int omv_camera_set_focus(omv_camera* c, float focus);
/* PinholeCamera::loadProperties (PinholeCamera.cpp:115-148): the "pinhole"
 * object of a camera JSON file, parsed by the caller; NULL = key absent
 * (aspect and y-fov then 0, focus unchanged).  half_h = focus *
 * tan(y-fov / 2), half_w = aspect * half_h, no clamping. */
int omv_camera_load_properties(omv_camera* c, const float* aspect, const float* focus, const float* yfov);
int omv_camera_get_desc(const omv_camera* c, omv_camera_desc* out);
/* getImpression (:110-113) for one pixel, and the whole normalised film
 * (W*H*4 floats, rgb + 1) after a render. */
int omv_camera_get_impression(const omv_camera* c, int x, int y, float out[4]);
int omv_camera_get_film(const omv_camera* c, float* out, size_t cap_floats);

/* ------------------------------------------------- renderer (BDPTRenderer) */
typedef struct omv_renderer omv_renderer;
int omv_renderer_create(omv_renderer** out);
int omv_renderer_destroy(omv_renderer* r);
int omv_renderer_set_parameters(omv_renderer* r, const omv_renderer_params* p);
int omv_renderer_set_scene(omv_renderer* r, omv_scene* s);
/* BDPTRenderer::render (BDPTRenderer.cpp:372-393): init the film, run every
 * sample on the GPU, normalise (PinholeCamera::stop).  Synchronous. */
int omv_renderer_render(omv_renderer* r, omv_camera* c);
/* omv_renderer_render with the normalised film (W*H*4 floats) written to
 * film_dev, a device buffer, instead of being read back into the camera
 * (no reference counterpart: for callers that keep the image on the GPU,
 * e.g. a loss on the device; the camera's film and omv_renderer_post are
 * not updated).  Synchronous; same film bit for bit. */
int omv_renderer_render_device(omv_renderer* r, omv_camera* c, float* film_dev);
/* False-colour pass (reference FC mode: main.cpp:261-270 forces spp=1,
 * B=b=1): the camera film gets the subject's colours; ids receives four int32
 * planes of W*H: global primitive id, mesh id, material id (-1 = missing
 * material), instance id; -1 where the ray escapes.  tie (may be NULL) gets 1
 * where the closest hit had an exact-t tie with a different triangle. */
int omv_renderer_render_fc(omv_renderer* r, omv_camera* c, int32_t* ids, uint8_t* tie);
/* The colour table the enumerated subjects (pid, gid, mid, iid, ipid) paint
 * with: EnumeratedFCSubject's construction (FCSubject.h:153-167) -- size
 * distinct colours 0xRRGGBB in [1, 0xFFFFFF], ascending, then shuffled --
 * with the draws and the shuffle taken from a generator seeded by
 * (seed, subject) instead of the reference's unseeded one, so that a table,
 * and with it an FC image, is reproducible.  render_fc uses the renderer's
 * params.seed.  Category sizes: pid = all mesh faces, gid = meshes, mid =
 * allocated materials, iid = instances, ipid = faces over all instances.
 * size > 0xFFFFFF is OMV_E_RANGE. */
int omv_fc_color_table(uint32_t seed, int subject, uint32_t size, uint32_t* out);
int omv_renderer_get_stats(const omv_renderer* r, omv_render_stats* out);
/* How omv_renderer_render schedules the work on the GPU (no reference
 * counterpart; both give the same film bit for bit):
 *   OMV_SCHEDULE_WAVEFRONT   (default) ray queues: gen, then per bounce a
 *                            traversal kernel and a shading kernel
 *                            (+ connection visibility and sum with lights);
 *   OMV_SCHEDULE_PER_SAMPLE  one persistent kernel, one lane per sample. */
#define OMV_SCHEDULE_WAVEFRONT   0
#define OMV_SCHEDULE_PER_SAMPLE  1
int omv_renderer_set_schedule(omv_renderer* r, int schedule);
int omv_renderer_get_schedule(const omv_renderer* r, int* schedule);
/* Launch tuning of the wavefront schedule (no reference counterpart; the
 * film does not depend on it). */
typedef struct omv_render_tuning {
    int trace_keep;           /* 1..64: persistent traversal, a wave refills its
                                 idle lanes once fewer than trace_keep are busy;
                                 0: one pass over the queue, one ray per lane   */
    int trace_blocks_per_cu;  /* resident traversal workgroups per CU
                                 (0 = occupancy limit)                          */
    int chunk_samples;        /* samples per wavefront chunk, rounded down to
                                 whole pixels (0 = default: 2^25)               */
    int traversal;            /* 0 = default = 1: every ray entry by entry,
                                 one lane each; 2 = the first bounce's camera
                                 rays (64 consecutive samples per wave) as
                                 wave packets: one stack of (entry, lane mask)
                                 per wave, node and triangle records read
                                 once per wave.  Same hits and film      */
    int sort_rays;            /* secondary camera rays reordered by
                                 (direction octant, origin cell of a 16^3
                                 grid over the scene) before each traversal:
                                 0 = default, 1 = never, 2 = always          */
    int chunk_streams;        /* renders of several chunks: 2 (= default 0):
                                 chunks alternate between two streams, so one
                                 chunk's late bounces (few rays, bound by
                                 traversal latency) run beside the next
                                 chunk's first bounces; 1: chunks one after
                                 another on one stream                        */
    int chunk_stagger;        /* with two streams: a chunk starts once the
                                 chunk before it has finished this many camera
                                 rounds (0 = default: 1)                      */
    int occlusion;            /* visibility rays (connections, environment and
                                 pinhole rays): 0 = default, the traversal ends
                                 at the first hit that decides occlusion (one
                                 before isConnected's acceptance radius
                                 around the connected point, or any hit for
                                 the rays that need none); 1 = closest hit
                                 as sceneIntersect finds it.  Same decisions,
                                 same film; scenes with alpha-MASK materials
                                 always take the closest hit                  */
    int xcd_bands;            /* traversal hand-out: n > 1 = the queue is cut
                                 into chunks dealt to n bands of workgroups
                                 (workgroup b in band b mod n, so a band lies
                                 on one XCD when n is a multiple of the XCD
                                 count): the rays an XCD traces at a time come
                                 from a few stretches of the queue; 1 = plain
                                 grid stride; 0 = default: four bands per XCD
                                 on scenes of at most 4 instances, else the
                                 plain stride                               */
    int tail_round;           /* camera rounds from this one to the last run as
                                 one launch, each path carried to its end by
                                 one lane (same film): 0 = default, 1 = never
                                 (every round as its own trace, shade and
                                 connect launches), n >= 2 = from round n;
                                 the default is 5 with point lights or NEE,
                                 else 4                                      */
} omv_render_tuning;
int omv_renderer_set_tuning(omv_renderer* r, const omv_render_tuning* t);
int omv_renderer_get_tuning(const omv_renderer* r, omv_render_tuning* t);

/* Film post pass of main.cpp:513-519 on the device, on the camera's last
 * film from this renderer: gammaCorrection(A = 1, gamma)
 * (ToneMapping.cpp:24-54), then when denoise != 0 the 3x3 median filter
 * SimpleDenoiserMedian<1> (Denoiser.h:24-68), and the 8-bit truncation of
 * saveImage (main.cpp:88-93).  out_f32 (W*H*4 floats) and out_u8 (W*H*4
 * bytes) are host or device pointers, either may be NULL. */
int omv_renderer_post(omv_renderer* r, omv_camera* c, float gamma, int denoise, float* out_f32, uint8_t* out_u8);

/* saveImage (main.cpp:86-120) with ImageSaver::flipVerticallyOnWrite(true)
 * (main.cpp:149): rgba is the W*H*4 gamma-corrected float image (host
 * memory, values in [0, 1]); ".png", ".bmp", ".tga", ".jpg" (quality 100)
 * write the 8-bit truncation 255*f, ".pfm" the floats.  The bytes equal
 * stb_image_write's (ImageSaver.h:23-97).  Another extension is OMV_E_ARG
 * ("This image extension is not supported"), a file error OMV_E_IO. */
int omv_save_image(const char* filename, const float* rgba, int width, int height);
/* The bytes omv_save_image would write for extension ext ("png", "bmp",
 * "tga", "jpg", "pfm"), into out (capacity cap); *len receives the size.
 * With out == NULL or cap too small only *len is set (OMV_E_RANGE when
 * out != NULL and cap < *len). */
int omv_encode_image(const char* ext, const float* rgba, int width, int height, uint8_t* out, size_t cap, size_t* len);

/* Allocation-free render.  No counterpart in the reference: this is for
 * callers that cannot allocate device memory while they render -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Renders the renderer's committed scene with its parameters from camera c
 * into film_dev (device pointer, W*H*4 floats: the normalised film
 * omv_renderer_render gives, bit for bit) on `stream` (a hipStream_t; NULL =
 * the null stream), as one persistent per-pixel kernel.  It calls no
 * hipMalloc / hipFree: its only device memory besides the scene and film_dev
 * is a fixed array of OMV_NOALLOC_SCRATCH_BYTES per GPU (counters, traversal
 * spill stacks, light vertices of the resident lanes), which each GPU the
 * library runs on holds until the process ends.  Every call on a GPU shares
 * that GPU's array, so calls that overlap on one GPU race.  stats == NULL:
 * the call only enqueues its work (no wait, usable under capture); else it
 * waits for the stream and fills samples, ray and traversal counts (the
 * times are 0).  Limits: min(light_bounces + 1, max_path_length) light
 * vertices kept <= OMV_NOALLOC_MAX_LIGHT_VERTICES (else OMV_E_RANGE);
 * OMV_EXT_LIGHT_SPLAT with point lights needs a W*H buffer of sums and is
 * refused (OMV_E_STATE). */
#define OMV_NOALLOC_SCRATCH_BYTES      (64u << 20)
#define OMV_NOALLOC_MIN_BLOCKS         256   /* blocks of 128 lanes the scratch must hold */
#define OMV_NOALLOC_MAX_LIGHT_VERTICES 68
int omv_renderer_render_noalloc(omv_renderer* r, const omv_camera* c, float* film_dev, void* stream, omv_render_stats* stats);

/* ---------------------------------------------------------- adjoint */
/* Gradient targets of omv_renderer_backward.  All pointers are device
 * pointers; gradients are ADDED to them.  texture_grads[i] holds w*h*4
 * floats for texture texture_ids[i] (the gradient with respect to the texel
 * values as floats, byte/255 for RGBA8); om_grads holds 6 floats per OM
 * material (d/dn rgb, d/dk rgb). */
typedef struct omv_adjoint_targets {
    int n_textures;
    const int* texture_ids;       /* host array */
    float* const* texture_grads;  /* host array of device pointers */
    int n_om_materials;
    const int* om_material_ids;   /* host array */
    float* om_grads;              /* device pointer, 6 * n_om_materials floats */
} omv_adjoint_targets;
/* Path-replay backward pass of omv_renderer_render (no reference
 * counterpart; the reference has no derivatives): given d_film = d loss /
 * d film (device pointer, W*H*4 floats, rgb used) for the camera's last
 * film, adds d loss / d parameter for the base-colour (rgb) and
 * metallic-roughness (G roughness, B metallic) texels of PBR materials and
 * the (n, k) of OpenMATERIAL mirrors.  Detached sampling: directions and
 * lobes are those the current parameters sample; the derivative is that of
 * the weights.  Scenes with point lights and the OMV_EXT_OM_ROUGH
 * extension return OMV_E_STATE. */
int omv_renderer_backward(omv_renderer* r, omv_camera* c, const float* d_film, const omv_adjoint_targets* t);

/* FP64 evaluation of the functional omv_renderer_backward differentiates
 * (no reference counterpart; for checking gradients by central finite
 * differences on the device's own paths): directions, lobes and hits are
 * sampled from the renderer's scene exactly as omv_renderer_render samples
 * them; every bounce weight is re-evaluated with eval's parameters (the
 * sampled PBR lobe's w * f_eval / f_scene, an OM mirror's Fresnel term of
 * eval's (n, k)), with material inputs, weights, throughput and radiance in
 * double.  eval is a committed scene with the same meshes, material kinds
 * and texture sizes (a copy with some texels or (n, k) changed).  film:
 * W*H*4 doubles (rgb + 1), host or device pointer.  Same scope as
 * omv_renderer_backward (OMV_E_STATE with point lights or an extension). */
int omv_renderer_render_detached(omv_renderer* r, omv_camera* c, const omv_scene* eval, double* film);

/* ------------------------------------------------------ batched jobs */
/* Many independent validation renders (each one BDPTRenderer::render call of
 * main.cpp:489-521 with its own scene, camera and parameters) run as ONE
 * persistent launch.  The reference has no batched form: its closest
 * counterpart is the per-viewpoint loop main.cpp:489-521 run once per asset
 * (the pathtracer/scripts shell scripts drive it per file).  Each job's film equals, bit
 * for bit, what omv_renderer_render gives for the same scene, camera and
 * parameters.  Scenes must stay committed and unchanged while a batch that
 * names them exists; cameras are copied at add time. */
typedef struct omv_batch omv_batch;
typedef struct omv_batch_tuning {
    int blocks_per_cu;   /* per-pixel kernel: resident workgroups per CU
                            (0 = occupancy limit)                          */
    int tile_waves;      /* per-pixel kernel: pixels per work item =
                            64 * tile_waves (>= 1)                         */
    int use_graph;       /* 1: replay a captured hipGraph, 0: plain launch */
    int streams;         /* 0: one persistent per-pixel kernel over all
                            jobs; N >= 1: each job's wavefront launch
                            sequence, jobs dealt over N streams so that one
                            job's last bounce rounds overlap other jobs    */
} omv_batch_tuning;
int omv_batch_create(omv_batch** out);
int omv_batch_destroy(omv_batch* b);
/* Adds a job; job ids count from 0 in call order. */
int omv_batch_add_job(omv_batch* b, const omv_scene* s, const omv_camera* c,
                      const omv_renderer_params* p, int* job_id);
int omv_batch_job_count(const omv_batch* b, int* n);
int omv_batch_set_tuning(omv_batch* b, const omv_batch_tuning* t);
int omv_batch_get_tuning(const omv_batch* b, omv_batch_tuning* t);
/* Times every candidate on the batch itself with plain launches (no graph
 * capture during the sweep), 'reps' runs each, keeps the fastest and reports
 * it in best (may be NULL): the jobs' wavefront sequences over 1, 2, 4 and 8
 * streams (no more streams than jobs) and the persistent per-pixel kernel at
 * each blocks_per_cu up to the occupancy limit x tile_waves 1, 2, 4.  The
 * starting use_graph is kept: the choice is captured once, at the next run. */
int omv_batch_autotune(omv_batch* b, int reps, omv_batch_tuning* best);
/* Renders every job.  Synchronous. */
int omv_batch_run(omv_batch* b);
/* Normalised film of one job (W*H*4 floats), as omv_camera_get_film. */
int omv_batch_get_film(const omv_batch* b, int job_id, float* out, size_t cap_floats);
/* Totals over all jobs of the last run; render_ms is the launch's device time. */
int omv_batch_get_stats(const omv_batch* b, omv_render_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* OMV_H */
