/*
 * refcpu.h -- C API of the CPU oracle (test infrastructure only).
 *
 * The oracle is a from-scratch CPU restatement of the reference pathtracer's
 * BDPT render loop.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product
 * (libomv.so) never links, calls or falls back to it.
 *
 * Scene set-up mirrors include/omv.h one to one (ref_ prefix); the POD
 * types (params, camera, PBR and IOR descriptors) are shared with it.
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stddef.h>
#include <stdint.h>
#include "../include/omv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_scene ref_scene;

ref_scene* ref_scene_create(void);
void ref_scene_destroy(ref_scene* s);
int  ref_scene_allocate(ref_scene* s, int meshes, int instances, int materials, int textures, int lights);
int  ref_scene_set_mesh(ref_scene* s, int mesh_id, int vert_n, int faces_n, int tex_n, int mat_id,
                        const uint32_t* faces, const float* vert, int vstride, const float* norm, int nstride,
                        const float* tans, int tanstride, const float* const* tex, int texstride);
int  ref_scene_set_instance(ref_scene* s, int inst_id, const float trans[16], int mesh_id);
int  ref_scene_set_material_om(ref_scene* s, int mat_id, float asset_temperature, int n_tables,
                               const omv_ior_table* tables);
int  ref_scene_set_material_om_nk(ref_scene* s, int mat_id, const float n_rgb[3], const float k_rgb[3]);
int  ref_scene_set_material_pbr(ref_scene* s, int mat_id, const omv_pbr_desc* d);
int  ref_scene_set_texture(ref_scene* s, int tex_id, int w, int h, const uint8_t* rgba);
int  ref_scene_set_texture_f32(ref_scene* s, int tex_id, int w, int h, const float* rgba);
int  ref_scene_set_light_point(ref_scene* s, int light_id, const float pos[3], const float intensity[3], float range);
int  ref_scene_set_background_hdr(ref_scene* s, int w, int h, const float* rgb, const float rotation[9], float scale);
int  ref_scene_set_background_constant(ref_scene* s, const float rgb[3]);
int  ref_scene_set_false_color_subject(ref_scene* s, const char* name);
int  ref_scene_commit(ref_scene* s, char* log, size_t log_cap);
int  ref_scene_mesh_info(ref_scene* s, int mesh_id, int* vert_n, int* faces_n, int* stride, int* mat_id);
int  ref_scene_get_vertices(ref_scene* s, int mesh_id, float* out, size_t cap_floats);
/* baked per-RGB complex IOR of an OM material after commit */
int  ref_scene_get_om_nk(ref_scene* s, int mat_id, float n_rgb[3], float k_rgb[3]);

/* Full render in float (fp64 = 0) or double (fp64 = 1).  film: W*H*4 floats,
 * normalised exactly as PinholeCamera::stop.  Rows [row0, row0+rows) only
 * (rows <= 0: all).  threads <= 0: params->number_of_cores.  stats may be NULL. */
int  ref_render(ref_scene* s, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64,
                int row0, int rows, int threads, float* film, omv_render_stats* stats);
/* Per-sample radiance (W*H*spp*3 floats, sample index (y*W+x)*spp+k) for rows [row0,row0+rows). */
int  ref_render_samples(ref_scene* s, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64,
                        int row0, int rows, int threads, float* rad);
/* False-colour pass, see omv_renderer_render_fc. */
int  ref_fc_color_table(uint32_t seed, int subject, uint32_t size, uint32_t* out);
int  ref_render_fc(ref_scene* s, const omv_renderer_params* p, const omv_camera_desc* cam,
                   float* film, int32_t* ids, uint8_t* tie);

/* Closest hit of one world-space ray; bvh = 0 brute force over all triangles,
 * 1 through the oracle's binary SAH BVH.  Returns 1 on hit. */
int  ref_closest_hit(ref_scene* s, const float org[3], const float dir[3], float tnear, float tfar, int bvh,
                     float* t, float* u, float* v, int* inst, int* prim, int* tie,
                     uint64_t* nodes, uint64_t* tris);

/* ---- unit-level restatements (pinned against golden vectors) ---- */
void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float ref_u01(uint32_t x);
void ref_fresnel(float n, float k, float cos_theta, float* rp, float* rs);
void ref_fresnel64(double n, double k, double cos_theta, double out[2]);
void ref_fresnel_grad(double n, double k, double cos_theta, double out[4]);
void ref_median3(float* img, int w, int h);
int  ref_env_tables(ref_scene* s, float* row_prob, int* row_alias, float* col_prob, int* col_alias, float* tex_pdf);
int  ref_env_sample(ref_scene* s, const float u[6], double dir[3], double* pdf, int* row, int* col);
double ref_env_pdf(ref_scene* s, const double dir[3]);
void ref_to_u8(const float* img, size_t n, uint8_t* out);
void ref_adjoint_plant(int mask);   /* test aid: plants a known mistake in the adjoint */
void ref_splat_plant(int mask);     /* test aid: plants a known mistake in the light-tracer splats */
int  ref_render_detached(ref_scene* sample, ref_scene* eval, const omv_renderer_params* p, const omv_camera_desc* cam,
                         int fp64, int threads, double* film);
int  ref_adjoint(ref_scene* s, const omv_renderer_params* p, const omv_camera_desc* cam, int fp64, int threads,
                 const float* dfilm, int ntex, const int* tex_ids, double* const* tex_grads, int nom, const int* om_ids,
                 double* om_grads);
void ref_evaluate_direct(const float I[3], const float O[3], const float N[3], const float color[4],
                         float metallic, float alpha, float out[3]);
/* BSDFModel::sampleIndirect with the uniforms given explicitly; returns the
 * number of uniforms consumed. */
int  ref_sample_indirect(const float I[3], const float N[3], const float T[3], const float color[4],
                         float metallic, float alpha, float eta, int transmissive,
                         const float* uniforms, float brdf[3], float O[3]);
void ref_uniform_hemisphere(float u0, float u1, float out[3]);
int  ref_uniform_sphere(const float* uniforms, int n_uniforms, float out[3]);
void ref_ggx_hemisphere(float u0, float u1, float alpha2, float out[3]);
void ref_cosine_sqrt_hemisphere(float u0, float u1, float out[3]);
void ref_texture_rgba8(int w, int h, const uint8_t* rgba, float u, float v, float out[4]);
void ref_hdr_radiance(int w, int h, const float* rgb, const float rotation[9], float scale,
                      const float dir[3], float out[3]);
void ref_camera_ray(const omv_camera_desc* cam, int x, int y, float u0, float u1, float org[3], float dir[3]);
int  ref_distance_random(int n, const float* weights, float u, float* pdf);
/* AssetMaterialIor::getIor on explicit tables; returns 0 or -1 if the
 * wavelength is outside every table (the reference throws). */
int  ref_get_ior(int n_tables, const omv_ior_table* tables, float temperature, float wavelength,
                 float* n, float* k);
/* tonemapping gammaCorrection (ToneMapping.cpp:24-54), in place on W*H*4. */
void ref_apply_texture_normal(const float* frame, const float* tn, int calls, float* out);
void ref_texel_normal(const float* texel, float scale, float* out);
void ref_om_next_direction(const float* inV, const float* normal, const float* tri, const float* nm, float* out);
void ref_gamma_correction(float* img, int w, int h, float A, float gamma);
/* Reference float helpers exposed for invariant tests. */
void ref_compute_camera(int w, int h, float yfov_deg, float* half_w, float* half_h, float* focus);
void ref_renderer_params_default(omv_renderer_params* p);

#ifdef __cplusplus
}
#endif
#endif
