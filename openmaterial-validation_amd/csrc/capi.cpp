// capi.cpp -- the extern "C" boundary declared in include/omv.h.
#include <cstring>
#include <new>
#include <string>

#include "fc_table.h"
#include "render_core.h"

struct omv_scene { omv::RenderScene s; };
struct omv_camera { omv::PinholeCamera c; omv_camera(int w, int h) : c(w, h) {} };
struct omv_renderer { omv::BDPTRenderer r; };
struct omv_batch { omv::BatchRunner b; };

#define OMV_NULLCHK(p) do { if (!(p)) { omv::setError(std::string(__func__) + ": null argument"); return OMV_E_ARG; } } while (0)

extern "C" {

const char* omv_last_error(void) { return omv::lastError(); }
const char* omv_version(void) { return "omv-mi355x 0.1 (gfx950)"; }

int omv_device_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) { omv::setError(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e)); return OMV_E_DEVICE; }
    return OMV_OK;
}

int omv_device_info(char* name, size_t cap, int* cus, size_t* mem) {
    int dev = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) { omv::setError(std::string("hipGetDeviceProperties: ") + hipGetErrorString(e)); return OMV_E_DEVICE; }
    if (name && cap) { std::strncpy(name, prop.gcnArchName, cap - 1); name[cap - 1] = 0; }
    if (cus) *cus = prop.multiProcessorCount;
    if (mem) *mem = prop.totalGlobalMem;
    return OMV_OK;
}

void omv_renderer_params_default(omv_renderer_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // RendererParameters defaults (RendererParameters.h:24-96)
    p->camera_bounces = 10;
    p->light_bounces = 10;
    p->max_path_length = 8;
    p->samples_per_pixel = 20;
    p->light_distance_attenuation = 1;
    p->light_scale_value = 100.0f;
    p->light_min_distance = 0.01f;
    p->ray_cut_pix_value = 0.002f;
    p->gamma = 0.5f;
    p->number_of_cores = 16;
    p->false_color_mode = 0;
    p->seed = 0;
    p->extensions = OMV_EXT_NONE;
    p->om_roughness = 0.3f;
}

int omv_scene_create(omv_scene** out) {
    OMV_NULLCHK(out);
    *out = new (std::nothrow) omv_scene();
    if (!*out) { omv::setError("omv_scene_create: out of memory"); return OMV_E_NOMEM; }
    return OMV_OK;
}
int omv_scene_destroy(omv_scene* s) { delete s; return OMV_OK; }
int omv_scene_allocate(omv_scene* s, int meshes, int instances, int materials, int textures, int lights) {
    OMV_NULLCHK(s);
    return s->s.allocate(meshes, instances, materials, textures, lights);
}
int omv_scene_set_mesh(omv_scene* s, int mesh_id, int vert_n, int faces_n, int tex_n, int mat_id, const uint32_t* faces,
                       const float* vert, int vstride, const float* norm, int nstride, const float* tans, int tanstride,
                       const float* const* tex, int texstride) {
    OMV_NULLCHK(s);
    return s->s.setMesh(mesh_id, vert_n, faces_n, tex_n, mat_id, faces, vert, vstride, norm, nstride, tans, tanstride, tex, texstride);
}
int omv_scene_set_instance(omv_scene* s, int inst_id, const float trans[16], int mesh_id) {
    OMV_NULLCHK(s);
    return s->s.setInstance(inst_id, trans, mesh_id);
}
int omv_scene_set_material_om(omv_scene* s, int mat_id, float t, int n_tables, const omv_ior_table* tables) {
    OMV_NULLCHK(s);
    return s->s.setMaterialOm(mat_id, t, n_tables, tables);
}
int omv_scene_set_material_om_nk(omv_scene* s, int mat_id, const float n[3], const float k[3]) {
    OMV_NULLCHK(s); OMV_NULLCHK(n); OMV_NULLCHK(k);
    return s->s.setMaterialOmNk(mat_id, n, k);
}
int omv_scene_set_material_pbr(omv_scene* s, int mat_id, const omv_pbr_desc* d) {
    OMV_NULLCHK(s); OMV_NULLCHK(d);
    return s->s.setMaterialPbr(mat_id, *d);
}
int omv_scene_set_texture(omv_scene* s, int tex_id, int w, int h, const uint8_t* rgba) {
    OMV_NULLCHK(s);
    return s->s.setTexture(tex_id, w, h, rgba);
}
int omv_scene_set_texture_f32(omv_scene* s, int tex_id, int w, int h, const float* rgba) {
    OMV_NULLCHK(s);
    return s->s.setTextureF32(tex_id, w, h, rgba);
}
int omv_scene_set_light_point(omv_scene* s, int light_id, const float pos[3], const float intensity[3], float range) {
    OMV_NULLCHK(s); OMV_NULLCHK(pos); OMV_NULLCHK(intensity);
    return s->s.setLightPoint(light_id, pos, intensity, range);
}
int omv_scene_set_background_hdr(omv_scene* s, int w, int h, const float* rgb, const float rot[9], float scale) {
    OMV_NULLCHK(s);
    return s->s.setBackgroundHdr(w, h, rgb, rot, scale);
}
int omv_scene_set_background_constant(omv_scene* s, const float rgb[3]) {
    OMV_NULLCHK(s); OMV_NULLCHK(rgb);
    return s->s.setBackgroundConstant(rgb);
}
int omv_scene_set_false_color_subject(omv_scene* s, const char* name) {
    OMV_NULLCHK(s); OMV_NULLCHK(name);
    return s->s.setFalseColorSubject(name);
}
int omv_scene_set_build_mode(omv_scene* s, int mode) {
    OMV_NULLCHK(s);
    return s->s.setBuildMode(mode);
}
int omv_scene_commit(omv_scene* s, char* log, size_t log_cap) {
    OMV_NULLCHK(s);
    std::string lg;
    int rc = s->s.commit(lg);
    if (log && log_cap) {
        std::strncpy(log, lg.c_str(), log_cap - 1);
        log[log_cap - 1] = 0;
    }
    return rc;
}
int omv_scene_mesh_info(omv_scene* s, int mesh_id, int* vert_n, int* faces_n, int* stride, int* mat_id) {
    OMV_NULLCHK(s);
    return s->s.meshInfo(mesh_id, vert_n, faces_n, stride, mat_id);
}
int omv_scene_get_vertices(omv_scene* s, int mesh_id, float* out, size_t cap) {
    OMV_NULLCHK(s);
    return s->s.vertices(mesh_id, out, cap);
}
int omv_scene_bvh_info(omv_scene* s, int* nodes, int* depth, int* records, size_t* bytes) {
    OMV_NULLCHK(s);
    return s->s.bvhInfo(nodes, depth, records, bytes);
}

int omv_camera_create(omv_camera** out, int w, int h) {
    OMV_NULLCHK(out);
    *out = new (std::nothrow) omv_camera(w, h);
    if (!*out) { omv::setError("omv_camera_create: out of memory"); return OMV_E_NOMEM; }
    return OMV_OK;
}
int omv_camera_destroy(omv_camera* c) { delete c; return OMV_OK; }
int omv_camera_set_viewpoint(omv_camera* c, const float pos[3], const float x[3], const float y[3], const float z[3]) {
    OMV_NULLCHK(c); OMV_NULLCHK(pos); OMV_NULLCHK(x); OMV_NULLCHK(y); OMV_NULLCHK(z);
    c->c.setViewPoint(pos, x, y, z);
    return OMV_OK;
}
int omv_camera_adjust_to_resolution(omv_camera* c) { OMV_NULLCHK(c); c->c.adjustToResolution(); return OMV_OK; }
int omv_camera_set_yfov(omv_camera* c, float fov) { OMV_NULLCHK(c); return c->c.setYFoV(fov); }
int omv_camera_set_real_height(omv_camera* c, float h, float a) { OMV_NULLCHK(c); c->c.setRealHeight(h, a); return OMV_OK; }
int omv_camera_set_focus(omv_camera* c, float f) { OMV_NULLCHK(c); c->c.desc.focus = f; return OMV_OK; }
int omv_camera_load_properties(omv_camera* c, const float* aspect, const float* focus, const float* yfov) {
    OMV_NULLCHK(c);
    c->c.loadProperties(aspect, focus, yfov);
    return OMV_OK;
}
int omv_camera_get_desc(const omv_camera* c, omv_camera_desc* out) {
    OMV_NULLCHK(c); OMV_NULLCHK(out);
    *out = c->c.desc;
    return OMV_OK;
}
// PinholeCamera::getImpression (PinholeCamera.cpp:110-113)
int omv_camera_get_impression(const omv_camera* c, int x, int y, float out[4]) {
    OMV_NULLCHK(c); OMV_NULLCHK(out);
    const omv_camera_desc& d = c->c.desc;
    if (x >= 0 && x < d.width && y >= 0 && y < d.height && !c->c.film.empty()) {
        std::memcpy(out, &c->c.film[((size_t)y * d.width + x) * 4], 4 * sizeof(float));
    } else {
        out[0] = out[1] = out[2] = 0.0f;
        out[3] = 1.0f;
    }
    return OMV_OK;
}
int omv_camera_get_film(const omv_camera* c, float* out, size_t cap) {
    OMV_NULLCHK(c); OMV_NULLCHK(out);
    if (cap < c->c.film.size() || c->c.film.empty()) { omv::setError("omv_camera_get_film: no film or buffer too small"); return OMV_E_ARG; }
    std::memcpy(out, c->c.film.data(), c->c.film.size() * sizeof(float));
    return OMV_OK;
}

int omv_renderer_create(omv_renderer** out) {
    OMV_NULLCHK(out);
    *out = new (std::nothrow) omv_renderer();
    if (!*out) { omv::setError("omv_renderer_create: out of memory"); return OMV_E_NOMEM; }
    return OMV_OK;
}
int omv_renderer_destroy(omv_renderer* r) { delete r; return OMV_OK; }
int omv_renderer_set_parameters(omv_renderer* r, const omv_renderer_params* p) {
    OMV_NULLCHK(r); OMV_NULLCHK(p);
    r->r.setParameters(*p);
    return OMV_OK;
}
int omv_renderer_set_scene(omv_renderer* r, omv_scene* s) {
    OMV_NULLCHK(r); OMV_NULLCHK(s);
    r->r.setScene(&s->s);
    return OMV_OK;
}
int omv_renderer_render(omv_renderer* r, omv_camera* c) {
    OMV_NULLCHK(r); OMV_NULLCHK(c);
    return r->r.render(c->c);
}
int omv_renderer_render_device(omv_renderer* r, omv_camera* c, float* film_dev) {
    OMV_NULLCHK(r); OMV_NULLCHK(c); OMV_NULLCHK(film_dev);
    return r->r.render(c->c, film_dev);
}
int omv_renderer_render_noalloc(omv_renderer* r, const omv_camera* c, float* film_dev, void* stream, omv_render_stats* stats) {
    OMV_NULLCHK(r); OMV_NULLCHK(c); OMV_NULLCHK(film_dev);
    return r->r.renderNoalloc(c->c, film_dev, static_cast<hipStream_t>(stream), stats);
}
int omv_renderer_render_fc(omv_renderer* r, omv_camera* c, int32_t* ids, uint8_t* tie) {
    OMV_NULLCHK(r); OMV_NULLCHK(c);
    return r->r.renderFc(c->c, ids, tie);
}
int omv_renderer_get_stats(const omv_renderer* r, omv_render_stats* out) {
    OMV_NULLCHK(r); OMV_NULLCHK(out);
    *out = r->r.stats();
    return OMV_OK;
}
int omv_renderer_post(omv_renderer* r, omv_camera* c, float gamma, int denoise, float* out_f32, uint8_t* out_u8) {
    OMV_NULLCHK(r); OMV_NULLCHK(c);
    return r->r.post(c->c, gamma, denoise != 0, out_f32, out_u8);
}
int omv_renderer_backward(omv_renderer* r, omv_camera* c, const float* d_film, const omv_adjoint_targets* t) {
    OMV_NULLCHK(r); OMV_NULLCHK(c); OMV_NULLCHK(t);
    return r->r.backward(c->c, d_film, *t);
}
int omv_renderer_render_detached(omv_renderer* r, omv_camera* c, const omv_scene* eval, double* film) {
    OMV_NULLCHK(r); OMV_NULLCHK(c); OMV_NULLCHK(eval); OMV_NULLCHK(film);
    return r->r.renderDetached(c->c, eval->s, film);
}
int omv_scene_env_tables(omv_scene* s, float* rp, int32_t* ra, float* cp, int32_t* ca, float* tp) {
    OMV_NULLCHK(s); OMV_NULLCHK(rp); OMV_NULLCHK(ra); OMV_NULLCHK(cp); OMV_NULLCHK(ca); OMV_NULLCHK(tp);
    return s->s.envTables(rp, ra, cp, ca, tp);
}
int omv_scene_update_texture(omv_scene* s, int tex_id, const void* texels) {
    OMV_NULLCHK(s);
    return s->s.updateTexture(tex_id, texels);
}
int omv_scene_update_material_om_nk(omv_scene* s, int mat_id, const float n[3], const float k[3]) {
    OMV_NULLCHK(s); OMV_NULLCHK(n); OMV_NULLCHK(k);
    return s->s.updateMaterialOmNk(mat_id, n, k);
}
int omv_renderer_set_schedule(omv_renderer* r, int schedule) {
    OMV_NULLCHK(r);
    return r->r.setSchedule(schedule);
}
int omv_renderer_get_schedule(const omv_renderer* r, int* schedule) {
    OMV_NULLCHK(r); OMV_NULLCHK(schedule);
    *schedule = r->r.schedule();
    return OMV_OK;
}
int omv_renderer_set_tuning(omv_renderer* r, const omv_render_tuning* t) {
    OMV_NULLCHK(r); OMV_NULLCHK(t);
    return r->r.setTuning(*t);
}
int omv_renderer_get_tuning(const omv_renderer* r, omv_render_tuning* t) {
    OMV_NULLCHK(r); OMV_NULLCHK(t);
    *t = r->r.tuning();
    return OMV_OK;
}

int omv_batch_create(omv_batch** out) {
    OMV_NULLCHK(out);
    *out = new (std::nothrow) omv_batch();
    if (!*out) { omv::setError("omv_batch_create: out of memory"); return OMV_E_NOMEM; }
    return OMV_OK;
}
int omv_batch_destroy(omv_batch* b) { delete b; return OMV_OK; }
int omv_batch_add_job(omv_batch* b, const omv_scene* s, const omv_camera* c, const omv_renderer_params* p, int* job_id) {
    OMV_NULLCHK(b); OMV_NULLCHK(s); OMV_NULLCHK(c); OMV_NULLCHK(p);
    int rc = b->b.addJob(&s->s, c->c.desc, *p);
    if (rc == OMV_OK && job_id) *job_id = b->b.jobCount() - 1;
    return rc;
}
int omv_batch_job_count(const omv_batch* b, int* n) {
    OMV_NULLCHK(b); OMV_NULLCHK(n);
    *n = b->b.jobCount();
    return OMV_OK;
}
int omv_batch_set_tuning(omv_batch* b, const omv_batch_tuning* t) {
    OMV_NULLCHK(b); OMV_NULLCHK(t);
    return b->b.setTuning(*t);
}
int omv_batch_get_tuning(const omv_batch* b, omv_batch_tuning* t) {
    OMV_NULLCHK(b); OMV_NULLCHK(t);
    *t = b->b.tuning();
    return OMV_OK;
}
int omv_batch_autotune(omv_batch* b, int reps, omv_batch_tuning* best) {
    OMV_NULLCHK(b);
    return b->b.autotune(reps, best);
}
int omv_batch_run(omv_batch* b) {
    OMV_NULLCHK(b);
    return b->b.run();
}
int omv_batch_get_film(const omv_batch* b, int job_id, float* out, size_t cap) {
    OMV_NULLCHK(b); OMV_NULLCHK(out);
    return b->b.film(job_id, out, cap);
}
int omv_batch_get_stats(const omv_batch* b, omv_render_stats* out) {
    OMV_NULLCHK(b); OMV_NULLCHK(out);
    *out = b->b.stats();
    return OMV_OK;
}

int omv_fc_color_table(uint32_t seed, int subject, uint32_t size, uint32_t* out) {
    if (size && !out) { omv::setError("omv_fc_color_table: null argument"); return OMV_E_ARG; }
    std::vector<uint32_t> t;
    if (!omv::fcColorTable(seed, subject, size, t)) { omv::setError("omv_fc_color_table: more than 0xFFFFFF colours"); return OMV_E_RANGE; }
    if (size) std::memcpy(out, t.data(), size * sizeof(uint32_t));
    return OMV_OK;
}

// saveImage (main.cpp:86-120)
int omv_save_image(const char* filename, const float* rgba, int width, int height) {
    OMV_NULLCHK(filename); OMV_NULLCHK(rgba);
    if (width <= 0 || height <= 0) { omv::setError("omv_save_image: empty image"); return OMV_E_ARG; }
    return omv::saveImage(filename, rgba, width, height);
}
int omv_encode_image(const char* ext, const float* rgba, int width, int height, uint8_t* out, size_t cap, size_t* len) {
    OMV_NULLCHK(ext); OMV_NULLCHK(rgba); OMV_NULLCHK(len);
    if (width <= 0 || height <= 0) { omv::setError("omv_encode_image: empty image"); return OMV_E_ARG; }
    std::vector<uint8_t> bytes;
    const int rc = omv::encodeImage(ext, rgba, width, height, bytes);
    if (rc != OMV_OK) return rc;
    *len = bytes.size();
    if (!out) return OMV_OK;
    if (cap < bytes.size()) { omv::setError("omv_encode_image: buffer too small"); return OMV_E_RANGE; }
    std::memcpy(out, bytes.data(), bytes.size());
    return OMV_OK;
}

}  // extern "C"
