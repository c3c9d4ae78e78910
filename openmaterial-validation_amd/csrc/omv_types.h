// omv_types.h -- device-resident scene layout shared by the host builder
// (scene.cpp, bvh.cpp) and the gfx950 kernels (kernels.hip).
//
// Everything lives in HBM as flat pools addressed by 32-bit offsets:
//   vertex pool   interleaved per vertex: pos3 normal3 tangent4 uv2*ch
//                 (the reference's RenderMesh layout, RenderMesh.cpp:49-72),
//                 one gather region per hit vertex
//   face pool     uint32 x3 per face, by original primitive id
//   BVH4 nodes    128-byte nodes (one L2 line), child boxes stored SoA
//                 inside the node so one lane's fetch is 8 x 16-byte loads
//   tri records   3 x float4 per triangle in leaf order: v0.xyz+primId,
//                 v1.xyz, v2.xyz (object space) for the watertight test
//   instances, meshes, materials, textures, lights: small tables
#ifndef OMV_TYPES_H
#define OMV_TYPES_H

#include <stdint.h>

#define OMV_MAX_UV_CHANNELS 8
#define OMV_BVH_LEAF_MAX 2   // triangles per BVH4 leaf: 2 traced fastest of 1-4 (100k- and 2M-triangle scenes)
#define OMV_STACK_LDS 28      // per-lane traversal stack entries kept in LDS
#define OMV_STACK_TOTAL 128   // LDS + global overflow; the builder proves the bound
#define OMV_EMPTY_CHILD ((int32_t)0x80000000)
#define OMV_MAX_NODES ((size_t)1 << 25)   // node byte offsets (node * 128) stay below 2^32

enum { OMV_MAT_NONE = 0, OMV_MAT_OM = 1, OMV_MAT_PBR = 2, OMV_MAT_DIFFUSE = 3 };
enum { OMV_TEX_RGBA8 = 0, OMV_TEX_RGBA32F = 1 };
enum { OMV_BG_NONE = 0, OMV_BG_CONSTANT = 1, OMV_BG_HDR = 2 };

// Traversal stack entries, 32 bits: the kind in bits 31..30 and below it a
// BLAS or TLAS node index, an instance id, or for a BLAS leaf
// (first triangle record << 3) | (count - 1).  OmvNode4::entry holds each
// child's entry exactly as the traversal pushes it.
#define OMV_ENT_SHIFT 30
#define OMV_ENT_BLAS_NODE 0u
#define OMV_ENT_BLAS_LEAF 1u
#define OMV_ENT_TLAS_NODE 2u
#define OMV_ENT_INST 3u
#define OMV_ENT(kind, payload) (((uint32_t)(kind) << OMV_ENT_SHIFT) | (uint32_t)(payload))
#define OMV_ENT_END 0xFFFFFFFFu    // end of an instance's BLAS (stack only)
#define OMV_ENT_NONE 0xFFFFFFFEu   // empty child slot (never pushed)
static_assert(OMV_BVH_LEAF_MAX >= 1 && OMV_BVH_LEAF_MAX <= 8, "leaf size must fit the entry's 3-bit count");

struct alignas(64) OmvNode4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];   // child boxes; an empty slot's planes are all +inf
    int32_t child[4];    // >= 0 internal node; < 0 leaf: -(first+1) (tri record / instance id); empty: OMV_EMPTY_CHILD
    uint32_t entry[4];   // the child's stack entry (OMV_ENT), OMV_ENT_NONE for an empty slot
};
static_assert(sizeof(OmvNode4) == 128, "BVH4 node must be one 128-byte line");

struct OmvMesh {
    uint32_t vtx_offset;      // float offset of vertex 0
    uint32_t stride;          // floats per vertex (10 + 2*channels)
    uint32_t face_offset;     // uint32 offset of face 0 (3 per face)
    uint32_t num_faces;
    int32_t  mat_id;          // material slot after commit (missing -> last slot)
    int32_t  face_base;       // global face id of face 0 (FC primitive id)
    int32_t  blas_root;       // node index, -1 if the mesh is invalid
    int32_t  valid;
    int32_t  tex_off[OMV_MAX_UV_CHANNELS];  // uv float offset inside a vertex per glTF channel, -1 absent
};

struct OmvInstance {
    float xfm[12];   // object -> world, row-major 3x4 (reference m_transformation rows 0..2)
    float inv[12];   // world -> object, the double-precision inverse rounded to float
    float nrm[9];    // normal matrix: inverse of the transposed 3x3 (RenderMesh.h:36-43)
    int32_t mesh_id; // -1 = invalid instance
    int32_t blas_root;   // the mesh's BLAS root node (a copy of OmvMesh::blas_root: one dependent load less per instance entry)
    int32_t pad;
};

struct OmvMaterial {
    int32_t type;
    int32_t has_base;        // OM with IOR data
    int32_t alpha_mode;
    int32_t double_sided;
    float   n[3], k[3];      // OM per-RGB complex IOR (baked getIor)
    float   bcf[4];          // base colour factor
    float   ef[3];           // emissive factor
    float   metal, rough, normal_scale, occ_strength, ior, alpha_cutoff;
    int32_t emissive;        // emission reachable: |ef| > eps and emissive map valid
    int32_t occ_separate;    // occlusion map differs from metallic-roughness map
    int32_t tex_map[5];      // texture ids or -1
    int32_t tex_ch[5];       // uv channel per slot
    float   tex_xf[5][9];    // KHR_texture_transform per slot
    float   diffuse[4];      // DiffuseColorMaterial colour
};

struct OmvTexture {
    int32_t w, h, fmt, pad;
    uint64_t offset;         // byte offset in the texel pool
};

struct OmvLight {
    float pos[3];
    float intensity[3];
    float range;             // FLT_MAX = infinite (RenderLight::getAttenuationDistance)
    float power;
};

// Everything a kernel needs, passed by value.
struct OmvSceneView {
    const float*       vtx;
    const uint32_t*    faces;
    const OmvMesh*     meshes;
    const OmvInstance* insts;
    const OmvMaterial* mats;
    const OmvTexture*  texs;
    const uint8_t*     texels;
    const OmvNode4*    nodes;
    const float*       tris;      // 12 floats per record
    const OmvLight*    lights;
    const float*       light_cdf;
    const float*       bg;        // 4 floats per texel (rgb + pad)
    int32_t num_meshes, num_insts, num_mats, num_texs, num_lights;
    int32_t tlas_root;
    int32_t total_faces;
    int32_t fc_subject;
    int32_t any_emissive;
    int32_t any_masked;   // some PBR material uses alpha MASK (re-traced cut-outs): no occlusion early-out
    int32_t closest_only; // 1: visibility rays take the closest hit too (the occlusion early-out off)
    int32_t lights_leave; // 1: every light lies above each child box of the TLAS root on some axis, so every
                          // light ray (positive-octant directions) misses the scene and no light round has rays
    float   scene_lo[3], scene_inv[3];   // world box of the instances, 1 / its extent (0 if flat)
    int32_t bg_type, bg_w, bg_h;
    float   bg_const[3];
    float   bg_rot[9];
    float   bg_scale;
    float   light_max;
    float   min_density, max_density;
    // environment importance sampling tables (HDR background), else null
    const float*       env_row_prob;
    const int32_t*     env_row_alias;
    const float*       env_col_prob;
    const int32_t*     env_col_alias;
    const float*       env_tex_pdf;
    // false-colour table of an enumerated subject (pid, gid, mid, iid, ipid)
    // and, for ipid, the first instance-face index of every instance
    const uint32_t*    fc_table;
    const uint32_t*    fc_inst_base;
    uint32_t           fc_table_size;
};

#endif
