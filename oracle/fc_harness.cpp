/*
 * fc_harness.cpp -- calls the reference pathtracer's own false-colour
 * subjects (validation/FCSubject.h, built with FC_VALIDATION from the
 * sources where they lie, see oracle/build_ref.sh) on scripted arguments and
 * prints what they paint as JSON golden vectors.  Test infrastructure only;
 * nothing here ships.
 *
 * Subjects (FCSubject.h): PrimitiveIDFC / GeometryIDFC / MaterialIDFC
 * (EnumeratedFCSubject, :153-245), MetallicCftFC (:262-282), RoughnessCftFC
 * (:286-306), MeshDensityFC (:310-330), SurfaceGradientFC (:334-352),
 * MaterialNameFC with the dictionary RenderScene::setFalseColorSubject
 * installs (RenderScene.cpp:142), InvertedNormalFC (:459-480).
 */
#include <cstdio>
#include <vector>

#include "FCSubject.h"

static bool firstItem = true;
static void f(float v) { std::printf("%.9g", (double)v); }
static void col(const fvec3 c) { std::printf("["); f(c[0]); std::printf(","); f(c[1]); std::printf(","); f(c[2]); std::printf("]"); }
static void key(const char* k) { std::printf("%s\"%s\": ", firstItem ? "" : ",\n", k); firstItem = false; }

/* the colour table of an enumerated subject, read through a member pointer
 * named inside a derived class (the table is a protected member) */
struct TablePeek : EnumeratedFCSubject {
    static std::vector<Enumerator> table(const EnumeratedFCSubject& s) { return s.*(&TablePeek::m_vHashedColors); }
};

static void enumerated(const char* name, const EnumeratedFCSubject& s, const std::vector<unsigned>& ids) {
    key(name);
    std::printf("{\"table\": [");
    auto t = TablePeek::table(s);
    for (size_t i = 0; i < t.size(); ++i) std::printf("%s%u", i ? "," : "", t[i]);
    std::printf("], \"paints\": [");
    for (size_t i = 0; i < ids.size(); ++i) {
        fvec3 c;
        s.paint(GenericFCArgument(ids[i]), c);
        std::printf("%s[%u,", i ? "," : "", ids[i]);
        col(c);
        std::printf("]");
    }
    std::printf("]}");
}

int main() {
    std::printf("{\n");
    /* clamped-float subjects over values inside, at and outside [0, 1] */
    const float vals[] = {-1.0f, -1e-7f, 0.0f, 1e-7f, 0.125f, 0.25f, 1.0f / 3.0f, 0.5f, 0.7f, 0.9f, 0.999f, 1.0f, 1.0000001f, 2.5f};
    MetallicCftFC mmp;
    RoughnessCftFC rmp;
    MeshDensityFC md;
    SurfaceGradientFC sg;
    const GenericFCSubject* subj[4] = {&mmp, &rmp, &md, &sg};
    const char* names[4] = {"mmp", "rmp", "md", "sg"};
    for (int k = 0; k < 4; ++k) {
        key(names[k]);
        std::printf("[");
        for (size_t i = 0; i < sizeof(vals) / sizeof(vals[0]); ++i) {
            fvec3 c;
            subj[k]->paint(GenericFCArgument(vals[i]), c);
            std::printf("%s[", i ? "," : "");
            f(vals[i]);
            std::printf(",");
            col(c);
            std::printf("]");
        }
        std::printf("]");
    }
    /* the default (zero) argument a material hands a subject it does not
     * support (OpenMaterial.cpp:139, DiffuseColorMaterial.cpp:62-65) */
    key("default_argument");
    std::printf("{");
    for (int k = 0; k < 4; ++k) {
        fvec3 c;
        subj[k]->paint(GenericFCArgument(), c);
        std::printf("%s\"%s\": ", k ? ", " : "", names[k]);
        col(c);
    }
    std::printf("}");
    /* material names: the dictionary of RenderScene.cpp:142; "om" is
     * OpenMaterial's, "pbr" PhysicallyBasedMaterial's, "mm" the missing
     * material's (DiffuseColorMaterial.cpp:71) */
    MaterialNameFC mn({{"om", {0.0f, 1.0f, 0.0f}}, {"pbr", {0.0f, 0.0f, 1.0f}}});
    key("mn");
    std::printf("{");
    const char* mats[3] = {"om", "pbr", "mm"};
    for (int i = 0; i < 3; ++i) {
        char buf[4] = {0, 0, 0, 0};
        for (int j = 0; j < 3 && mats[i][j]; ++j) buf[j] = mats[i][j];
        fvec3 c;
        mn.paint(GenericFCArgument(buf), c);
        std::printf("%s\"%s\": ", i ? ", " : "", mats[i]);
        col(c);
    }
    std::printf("}");
    InvertedNormalFC in;
    key("in");
    std::printf("{");
    for (int b = 0; b < 2; ++b) {
        fvec3 c;
        in.paint(GenericFCArgument(b != 0), c);
        std::printf("%s\"%d\": ", b ? ", " : "", b);
        col(c);
    }
    std::printf("}");
    /* enumerated subjects: their (random) tables and paints of ids inside
     * and past the table; mid also of -1, the missing material */
    PrimitiveIDFC pid(7);
    GeometryIDFC gid(3);
    MaterialIDFC mid(4);
    enumerated("pid", pid, {0, 1, 2, 3, 4, 5, 6, 7, 8, 1000});
    enumerated("gid", gid, {0, 1, 2, 3, 4});
    enumerated("mid", mid, {0, 1, 2, 3, 4, 5, 0xFFFFFFFFu});
    std::printf("\n}\n");
    return 0;
}
