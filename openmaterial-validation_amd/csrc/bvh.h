// bvh.h -- host BVH builder: binned-SAH binary tree collapsed into 4-wide
// nodes (OmvNode4).  Replaces the Embree scene build of the reference
// (RenderScene::buildEmbreeScene, pathtracer/src/scene/RenderScene.cpp:225-272):
// one BLAS per mesh, one TLAS over the instances.
#ifndef OMV_BVH_H
#define OMV_BVH_H

#include <stdint.h>
#include <vector>

#include "omv_types.h"

namespace omv {

struct BuildBox { float lo[3], hi[3]; };

struct Bvh4Result {
    int root = -1;          // node index in the output pool
    int depth = 0;          // BVH4 levels on the longest root-to-leaf path
    std::vector<int> order; // primitive ids in leaf order
};

// Builds over `boxes` (one per primitive), appending nodes to `pool`.
// Leaves hold up to leafMax primitives; leaf slots encode
// child = -(leafBase + first + 1), where first indexes `order`.  Entries:
// a BLAS (tlas = false) gets OMV_ENT_BLAS_NODE / OMV_ENT_BLAS_LEAF
// ((leafBase + first) << 3 | count - 1); a TLAS (leafMax 1) gets
// OMV_ENT_TLAS_NODE / OMV_ENT_INST (leafBase + first, an index into `order`
// that the caller replaces by the instance id).
Bvh4Result buildBvh4(const std::vector<BuildBox>& boxes, int leafMax, int leafBase, std::vector<OmvNode4>& pool, bool tlas);

}  // namespace omv
#endif
