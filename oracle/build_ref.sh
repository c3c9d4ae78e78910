#!/usr/bin/env bash
# Builds oracle/_ref/golden_harness from the reference pathtracer's own
# Embree-free sources, compiled where they lie (REF, default /root/reference).
# Outputs go only into oracle/_ref/.  The Embree-dependent translation units
# (OpenMaterial.cpp, PhysicallyBasedMaterial.cpp, RenderMesh.cpp,
# BDPTRenderer.cpp, RenderScene.cpp, main.cpp) are not built: Embree 4.1.0 is
# fetched at CMake configure time and is absent offline.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${REF:-/root/reference}"
SRC="$REF/pathtracer/src"
OUT="$HERE/_ref"
mkdir -p "$OUT"
if [ ! -d "$SRC" ]; then
  echo "reference sources not found at $SRC" >&2
  exit 2
fi
INC=(-I"$SRC" -I"$SRC/utilities" -I"$SRC/brdf" -I"$SRC/materials" -I"$SRC/scene" -I"$SRC/sensors"
     -I"$SRC/lights" -I"$SRC/backgrounds" -I"$SRC/tonemapping" -I"$SRC/assets" -I"$SRC/assets/materials"
     -I"$SRC/validation" -I"$REF/external" -I"$REF/external/json" -I"$REF/external/tinygltf"
     -I"$REF/external/doctest")
TUS=(materials/BitmapTexture.cpp backgrounds/HdrBackground.cpp sensors/PinholeCamera.cpp
     sensors/RenderSensor.cpp sensors/RenderRay.cpp lights/RenderLightPoint.cpp lights/RenderLight.cpp
     tonemapping/ToneMapping.cpp assets/materials/AssetMaterialIor.cpp assets/AssetDefs.cpp
     utilities/ReferenceLink.cpp utilities/Utils.cpp utilities/Random.cpp)
SRCS=()
for t in "${TUS[@]}"; do SRCS+=("$SRC/$t"); done
# -O0 and no FMA contraction: the reference's own CMake default for GNU is -O0
# BitmapTexture::texture(const SurfacePoint*,...) refers to SurfacePoint::getTexCoords, which
# lives in an Embree-dependent TU; the harness never calls that overload, so
# unused sections are dropped at link time.
g++ -std=c++11 -O0 -ffp-contract=off -ffunction-sections -Wl,--gc-sections -DDOCTEST_CONFIG_DISABLE "${INC[@]}" \
    "$HERE/golden_harness.cpp" "$HERE/golden_stb.cpp" "${SRCS[@]}" -o "$OUT/golden_harness"
echo "built $OUT/golden_harness"

# The false-colour subjects (validation/FCSubject.h, header-only paints)
# with FC_VALIDATION, the reference's own validation build flag.
g++ -std=c++11 -O0 -ffp-contract=off -DFC_VALIDATION -DDOCTEST_CONFIG_DISABLE "${INC[@]}" \
    "$HERE/fc_harness.cpp" "$SRC/utilities/Random.cpp" -o "$OUT/fc_harness"
echo "built $OUT/fc_harness"

# The glTF / OpenMATERIAL asset loader (AssetGeometry and what it pulls in);
# the reference's external/tiny_gltf.cpp is its tinygltf + stb_image TU.
ATUS=(assets/AssetGeometry.cpp assets/ImageData.cpp assets/Mesh.cpp assets/AssetDefs.cpp
      assets/materials/AssetMaterial.cpp assets/materials/PbrMaterial.cpp assets/materials/AssetMaterialIor.cpp
      utilities/ReferenceLink.cpp utilities/Utils.cpp utilities/Random.cpp utilities/ImageSaver.cpp lights/PointLight.cpp)
ASRCS=("$REF/external/tiny_gltf.cpp")
for t in "${ATUS[@]}"; do ASRCS+=("$SRC/$t"); done
g++ -std=c++11 -O1 -ffp-contract=off -DDOCTEST_CONFIG_DISABLE -Dcimg_display=0 "${INC[@]}" -I"$REF/external/CImg" \
    "$HERE/asset_harness.cpp" "${ASRCS[@]}" -lpthread -o "$OUT/asset_harness"
echo "built $OUT/asset_harness"
