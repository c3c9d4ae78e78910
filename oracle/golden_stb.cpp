/* The stb_image implementation the reference's HDRBackground links against
 * (the reference compiles it inside external/tiny_gltf.cpp); compiled from
 * the reference's vendored header for the golden harness only. */
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"
