// ref_stb.cpp - exposes the reference loader's colour-to-grey conversion
// (stbi__convert_format of include/stb_image.h in
// carbotaniuman/sycl-apriltag-test, the routine behind
// stbi_load(..., STBI_grey) at src/main.cpp:265) through a C ABI, so that
// tests can check the oracle's conversion against the reference's own code.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/ with
// the reference's include directory on the include path; nothing of the
// reference is copied here. Only tests/ load the result.
#include <cstdint>
#include <cstdlib>
#include <cstring>

#define STB_IMAGE_IMPLEMENTATION
#define STBI_NO_STDIO
#include "stb_image.h"

extern "C" {

// n pixels of `channels` bytes (3 or 4) -> n grey bytes, as the loader
// converts a decoded image to one channel.
int ref_stbi_grey(const uint8_t *pixels, int n, int channels, uint8_t *out) {
    unsigned char *copy = static_cast<unsigned char *>(std::malloc((size_t)n * channels));
    if (!copy) return -1;
    std::memcpy(copy, pixels, (size_t)n * channels);
    unsigned char *g = stbi__convert_format(copy, channels, 1, (unsigned)n, 1u);  // frees copy
    if (!g) return -1;
    std::memcpy(out, g, (size_t)n);
    std::free(g);
    return 0;
}

}  // extern "C"
