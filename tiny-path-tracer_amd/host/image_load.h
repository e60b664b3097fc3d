// Image files for image_texture (the reference loads them with the vendored
// stb_image, src/utils.cc:236-240): baseline sequential JPEG and 8-bit PPM
// (P6, P3), decoded to 8-bit RGB.
#ifndef TPT_HOST_IMAGE_LOAD_H
#define TPT_HOST_IMAGE_LOAD_H

#include <cstdint>
#include <string>
#include <vector>

namespace tpt {

// Decodes `path` into rgb (width*height*3 bytes, rows top first, as stb_image
// returns them).  Returns false with *err set for a missing file or an
// unsupported format (progressive or arithmetic-coded JPEG, 12-bit samples).
bool load_image_rgb8(const std::string &path, std::vector<uint8_t> *rgb, int *width, int *height, std::string *err);

}  // namespace tpt

#endif  // TPT_HOST_IMAGE_LOAD_H
