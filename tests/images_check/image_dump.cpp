// Host-only check program for the front end's image readers
// (qmiv_amd/csrc/ivqm_images.cpp): decodes one PNG or BMP file of a given size
// and writes its R, G and B planes (w*h bytes each) to stdout, or the reader's
// message to stderr with exit status 1.
//   image_dump PNG|BMP file width height
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../qmiv_amd/csrc/ivqm_images.h"

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: image_dump PNG|BMP file width height\n");
    return 2;
  }
  const int w = atoi(argv[3]), h = atoi(argv[4]);
  if (w <= 0 || h <= 0) return 2;
  std::vector<uint8_t> rgb((size_t)w * h * 3, 0);
  const std::string err = strcmp(argv[1], "BMP") == 0 ? ivqm::images::read_bmp(argv[2], w, h, rgb.data())
                                                      : ivqm::images::read_png(argv[2], w, h, rgb.data());
  if (!err.empty()) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  fwrite(rgb.data(), 1, rgb.size(), stdout);
  return 0;
}
