// Image-list input of the QMIV command-line front end (-ff PNG / BMP): the
// reference reads one image file per frame through xSeqPNG / xSeqBMP
// (libs/pmbb-slst/src/xSeqLST.cpp); pictures are 8-bit RGB, stored here as
// three W x H planes R, G, B (the packed layout xSeqLST.cpp:224-238 writes).
#pragma once
#include <stdint.h>

#include <string>

namespace ivqm {
namespace images {

// File name of frame `idx` from a pattern with one integer field ("{}",
// "{:d}", "{:04d}", "{:4}"), as fmt::format(pattern, idx) gives it
// (xSeqLST.h:46). A pattern without a field names one file.
std::string format_name(const std::string& pattern, int idx);
bool has_field(const std::string& pattern);

// Sequence of image files: the first index is 0 or 1 (the first that
// exists), then every consecutive index that exists (xSeqLST.cpp:92-127).
// Returns the number of frames (0: none found) and the first index.
int count_frames(const std::string& pattern, int* first_idx);

// Decodes one file into planes R, G, B (each w x h bytes). Returns an empty
// string on success, else the reason.
//   PNG: colour types gray, gray + alpha, RGB, RGBA and palette at 8 bits per
//        sample, not interlaced or Adam7-interlaced; alpha is dropped (as a
//        decode to RGB8 does).
//   BMP: BI_RGB 24 or 32 bits per pixel, bottom-up or top-down rows padded
//        to 4 bytes (xSeqLST.cpp:337-415).
std::string read_png(const std::string& file, int w, int h, uint8_t* rgb);
std::string read_bmp(const std::string& file, int w, int h, uint8_t* rgb);

}  // namespace images
}  // namespace ivqm
