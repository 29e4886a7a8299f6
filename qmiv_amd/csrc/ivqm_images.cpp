// PNG / BMP frame readers of the command-line front end (see ivqm_images.h).
#include "ivqm_images.h"

#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <cstdlib>
#include <vector>

namespace ivqm {
namespace images {

namespace {
uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t le16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

bool read_file(const std::string& file, std::vector<uint8_t>& out) {
  FILE* f = fopen(file.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  const bool ok = n >= 0 && fread(out.data(), 1, out.size(), f) == out.size();
  fclose(f);
  return ok;
}
bool exists(const std::string& file) {
  FILE* f = fopen(file.c_str(), "rb");
  if (f) fclose(f);
  return f != nullptr;
}
// PNG filter predictor (PNG specification, section 9.4)
uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
  return (uint8_t)(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}
}  // namespace

std::string format_name(const std::string& pattern, int idx) {
  const size_t a = pattern.find('{'), b = a == std::string::npos ? a : pattern.find('}', a);
  if (a == std::string::npos || b == std::string::npos) return pattern;
  std::string spec = pattern.substr(a + 1, b - a - 1);  // "", ":d", ":04d", ":4"
  int width = 0;
  bool zero = false;
  if (!spec.empty() && spec[0] == ':') {
    size_t k = 1;
    if (k < spec.size() && spec[k] == '0') { zero = true; k++; }
    while (k < spec.size() && spec[k] >= '0' && spec[k] <= '9') width = width * 10 + (spec[k++] - '0');
  }
  std::string num = std::to_string(idx);
  if ((int)num.size() < width) num = std::string(width - num.size(), zero ? '0' : ' ') + num;
  return pattern.substr(0, a) + num + pattern.substr(b + 1);
}
bool has_field(const std::string& pattern) { return format_name(pattern, 0) != pattern; }

int count_frames(const std::string& pattern, int* first_idx) {
  *first_idx = 0;
  if (!has_field(pattern)) return exists(pattern) ? 1 : 0;
  int first = -1;
  for (int i = 0; i <= 1 && first < 0; i++)
    if (exists(format_name(pattern, i))) first = i;
  if (first < 0) return 0;
  int n = 0;
  while (n < 65535 && exists(format_name(pattern, first + n))) n++;
  *first_idx = first;
  return n;
}

std::string read_png(const std::string& file, int w, int h, uint8_t* rgb) {
  std::vector<uint8_t> d;
  if (!read_file(file, d)) return "cannot read " + file;
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  if (d.size() < 8 || memcmp(d.data(), sig, 8) != 0) return "not a PNG file: " + file;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte;
  for (size_t p = 8; p + 12 <= d.size();) {
    const uint32_t len = be32(&d[p]);
    if (p + 12 + (size_t)len > d.size()) return "truncated PNG chunk: " + file;
    const uint8_t* t = &d[p + 4];
    const uint8_t* c = &d[p + 8];
    if (!memcmp(t, "IHDR", 4) && len >= 13) {
      W = be32(c);
      H = be32(c + 4);
      depth = c[8];
      ctype = c[9];
      interlace = c[12];
    } else if (!memcmp(t, "PLTE", 4)) {
      plte.assign(c, c + len);
    } else if (!memcmp(t, "IDAT", 4)) {
      idat.insert(idat.end(), c, c + len);
    } else if (!memcmp(t, "IEND", 4)) {
      break;
    }
    p += 12 + (size_t)len;
  }
  if ((int)W != w || (int)H != h) return "PNG size does not match: " + file;
  if (depth != 8) return "only 8-bit PNG is supported: " + file;
  if (interlace != 0 && interlace != 1) return "unknown PNG interlace method: " + file;
  const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
  if (!ch) return "unsupported PNG colour type: " + file;
  if (ctype == 3 && plte.size() < 3) return "PNG palette missing: " + file;
  // Passes: the whole image, or the seven Adam7 sub-images (PNG specification,
  // section 8.2): pass p holds the pixels at (x0 + i*dx, y0 + j*dy), each pass
  // stored as its own filtered image; an empty pass stores nothing.
  static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  static const int kWhole[1][4] = {{0, 0, 1, 1}};
  const int npass = interlace ? 7 : 1;
  const int(*pass)[4] = interlace ? kAdam7 : kWhole;
  size_t total = 0;
  for (int p = 0; p < npass; p++) {
    const size_t pw = ((size_t)w + pass[p][2] - 1 - pass[p][0]) / pass[p][2];
    const size_t ph = ((size_t)h + pass[p][3] - 1 - pass[p][1]) / pass[p][3];
    if (pw && ph) total += (pw * ch + 1) * ph;
  }
  std::vector<uint8_t> raw(total);
  uLongf rawlen = (uLongf)raw.size();
  if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size())
    return "PNG data does not inflate to the image size: " + file;
  const size_t plane = (size_t)w * h;
  size_t pos = 0;
  for (int p = 0; p < npass; p++) {
    const int x0 = pass[p][0], y0 = pass[p][1], dx = pass[p][2], dy = pass[p][3];
    const size_t pw = ((size_t)w + dx - 1 - x0) / dx, ph = ((size_t)h + dy - 1 - y0) / dy;
    if (!pw || !ph) continue;
    const size_t stride = pw * ch;
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    for (size_t j = 0; j < ph; j++, pos += stride + 1) {
      const uint8_t f = raw[pos];
      const uint8_t* s = &raw[pos + 1];
      for (size_t i = 0; i < stride; i++) {
        const int a = i >= (size_t)ch ? cur[i - ch] : 0, b = prev[i], c = i >= (size_t)ch ? prev[i - ch] : 0;
        int v = s[i];
        if (f == 1) v += a;
        else if (f == 2) v += b;
        else if (f == 3) v += (a + b) >> 1;
        else if (f == 4) v += paeth(a, b, c);
        else if (f != 0) return "bad PNG filter: " + file;
        cur[i] = (uint8_t)v;
      }
      const size_t y = y0 + j * dy;
      for (size_t i = 0; i < pw; i++) {
        uint8_t r, g, bl;
        const uint8_t* px = &cur[i * ch];
        if (ctype == 0 || ctype == 4) r = g = bl = px[0];
        else if (ctype == 3) {
          const size_t k = (size_t)px[0] * 3;
          if (k + 2 >= plte.size()) return "PNG palette index out of range: " + file;
          r = plte[k]; g = plte[k + 1]; bl = plte[k + 2];
        } else { r = px[0]; g = px[1]; bl = px[2]; }
        const size_t o = y * w + x0 + i * dx;
        rgb[o] = r;
        rgb[plane + o] = g;
        rgb[2 * plane + o] = bl;
      }
      prev.swap(cur);
    }
  }
  return "";
}

std::string read_bmp(const std::string& file, int w, int h, uint8_t* rgb) {
  std::vector<uint8_t> d;
  if (!read_file(file, d)) return "cannot read " + file;
  if (d.size() < 54 || le16(&d[0]) != 0x4D42) return "BitmapFileHeader content is invalid: " + file;
  const uint32_t off = le32(&d[10]);
  const int32_t W = (int32_t)le32(&d[18]), H = (int32_t)le32(&d[22]);
  const int planes = le16(&d[26]), bpp = le16(&d[28]);
  const uint32_t comp = le32(&d[30]);
  if (off < 54 || planes != 1) return "BitmapInfoHeader content is invalid: " + file;
  if (bpp != 24 && bpp != 32) return "Only 24 or 32 bit (non-palette) is suported: " + file;
  if (comp != 0) return "Only BI_RGB compression method is supported: " + file;
  const int AbsH = H < 0 ? -H : H, nc = bpp == 32 ? 4 : 3;
  if (W != w || AbsH != h) return "BMP size does not match: " + file;
  const size_t line = ((size_t)W * nc + 3) / 4 * 4, plane = (size_t)w * h;
  if ((size_t)off + line * AbsH > d.size()) return "BMP row read error: " + file;
  for (int r = 0; r < AbsH; r++) {
    const int y = H > 0 ? AbsH - 1 - r : r;  // bottom-up unless the height is negative
    const uint8_t* s = &d[off + line * r];
    for (int x = 0; x < w; x++) {
      const size_t o = (size_t)y * w + x;
      rgb[2 * plane + o] = s[x * nc + 0];  // B
      rgb[plane + o] = s[x * nc + 1];      // G
      rgb[o] = s[x * nc + 2];              // R
    }
  }
  return "";
}

}  // namespace images
}  // namespace ivqm
