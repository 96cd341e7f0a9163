// image_io.cpp - render output files (the reference shows its image only in
// a GL window and writes no file).
//
//   PFM  the float framebuffer: "PF", width height, scale -1 (little-endian),
//        rows bottom to top, which is the order of the library's rgb buffers
//        (row 0 = bottom, Render.cu:220-222).
//   PNG  the reference's displayed bytes (pic2RGBW, Render.cu:462-481, and the
//        STATIC / NORMAL / DEPTH cases of Render.cu:232-258):
//        uchar(254.99 * clamp01(sqrt(x))) with gamma, uchar(254.99 * clamp01(x))
//        without; 8-bit RGBA, alpha 255, written top row first.  The zlib
//        stream uses stored (uncompressed) deflate blocks, so the writer needs
//        no compressor: CRC-32 (ISO 3309 polynomial 0xEDB88320) per chunk and
//        Adler-32 over the raw scanlines, per RFC 1950/1951 and the PNG spec.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../../include/cpt.h"

namespace {

uint32_t crcTable[256];
bool crcReady = false;

void crcInit() {
  for (uint32_t n = 0; n < 256; n++) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    crcTable[n] = c;
  }
  crcReady = true;
}

uint32_t crc32(uint32_t crc, const uint8_t *p, size_t n) {
  crc = ~crc;
  for (size_t i = 0; i < n; i++) crc = crcTable[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
  return ~crc;
}

void put32be(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

// length, type, data, CRC over type + data
void chunk(std::vector<uint8_t> &out, const char type[4], const std::vector<uint8_t> &data) {
  put32be(out, (uint32_t)data.size());
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  put32be(out, crc32(0, out.data() + at, 4 + data.size()));
}

uint8_t displayByte(float v, bool gamma) {
  double x = (double)v;
  if (gamma) x = std::sqrt(x);
  x = x > 1.0 ? 1.0 : (x < 0.0 ? 0.0 : x);
  if (!(x == x)) x = 0.0;  // NaN (sqrt of a negative) shows black
  return static_cast<uint8_t>(254.99 * x);
}

}  // namespace

extern "C" {

int cpt_write_pfm(const char *path, const float *rgb, int32_t w, int32_t h) {
  if (!path || !rgb || w <= 0 || h <= 0) return CPT_ERR_ARG;
  FILE *f = std::fopen(path, "wb");
  if (!f) return CPT_ERR_IO;
  bool ok = std::fprintf(f, "PF\n%d %d\n-1.0\n", (int)w, (int)h) > 0;
  // the library's rows already run bottom to top, as PFM stores them; the
  // floats are written little-endian (scale -1)
  const size_t n = (size_t)w * h * 3;
  const uint16_t probe = 1;
  if (*(const uint8_t *)&probe == 1) {
    ok = ok && std::fwrite(rgb, sizeof(float), n, f) == n;
  } else {
    std::vector<uint8_t> b(n * 4);
    for (size_t i = 0; i < n; i++) {
      uint32_t u;
      std::memcpy(&u, rgb + i, 4);
      for (int k = 0; k < 4; k++) b[4 * i + k] = (uint8_t)(u >> (8 * k));
    }
    ok = ok && std::fwrite(b.data(), 1, b.size(), f) == b.size();
  }
  ok = std::fclose(f) == 0 && ok;
  return ok ? CPT_OK : CPT_ERR_IO;
}

int cpt_write_png(const char *path, const float *rgb, int32_t w, int32_t h, int32_t gamma) {
  if (!path || !rgb || w <= 0 || h <= 0 || (int64_t)w * 4 + 1 > 0x7fffffff) return CPT_ERR_ARG;
  if (!crcReady) crcInit();
  // raw scanlines, filter byte 0, top row (the library's last row) first
  const size_t stride = (size_t)w * 4 + 1;
  std::vector<uint8_t> raw(stride * (size_t)h);
  for (int32_t r = 0; r < h; r++) {
    uint8_t *line = raw.data() + stride * (size_t)r;
    line[0] = 0;
    const float *src = rgb + (size_t)(h - 1 - r) * w * 3;
    for (int32_t x = 0; x < w; x++) {
      for (int c = 0; c < 3; c++) line[1 + 4 * x + c] = displayByte(src[3 * x + c], gamma != 0);
      line[1 + 4 * x + 3] = 255;
    }
  }
  // zlib stream: header (deflate, 32K window, no dictionary, check bits),
  // stored blocks of at most 65535 bytes, Adler-32 of the raw data
  std::vector<uint8_t> z;
  z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
  z.push_back(0x78);
  z.push_back(0x01);
  size_t pos = 0;
  do {
    const size_t len = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
    const bool last = pos + len == raw.size();
    z.push_back(last ? 1 : 0);  // BFINAL, BTYPE = 00 (stored)
    z.push_back((uint8_t)(len & 0xff));
    z.push_back((uint8_t)(len >> 8));
    z.push_back((uint8_t)(~len & 0xff));
    z.push_back((uint8_t)((~len >> 8) & 0xff));
    z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
    pos += len;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;
  for (uint8_t v : raw) {
    a = (a + v) % 65521u;
    b = (b + a) % 65521u;
  }
  put32be(z, (b << 16) | a);

  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  put32be(ihdr, (uint32_t)w);
  put32be(ihdr, (uint32_t)h);
  ihdr.push_back(8);  // bit depth
  ihdr.push_back(6);  // colour type: RGBA
  ihdr.push_back(0);  // compression: deflate
  ihdr.push_back(0);  // filter method
  ihdr.push_back(0);  // no interlace
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  FILE *f = std::fopen(path, "wb");
  if (!f) return CPT_ERR_IO;
  bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  ok = std::fclose(f) == 0 && ok;
  return ok ? CPT_OK : CPT_ERR_IO;
}

}  // extern "C"
