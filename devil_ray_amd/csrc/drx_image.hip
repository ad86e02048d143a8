// Framebuffer::save (rendering/framebuffer.cpp:44-51) through PNGEncoder::encode(const float32 *)
// (utils/png_encoder.cpp:56-95): float RGBA (4 floats per pixel, row 0 at the bottom of the image) ->
// 8-bit RGBA PNG with row 0 written last, each channel (uint8)(v * 255).  Host code only.
//
// The PNG is written with stored (uncompressed) deflate blocks, so the encoder needs no compression
// library; decoders read the same pixels as from dray's lodepng output (file bytes differ).
#include "drx_common.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace drx
{
namespace
{

struct CrcTable
{
  uint32_t t[256];
  CrcTable ()
  {
    for (uint32_t n = 0; n < 256; ++n)
    {
      uint32_t c = n;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[n] = c;
    }
  }
};

uint32_t crc32_of (const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu)
{
  static const CrcTable table;   // CRC-32 (PNG, ISO 3309), reflected polynomial 0xEDB88320
  for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
  return c;
}

void put_be32 (std::vector<uint8_t> &v, uint32_t x)
{
  v.push_back ((uint8_t)(x >> 24));
  v.push_back ((uint8_t)(x >> 16));
  v.push_back ((uint8_t)(x >> 8));
  v.push_back ((uint8_t)x);
}

void put_chunk (std::vector<uint8_t> &out, const char type[4], const std::vector<uint8_t> &data)
{
  put_be32 (out, (uint32_t)data.size ());
  const size_t start = out.size ();
  out.insert (out.end (), type, type + 4);
  out.insert (out.end (), data.begin (), data.end ());
  put_be32 (out, crc32_of (out.data () + start, out.size () - start) ^ 0xFFFFFFFFu);
}

// (uint8)(v * 255.f) as dray does; clamped first so out-of-range values stay defined.
inline uint8_t quantise (float v)
{
  float s = v * 255.f;
  if (!(s > 0.f)) return 0;
  if (s > 255.f) s = 255.f;
  return (uint8_t)s;
}

} // namespace

int encode_png (const float *rgba, int32_t w, int32_t h, std::vector<uint8_t> &out)
{
  // scanlines: filter byte 0 + w * 4 bytes, top row of the PNG = last row of the framebuffer
  const size_t row = 1 + 4 * (size_t)w;
  std::vector<uint8_t> raw (row * (size_t)h);
  for (int32_t y = 0; y < h; ++y)
  {
    uint8_t *d = raw.data () + row * (size_t)(h - 1 - y);
    d[0] = 0;
    const float *s = rgba + 4 * (size_t)w * (size_t)y;
    for (size_t i = 0; i < 4 * (size_t)w; ++i) d[1 + i] = quantise (s[i]);
  }
  // zlib stream of stored deflate blocks (RFC 1950/1951) + Adler-32
  std::vector<uint8_t> z;
  z.reserve (raw.size () + raw.size () / 65535 * 5 + 16);
  z.push_back (0x78);
  z.push_back (0x01);
  size_t pos = 0;
  do
  {
    const size_t len = std::min<size_t> (65535, raw.size () - pos);
    const bool last = pos + len == raw.size ();
    z.push_back (last ? 1 : 0);
    z.push_back ((uint8_t)len);
    z.push_back ((uint8_t)(len >> 8));
    z.push_back ((uint8_t)~len);
    z.push_back ((uint8_t)(~len >> 8));
    z.insert (z.end (), raw.begin () + pos, raw.begin () + pos + len);
    pos += len;
  } while (pos < raw.size ());
  uint32_t a = 1, b = 0;
  for (uint8_t x : raw)
  {
    a = (a + x) % 65521u;
    b = (b + a) % 65521u;
  }
  put_be32 (z, (b << 16) | a);

  static const uint8_t sig[8] = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n' };
  out.assign (sig, sig + 8);
  std::vector<uint8_t> ihdr;
  put_be32 (ihdr, (uint32_t)w);
  put_be32 (ihdr, (uint32_t)h);
  const uint8_t tail[5] = { 8, 6, 0, 0, 0 };   // 8 bits, RGBA, deflate, filter 0, no interlace
  ihdr.insert (ihdr.end (), tail, tail + 5);
  put_chunk (out, "IHDR", ihdr);
  put_chunk (out, "IDAT", z);
  put_chunk (out, "IEND", {});
  return DRX_OK;
}

} // namespace drx

using namespace drx;

extern "C" {

int drx_save_png (const float *rgba, int32_t width, int32_t height, const char *path)
{
  if (!rgba || !path || width <= 0 || height <= 0 || (int64_t)width * height > (int64_t (1) << 28))
    return fail (DRX_ERR_ARG, "drx_save_png: bad arguments");
  std::vector<uint8_t> png;
  encode_png (rgba, width, height, png);
  FILE *f = std::fopen (path, "wb");
  if (!f) return fail (DRX_ERR_ARG, "drx_save_png: cannot open %s", path);
  const size_t wrote = std::fwrite (png.data (), 1, png.size (), f);
  const int closed = std::fclose (f);
  if (wrote != png.size () || closed != 0) return fail (DRX_ERR_ARG, "drx_save_png: write to %s failed", path);
  return DRX_OK;
}

} // extern "C"
