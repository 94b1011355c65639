#include "ImageUtils.h"

#include <algorithm>
#include <cctype>
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace beams
{
namespace io
{
namespace
{
uint32_t Crc32(const uint8_t* data, size_t n, uint32_t crc = 0)
{
  // CRC-32 (ISO-HDLC, reflected polynomial 0xEDB88320), as PNG chunks carry it; the table
  // is a function-local static, built once even when frames are saved from several threads
  static const std::array<uint32_t, 256> table = [] {
    std::array<uint32_t, 256> t{};
    for (uint32_t i = 0; i < 256; ++i)
    {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k)
        c = (c & 1u) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
      t[i] = c;
    }
    return t;
  }();
  crc = ~crc;
  for (size_t i = 0; i < n; ++i)
    crc = table[(crc ^ data[i]) & 0xFFu] ^ (crc >> 8);
  return ~crc;
}

uint32_t Adler32(const std::vector<uint8_t>& data)
{
  uint32_t a = 1, b = 0;
  for (uint8_t v : data)
  {
    a = (a + v) % 65521u;
    b = (b + a) % 65521u;
  }
  return (b << 16) | a;
}

void PutU32(std::vector<uint8_t>& out, uint32_t v)
{
  out.push_back(static_cast<uint8_t>(v >> 24));
  out.push_back(static_cast<uint8_t>(v >> 16));
  out.push_back(static_cast<uint8_t>(v >> 8));
  out.push_back(static_cast<uint8_t>(v));
}

void PutChunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data)
{
  PutU32(out, static_cast<uint32_t>(data.size()));
  std::vector<uint8_t> body(type, type + 4);
  body.insert(body.end(), data.begin(), data.end());
  out.insert(out.end(), body.begin(), body.end());
  PutU32(out, Crc32(body.data(), body.size()));
}

bool Ends(const std::string& s, const std::string& suffix)
{
  if (s.size() < suffix.size())
    return false;
  std::string tail = s.substr(s.size() - suffix.size());
  std::transform(tail.begin(), tail.end(), tail.begin(),
                 [](unsigned char ch) { return static_cast<char>(std::tolower(ch)); });
  return tail == suffix;
}
} // namespace

uint8_t QuantizeChannel(float v)
{
  const float c = std::isnan(v) ? 0.0f : std::min(std::max(v, 0.0f), 1.0f);
  return static_cast<uint8_t>(std::lround(c * 255.0f));
}

WriteResult SaveImage(const std::vector<float>& rgba, int width, int height,
                      const std::string& fileName)
{
  WriteResult r;
  if (width <= 0 || height <= 0 || rgba.size() != static_cast<size_t>(width) * height * 4)
  {
    r.Error = "SaveImage: need width*height*4 colour values";
    return r;
  }
  std::vector<uint8_t> file;
  if (Ends(fileName, ".png"))
  {
    // scanlines: filter byte 0 then RGBA8
    std::vector<uint8_t> raw;
    raw.reserve(static_cast<size_t>(height) * (1 + 4 * static_cast<size_t>(width)));
    for (int y = 0; y < height; ++y)
    {
      raw.push_back(0);
      for (int x = 0; x < width; ++x)
        for (int c = 0; c < 4; ++c)
          raw.push_back(QuantizeChannel(rgba[(static_cast<size_t>(y) * width + x) * 4 + c]));
    }
    // zlib stream of stored (uncompressed) deflate blocks of at most 65535 bytes
    std::vector<uint8_t> z = { 0x78, 0x01 };
    size_t pos = 0;
    do
    {
      const size_t n = std::min<size_t>(65535, raw.size() - pos);
      const bool last = pos + n == raw.size();
      z.push_back(last ? 1 : 0);
      z.push_back(static_cast<uint8_t>(n & 0xFF));
      z.push_back(static_cast<uint8_t>(n >> 8));
      z.push_back(static_cast<uint8_t>(~n & 0xFF));
      z.push_back(static_cast<uint8_t>((~n >> 8) & 0xFF));
      z.insert(z.end(), raw.begin() + static_cast<std::ptrdiff_t>(pos),
               raw.begin() + static_cast<std::ptrdiff_t>(pos + n));
      pos += n;
    } while (pos < raw.size());
    PutU32(z, Adler32(raw));
    const uint8_t sig[8] = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n' };
    file.assign(sig, sig + 8);
    std::vector<uint8_t> ihdr;
    PutU32(ihdr, static_cast<uint32_t>(width));
    PutU32(ihdr, static_cast<uint32_t>(height));
    ihdr.push_back(8); // bit depth
    ihdr.push_back(6); // RGBA
    ihdr.push_back(0); // deflate
    ihdr.push_back(0); // adaptive filtering (filter type 0 on every line)
    ihdr.push_back(0); // no interlace
    PutChunk(file, "IHDR", ihdr);
    PutChunk(file, "IDAT", z);
    PutChunk(file, "IEND", {});
  }
  else if (Ends(fileName, ".ppm") || Ends(fileName, ".pnm"))
  {
    const std::string header = "P6\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
    file.assign(header.begin(), header.end());
    for (size_t i = 0; i < static_cast<size_t>(width) * height; ++i)
      for (int c = 0; c < 3; ++c)
        file.push_back(QuantizeChannel(rgba[i * 4 + c]));
  }
  else
  {
    r.Error = "SaveImage: unknown image type (use .png, .ppm or .pnm): " + fileName;
    return r;
  }
  std::ofstream f(fileName, std::ios::binary);
  if (!f.write(reinterpret_cast<const char*>(file.data()), static_cast<std::streamsize>(file.size())))
  {
    r.Error = "SaveImage: cannot write " + fileName;
    return r;
  }
  r.Success = true;
  return r;
}

std::string FrameFileName(const std::string& sceneId, int rank, int num, const std::string& extension)
{
  std::stringstream ss;
  ss << sceneId << "_" << rank << "_" << std::setw(3) << std::setfill('0') << num << extension;
  return ss.str();
}
} // namespace io
} // namespace beams
