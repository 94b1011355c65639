// Checks of the C++ host surface (beams::rendering::* over the C ABI).
//   test_host cpu   config parsing, DataSetFactory, error paths (no GPU needed)
//   test_host gpu   SceneFactory -> Scene::Ready -> RenderFrame, and Render(Ray) against
//                   RenderImage on the same camera rays (bit equality)
#include "../../include/beams_mi355x.h"
#include "beams/SceneFactory.h"
#include "beams/io/DataSetUtils.h"
#include "beams/io/ImageUtils.h"
#include "beams/rendering/ColorTable.h"

#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

static int failures = 0;
#define CHECK(cond)                                                     \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

static const char* kToml = R"(
default_preset = "small"   # comment
[presets.small]
dataset = "blobs"
size = 24
seed = 7
blobs = 8
dtype = "f32"
map_size = [23, 23, 23]
camera_azimuth = 35.0
camera_elevation = 12.5
light_type = "point"
light_position = [2.2, 1.7, -0.8]
light_color = [1, 1, 1]
light_intensity = 2.0
colormap_seed = 8
image = [96, 64]
sample_distance = 0.0217
[presets.half]
dataset = "blobs"
size = 16
dtype = "bf16"
map_enabled = false
)";

static std::string TempPath(const std::string& tag)
{
  return "/tmp/beams_io_" + std::to_string(getpid()) + "_" + tag;
}

static bool SameDataSet(const beams::DataSet& a, const beams::DataSet& b)
{
  if (a.CellSet.PointDims != b.CellSet.PointDims || a.Fields.size() != b.Fields.size())
    return false;
  for (int d = 0; d < 3; ++d)
    if (a.Coords.Axes[d] != b.Coords.Axes[d])
      return false;
  for (size_t k = 0; k < a.Fields.size(); ++k)
    if (a.Fields[k].Name != b.Fields[k].Name || a.Fields[k].Assoc != b.Fields[k].Assoc ||
        a.Fields[k].Values != b.Fields[k].Values)
      return false;
  return true;
}

// VTK legacy files and chunking (pilot/io/DataSetUtils.cxx, convert/Chunk.cxx)
static void test_io(const beams::DataSet& ds, const beams::DataSet& bf16)
{
  namespace io = beams::io;
  for (auto ft : { io::FileType::BINARY, io::FileType::ASCII }) {
    const std::string fn = TempPath(ft == io::FileType::BINARY ? "bin.vtk" : "asc.vtk");
    CHECK(io::DataSetUtils::Write(ds, fn, ft).Success);
    const io::ReadResult r = io::DataSetUtils::Read(fn);
    CHECK(r.Success);
    CHECK(SameDataSet(ds, r.Value));
    std::remove(fn.c_str());
  }
  // rectilinear coordinates and a cell field
  beams::DataSet rect = ds;
  for (size_t i = 0; i < rect.Coords.Axes[1].size(); ++i)
    rect.Coords.Axes[1][i] = 0.5f * (float)(i * i) / 23.0f;
  beams::Field cf;
  cf.Name = "cellid";
  cf.Assoc = beams::Association::Cells;
  for (int i = 0; i < 23 * 23 * 23; ++i) cf.Values.push_back((float)i * 0.25f);
  rect.Fields.push_back(cf);
  const std::string fr = TempPath("rect.vtk");
  CHECK(io::DataSetUtils::Write(rect, fr).Success);
  {
    std::ifstream in(fr);
    std::string l1, l2, l3, l4;
    std::getline(in, l1), std::getline(in, l2), std::getline(in, l3), std::getline(in, l4);
    CHECK(l4 == "DATASET RECTILINEAR_GRID");
  }
  const io::ReadResult rr = io::DataSetUtils::Read(fr);
  CHECK(rr.Success && SameDataSet(rect, rr.Value));
  std::remove(fr.c_str());
  // BF16 storage is written as its decoded values
  const std::string fb = TempPath("bf16.vtk");
  CHECK(io::DataSetUtils::Write(bf16, fb).Success);
  const io::ReadResult rb = io::DataSetUtils::Read(fb);
  CHECK(rb.Success && rb.Value.Fields.size() == 1);
  bool same = rb.Success;
  for (size_t i = 0; same && i < bf16.Fields[0].Half.size(); ++i) {
    const uint32_t bits = (uint32_t)bf16.Fields[0].Half[i] << 16;
    float v;
    std::memcpy(&v, &bits, 4);
    same = rb.Value.Fields[0].Values[i] == v;
  }
  CHECK(same);
  std::remove(fb.c_str());
  // a file written by hand: ASCII, SPACING, FIELD data skipped, unsigned_char scalars
  const std::string fh = TempPath("hand.vtk");
  {
    std::ofstream os(fh);
    os << "# vtk DataFile Version 2.0\nhand\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS 2 2 2\n"
          "ORIGIN 1 2 3\nSPACING 0.5 0.5 2\nPOINT_DATA 8\nFIELD FieldData 1\nlabels 1 8 int\n"
          "0 1 2 3 4 5 6 7\nSCALARS density unsigned_char\nLOOKUP_TABLE default\n"
          "0 10 20 30 40 50 60 255\n";
  }
  const io::ReadResult rh = io::DataSetUtils::Read(fh);
  CHECK(rh.Success);
  CHECK(rh.Success && rh.Value.Coords.Axes[2][1] == 5.0f && rh.Value.Fields[0].Values[7] == 255.0f);
  std::remove(fh.c_str());
  // errors
  CHECK(!io::DataSetUtils::Read(TempPath("missing.vtk")).Success);
  const std::string fe = TempPath("poly.vtk");
  {
    std::ofstream os(fe);
    os << "# vtk DataFile Version 3.0\nx\nASCII\nDATASET POLYDATA\nPOINTS 0 float\n";
  }
  const io::ReadResult re = io::DataSetUtils::Read(fe);
  CHECK(!re.Success && re.Error.find("POLYDATA") != std::string::npos);
  std::remove(fe.c_str());
  // chunking: 2 x 3 x 1 blocks sharing their boundary planes, values copied exactly
  const auto ranges = io::ChunkPointRanges(ds.CellSet.PointDims, 2, 3, 1);
  CHECK(ranges.size() == 6);
  CHECK(ranges[0][0] == 0 && ranges[0][1] == 11 && ranges[1][0] == 11 && ranges[1][1] == 23);
  CHECK(ranges[5][2] == 15 && ranges[5][3] == 23 && ranges[5][5] == 23);
  const auto chunks = io::ChunkDataSet(ds, 2, 3, 1);
  CHECK(chunks.size() == 6);
  const auto& c5 = chunks[5];
  const auto& src = ds.Fields[0].Values;
  bool ok = c5.CellSet.PointDims[0] == 13 && c5.CellSet.PointDims[1] == 9;
  for (int64_t k = 0; ok && k < c5.CellSet.PointDims[2]; ++k)
    for (int64_t j = 0; ok && j < c5.CellSet.PointDims[1]; ++j)
      for (int64_t i = 0; ok && i < c5.CellSet.PointDims[0]; ++i)
        ok = c5.Fields[0].Values[(size_t)((k * 9 + j) * 13 + i)] == src[(size_t)((k * 24 + 15 + j) * 24 + 11 + i)];
  CHECK(ok);
  const std::string prefix = TempPath("chunk");
  CHECK(io::SaveChunksToDisk(chunks, prefix).Success);
  const io::ReadResult rc = io::DataSetUtils::Read(prefix + "_5.vtk");
  CHECK(rc.Success && SameDataSet(c5, rc.Value));
  for (size_t i = 0; i < chunks.size(); ++i) std::remove((prefix + "_" + std::to_string(i) + ".vtk").c_str());
  // a file scene: preset dataset = "file" with the per-rank pattern
  const std::string fs = TempPath("scene_r0_of_1.vtk");
  CHECK(io::DataSetUtils::Write(ds, fs).Success);
  beams::Config cfg;
  const std::string toml = "[presets.f]\ndataset = \"file\"\nfile_pattern = \"" +
                           TempPath("scene_r%r_of_%s.vtk") + "\"\n";
  CHECK(cfg.Deserialize(toml).Success);
  beams::DataSet fromFile;
  std::string fname;
  CHECK(beams::DataSetFactory::CreateDataSetFromPreset(cfg.Presets.at("f"), fromFile, fname).Success);
  CHECK(fname == ds.Fields[0].Name && SameDataSet(ds, fromFile));
  std::remove(fs.c_str());
  bool threw = false;
  try {
    io::ChunkPointRanges(ds.CellSet.PointDims, 0, 1, 1);
  } catch (const beams::ErrorBadValue&) {
    threw = true;
  }
  CHECK(threw);
}

// The synthetic frame written by test_images; tests/test_host_cpp.py decodes the files
// with zlib and recomputes the same values.
static float FrameValue(int x, int y, int c)
{
  if (x == 0 && y == 0) {
    const float special[4] = { -0.5f, 1.7f, NAN, 0.5f };
    return special[c];
  }
  return static_cast<float>((x * 37 + y * 101 + c * 53) % 256) / 255.0f;
}

// ResampleChunks: convert's chunk grids (dims / chunks cells per chunk, not the source's
// spacing) probed from a linear field, which trilinear interpolation reproduces exactly.
static void test_resample_chunks()
{
  namespace io = beams::io;
  const std::array<int64_t, 3> dims = { 20, 14, 11 };
  beams::DataSet ds;
  ds.CellSet.PointDims = dims;
  ds.Coords = beams::CoordinateSystem::Uniform(dims, { 0.0, 0.0, 0.0 },
                                               { 1.0 / 19, 1.0 / 13, 1.0 / 10 });
  auto lin = [](double x, double y, double z) { return 2.0 * x + 3.0 * y - z + 0.5; };
  beams::Field pf, cf;
  pf.Name = "lin";
  cf.Name = "cellid";
  cf.Assoc = beams::Association::Cells;
  for (int64_t k = 0; k < dims[2]; ++k)
    for (int64_t j = 0; j < dims[1]; ++j)
      for (int64_t i = 0; i < dims[0]; ++i) {
        pf.Values.push_back(static_cast<float>(lin(ds.Coords.Axes[0][i], ds.Coords.Axes[1][j], ds.Coords.Axes[2][k])));
        if (i < dims[0] - 1 && j < dims[1] - 1 && k < dims[2] - 1)
          cf.Values.push_back(static_cast<float>((k * (dims[1] - 1) + j) * (dims[0] - 1) + i));
      }
  ds.Fields = { pf, cf };
  const auto chunks = io::ResampleChunks(ds, 2, 3, 2);
  CHECK(chunks.size() == 12);
  bool shape = true, linear = true, cells = true, finite = true;
  double maxerr = 0.0;
  for (size_t c = 0; c < chunks.size(); ++c) {
    const auto& ch = chunks[c];
    const int cx = static_cast<int>(c % 2), cy = static_cast<int>((c / 2) % 3), cz = static_cast<int>(c / 6);
    shape &= ch.CellSet.PointDims == std::array<int64_t, 3>{ 11, 5, 6 };
    // origin = chunk index x bounds / chunks; spacing = (1 / chunks) / (dims / chunks)
    shape &= std::fabs(ch.Coords.Axes[0][0] - cx * 0.5f) < 1e-6f && std::fabs(ch.Coords.Axes[1][0] - cy / 3.0f) < 1e-6f &&
             std::fabs(ch.Coords.Axes[2][0] - cz * 0.5f) < 1e-6f;
    shape &= std::fabs(ch.Coords.Axes[0][1] - ch.Coords.Axes[0][0] - 0.05f) < 1e-6f &&
             std::fabs(ch.Coords.Axes[1][1] - ch.Coords.Axes[1][0] - 1.0f / 12) < 1e-6f;
    const auto& v = ch.GetField("lin").Values;
    const auto& id = ch.GetField("cellid").Values;
    size_t o = 0;
    for (int64_t k = 0; k < 6; ++k)
      for (int64_t j = 0; j < 5; ++j)
        for (int64_t i = 0; i < 11; ++i, ++o) {
          const double x = ch.Coords.Axes[0][i], y = ch.Coords.Axes[1][j], z = ch.Coords.Axes[2][k];
          finite &= std::isfinite(v[o]);
          maxerr = std::fmax(maxerr, std::fabs(v[o] - lin(std::fmin(x, 1.0), std::fmin(y, 1.0), std::fmin(z, 1.0))));
          // the cell field takes the value of the source cell holding the point
          const int64_t ci = std::min<int64_t>(static_cast<int64_t>(std::floor(x * 19 + 1e-9)), 18);
          const int64_t cj = std::min<int64_t>(static_cast<int64_t>(std::floor(y * 13 + 1e-9)), 12);
          const int64_t ck = std::min<int64_t>(static_cast<int64_t>(std::floor(z * 10 + 1e-9)), 9);
          const float want = static_cast<float>((ck * 13 + cj) * 19 + ci);
          cells &= id[o] == want || std::fabs(x * 19 - std::round(x * 19)) < 1e-4 ||
                   std::fabs(y * 13 - std::round(y * 13)) < 1e-4 || std::fabs(z * 10 - std::round(z * 10)) < 1e-4;
        }
  }
  linear = maxerr < 2e-5;
  CHECK(shape);
  CHECK(finite);
  CHECK(linear);
  CHECK(cells);
  bool threw = false;
  try {
    io::ResampleChunks(ds, 21, 1, 1);
  } catch (const beams::ErrorBadValue&) {
    threw = true;
  }
  CHECK(threw);

  // The chunk size is the FP64 bounds length over the chunk count, rounded once to FP32
  // (Chunk.cxx: bounds.X.Length() / chunksX).  On [0.1431428, 2.2749565] in 5 chunks an FP32
  // subtraction and an FP32 division round twice, to 0.42636275 instead of 0.42636272.
  {
    beams::DataSet r;
    r.CellSet.PointDims = { 11, 2, 2 };
    const float a = 0.1431428f, b = 2.2749565f;
    for (int i = 0; i < 11; ++i) r.Coords.Axes[0].push_back(a + (b - a) * static_cast<float>(i) / 10.0f);
    r.Coords.Axes[0].back() = b;
    r.Coords.Axes[1] = { 0.0f, 1.0f };
    r.Coords.Axes[2] = { 0.0f, 1.0f };
    beams::Field f;
    f.Name = "zero";
    f.Values.assign(11 * 2 * 2, 0.0f);
    r.Fields = { f };
    const auto rc = io::ResampleChunks(r, 5, 1, 1);
    const float once = static_cast<float>((static_cast<double>(b) - static_cast<double>(a)) / 5.0);
    const float twice = (b - a) / 5.0f;
    CHECK(once != twice);
    bool origins = rc.size() == 5;
    for (size_t c = 0; origins && c < rc.size(); ++c)
      origins = rc[c].Coords.Axes[0][0] == a + static_cast<float>(c) * once &&
                rc[c].Coords.Axes[0][1] == a + static_cast<float>(c) * once + once / 2.0f;
    CHECK(origins);
  }
}

static void test_images()
{
  namespace io = beams::io;
  CHECK(io::QuantizeChannel(0.0f) == 0 && io::QuantizeChannel(1.0f) == 255);
  CHECK(io::QuantizeChannel(0.5f) == 128 && io::QuantizeChannel(-3.0f) == 0);
  CHECK(io::QuantizeChannel(2.0f) == 255 && io::QuantizeChannel(NAN) == 0);
  CHECK(io::FrameFileName("blobs", 3, 7) == "blobs_3_007.png");
  CHECK(io::FrameFileName("s", 0, 1234, ".ppm") == "s_0_1234.ppm");

  const int w = 5, h = 3;
  std::vector<float> rgba(static_cast<size_t>(w) * h * 4);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      for (int c = 0; c < 4; ++c) rgba[(static_cast<size_t>(y) * w + x) * 4 + c] = FrameValue(x, y, c);
  const char* dir = std::getenv("BEAMS_TEST_IMAGE_DIR");
  const std::string base = dir ? std::string(dir) + "/frame" : TempPath("frame");
  CHECK(io::SaveImage(rgba, w, h, base + ".png").Success);
  CHECK(io::SaveImage(rgba, w, h, base + ".ppm").Success);
  CHECK(!io::SaveImage(rgba, w, h, base + ".jpg").Success);
  CHECK(!io::SaveImage(rgba, w, h + 1, base + ".png").Success);
  beams::rendering::Canvas canvas(w, h);
  canvas.Color = rgba;
  bool threw = false;
  try {
    canvas.SaveAs(base + ".bmp");
  } catch (const beams::ErrorBadValue& e) {
    threw = std::strstr(e.what(), "unknown image type") != nullptr;
  }
  CHECK(threw);

  std::ifstream pf(base + ".ppm", std::ios::binary);
  std::string ppm((std::istreambuf_iterator<char>(pf)), std::istreambuf_iterator<char>());
  const std::string header = "P6\n5 3\n255\n";
  CHECK(ppm.size() == header.size() + 45 && ppm.compare(0, header.size(), header) == 0);
  if (ppm.size() == header.size() + 45) {
    bool same = true;
    for (int i = 0; i < 15; ++i)
      for (int c = 0; c < 3; ++c)
        same &= static_cast<uint8_t>(ppm[header.size() + 3 * i + c]) ==
                io::QuantizeChannel(FrameValue(i % w, i / w, c));
    CHECK(same);
  }
  // PNG: signature, IHDR, and the IEND chunk whose CRC-32 is the well-known AE 42 60 82
  std::ifstream nf(base + ".png", std::ios::binary);
  std::string png((std::istreambuf_iterator<char>(nf)), std::istreambuf_iterator<char>());
  const unsigned char sig[8] = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n' };
  CHECK(png.size() > 45 && std::memcmp(png.data(), sig, 8) == 0);
  CHECK(png.compare(12, 4, "IHDR") == 0 && png[19] == 5 && png[23] == 3 && png[24] == 8 && png[25] == 6);
  const unsigned char iend[12] = { 0, 0, 0, 0, 'I', 'E', 'N', 'D', 0xAE, 0x42, 0x60, 0x82 };
  CHECK(png.size() > 12 && std::memcmp(png.data() + png.size() - 12, iend, 12) == 0);
  // a frame of more than one 65535-byte stored deflate block
  const int bw = 160, bh = 120;
  std::vector<float> big(static_cast<size_t>(bw) * bh * 4);
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x)
      for (int c = 0; c < 4; ++c) big[(static_cast<size_t>(y) * bw + x) * 4 + c] = FrameValue(x, y, c);
  CHECK(io::SaveImage(big, bw, bh, base + "_big.png").Success);
  if (!dir) {
    std::remove((base + "_big.png").c_str());
    std::remove((base + ".png").c_str());
    std::remove((base + ".ppm").c_str());
  }
}

static void test_color_tables()
{
  namespace r = beams::rendering;
  // Moreland's Msh coordinates of the cool and warm endpoints: (80, 1.08, -1.10), (80, 1.08, 0.50)
  const auto cool = r::RgbToMsh({ 0.230, 0.299, 0.754 });
  const auto warm = r::RgbToMsh({ 0.706, 0.016, 0.150 });
  CHECK(std::fabs(cool[0] - 80.0) < 0.1 && std::fabs(cool[1] - 1.08) < 0.005 && std::fabs(cool[2] + 1.10) < 0.005);
  CHECK(std::fabs(warm[0] - 80.0) < 0.1 && std::fabs(warm[1] - 1.08) < 0.005 && std::fabs(warm[2] - 0.50) < 0.005);
  // Msh -> RGB inverts RGB -> Msh
  const auto back = r::MshToRgb(cool);
  CHECK(std::fabs(back[0] - 0.230) < 1e-5 && std::fabs(back[1] - 0.299) < 1e-5 && std::fabs(back[2] - 0.754) < 1e-5);

  // Moreland's published cool-to-warm byte table at eighths of the range (one count of
  // rounding slack): the hue spin, the white midpoint and the Lab transforms all show here
  const int table[9][3] = { { 59, 76, 192 },   { 98, 130, 234 },  { 141, 176, 254 },
                            { 184, 208, 249 }, { 221, 221, 221 }, { 245, 196, 173 },
                            { 244, 154, 123 }, { 222, 96, 77 },   { 180, 4, 38 } };
  r::ColorTable ct("Cool to Warm");
  CHECK(ct.Valid());
  bool near = true;
  for (int i = 0; i <= 8; ++i) {
    const auto c = ct.MapColor(i / 8.0);
    for (int k = 0; k < 3; ++k) near &= std::fabs(c[k] * 255.0 - table[i][k]) <= 1.0;
  }
  CHECK(near);
  CHECK(!r::ColorTable("no such preset").Valid());

  // alpha points: piecewise linear, constant outside, later point at the same x replaces
  ct.AddPointAlpha(0.0, 0.0f);
  ct.AddPointAlpha(1.0, 0.8f);
  ct.AddPointAlpha(0.5, 0.1f);
  ct.AddPointAlpha(0.5, 0.2f);
  CHECK(std::fabs(ct.MapAlpha(0.25) - 0.1) < 1e-7 && std::fabs(ct.MapAlpha(0.75) - 0.5) < 1e-7);
  CHECK(ct.MapAlpha(-1.0) == 0.0 && std::fabs(ct.MapAlpha(2.0) - 0.8) < 1e-7);
  const auto s = ct.Sample(1024);
  CHECK(s.size() == 1024);
  CHECK(s[0][0] == 59 && s[0][1] == 76 && s[0][2] == 192 && s[0][3] == 0);
  CHECK(s[1023][0] == 180 && s[1023][1] == 4 && s[1023][2] == 38 && s[1023][3] == 204);
  const auto m = ct.SampleColorMap();
  CHECK(m.size() == 1024 && m[1023].W == 204.0f / 255.0f);

  // the grayscale preset is a linear ramp: entry i of n holds round(255 i / (n - 1))
  const auto g = r::ColorTable("grayscale").Sample(256);
  bool ramp = g.size() == 256;
  for (int i = 0; ramp && i < 256; ++i) ramp = g[i][0] == i && g[i][1] == i && g[i][2] == i && g[i][3] == 255;
  CHECK(ramp);

  // a scene built from a named table with alpha points gets the 1024-entry sampled map
  beams::Config cfg;
  CHECK(cfg.Deserialize("[presets.t]\ndataset = \"blobs\"\nsize = 8\ncolor_table = \"cool to warm\"\n"
                        "point_alphas = [0.0, 0.0, 1.0, 0.5]\n").Success);
  CHECK(cfg.Presets.at("t").ColorTable.PointAlphas.size() == 2);
  auto scene = beams::SceneFactory::CreateFromPreset(cfg.Presets.at("t"));
  CHECK(scene->ColorMap.size() == 1024 && scene->ColorMap[1023].W == 128.0f / 255.0f);
  CHECK(scene->ColorMap[0].X == 59.0f / 255.0f && scene->ColorMap[0].W == 0.0f);
  CHECK(!cfg.Deserialize("[presets.t]\npoint_alphas = [0.0, 0.0, 1.0]\n").Success);
  CHECK(!cfg.Deserialize("[presets.t]\npoint_alphas = [0.0, nan, 1.0, 0.5]\n").Success);
  beams::Config named;
  CHECK(named.Deserialize("[presets.u]\ndataset = \"blobs\"\nsize = 8\ncolor_table = \"nope\"\n").Success);
  bool threw = false;
  try {
    beams::SceneFactory::CreateFromPreset(named.Presets.at("u"));
  } catch (const beams::ErrorBadValue& e) {
    threw = std::strstr(e.what(), "nope") != nullptr;
  }
  CHECK(threw);
}

static void test_cpu()
{
  test_resample_chunks();
  test_images();
  test_color_tables();

  beams::Config cfg;
  beams::Result r = cfg.Deserialize(kToml);
  CHECK(r.Success);
  CHECK(cfg.DefaultPresetName == "small");
  CHECK(cfg.Presets.size() == 2);
  const beams::Preset& p = cfg.Presets.at("small");
  CHECK(p.OpacityMap.Size[1] == 23);
  CHECK(std::fabs(p.Camera.Elevation - 12.5f) < 1e-6f);
  CHECK(p.Lights.size() == 1 && std::fabs(p.Lights[0].Position[2] + 0.8f) < 1e-6f);
  CHECK(p.ImageSize[0] == 96 && p.ImageSize[1] == 64);
  CHECK(!cfg.Presets.at("half").OpacityMap.Enabled);

  beams::Config bad;
  CHECK(!bad.Deserialize("[presets.x]\nunknown_key = 3\n").Success);
  CHECK(!bad.Deserialize("[presets.x]\nmap_size = [1, 2]\n").Success);
  CHECK(!bad.Deserialize("size = 3\n").Success);
  CHECK(!bad.Deserialize("[other]\n").Success);

  beams::DataSet ds;
  std::string fname;
  CHECK(beams::DataSetFactory::CreateDataSetFromPreset(p, ds, fname).Success);
  CHECK(ds.CellSet.PointDims[0] == 24 && ds.Coords.Axes[0].size() == 24);
  const auto& f = ds.GetField(fname);
  float mn = 1e9f, mx = -1e9f;
  for (float v : f.Values) {
    mn = std::fmin(mn, v);
    mx = std::fmax(mx, v);
  }
  CHECK(mn == 0.0f && mx == 1.0f);
  beams::DataSet ds2;
  CHECK(beams::DataSetFactory::CreateDataSetFromPreset(cfg.Presets.at("half"), ds2, fname).Success);
  CHECK(ds2.Fields[0].Type == beams::DType::BF16 && ds2.Fields[0].Half.size() == 16 * 16 * 16);
  test_io(ds, ds2);

  beams::rendering::LightedVolumeRenderer lvr;
  bool threw = false;
  try {
    lvr.SetSampleDistance(-1.0f);
  } catch (const beams::ErrorBadValue&) {
    threw = true;
  }
  CHECK(threw);
  beams::rendering::Ray<float> rays;
  rays.Resize(4);
  threw = false;
  try {
    lvr.Render(rays);  // no SetData yet
  } catch (const beams::ErrorBadValue& e) {
    threw = std::strstr(e.what(), "SetData") != nullptr;
  }
  CHECK(threw);
}

static void test_gpu()
{
  beams::Config cfg;
  CHECK(cfg.Deserialize(kToml).Success);
  const beams::Preset& p = cfg.Presets.at(cfg.DefaultPresetName);
  auto scene = beams::SceneFactory::CreateFromPreset(p);
  beams::rendering::Canvas canvas(p.ImageSize[0], p.ImageSize[1]);
  scene->CanvasPtr = &canvas;
  CHECK(scene->Ready().Success);
  scene->RenderFrame();
  double amax = 0.0;
  for (size_t i = 3; i < canvas.Color.size(); i += 4) amax = std::fmax(amax, canvas.Color[i]);
  CHECK(amax > 0.5);  // the background is blended under the image (alpha -> 1 everywhere)
  const std::string frame = TempPath(beams::io::FrameFileName(scene->Id, 0, 0));
  canvas.SaveAs(frame);
  std::ifstream ff(frame, std::ios::binary);
  std::string pngHead(8, '\0');
  CHECK(ff.read(&pngHead[0], 8) && pngHead.compare(1, 3, "PNG") == 0);
  std::remove(frame.c_str());

  // Render(Ray<Float32>) on the camera's rays equals RenderImage without background.
  beams::rendering::LightedVolumeRenderer& lvr = scene->Mapper.GetRenderer();
  const int w = canvas.Width, h = canvas.Height;
  beams::rendering::Ray<float> rays;
  rays.Resize(static_cast<size_t>(w) * h);
  beams_camera cam{};
  for (int i = 0; i < 3; ++i) {
    cam.eye[i] = scene->Cam.Position[i];
    cam.look_at[i] = scene->Cam.LookAt[i];
    cam.up[i] = scene->Cam.ViewUp[i];
  }
  cam.fov_y_deg = scene->Cam.FieldOfView;
  float* o[3] = { rays.Origin[0].data(), rays.Origin[1].data(), rays.Origin[2].data() };
  float* d[3] = { rays.Dir[0].data(), rays.Dir[1].data(), rays.Dir[2].data() };
  CHECK(beams_gen_camera_rays(&cam, w, h, o, d, nullptr, nullptr) == BEAMS_OK);
  lvr.Render(rays);
  std::vector<float> img(static_cast<size_t>(w) * h * 4), depth(static_cast<size_t>(w) * h);
  lvr.RenderImage(cam.eye, cam.look_at, cam.up, cam.fov_y_deg, w, h, nullptr, img.data(),
                  depth.data(), false);
  CHECK(std::memcmp(img.data(), rays.Buffer.data(), img.size() * sizeof(float)) == 0);
  auto t = lvr.GetTimings();
  CHECK(t[BEAMS_PHASE_MAP] > 0.0f && t[BEAMS_PHASE_MARCH] > 0.0f);
  auto m = lvr.GetTransmittanceMap();
  CHECK(m.size() == 24u * 24u * 24u);
}

int main(int argc, char** argv)
{
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  try {
    if (mode == "cpu" || mode == "all") test_cpu();
    if (mode == "gpu" || mode == "all") test_gpu();
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    ++failures;
  }
  std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
  return failures ? 1 : 0;
}
