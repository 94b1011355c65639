// Checks of the C++ host surface (beams::rendering::* over the C ABI).
//   test_host cpu   config parsing, DataSetFactory, error paths (no GPU needed)
//   test_host gpu   SceneFactory -> Scene::Ready -> RenderFrame, and Render(Ray) against
//                   RenderImage on the same camera rays (bit equality)
#include "../../include/beams_mi355x.h"
#include "beams/SceneFactory.h"

#include <cmath>
#include <cstdio>
#include <cstring>
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

static void test_cpu()
{
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
