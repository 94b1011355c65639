#include "SceneFactory.h"

#include "io/DataSetUtils.h"

#include "../../../include/beams_mi355x.h"

#include <cstdlib>

namespace beams
{
namespace
{
std::string Param(const Preset& p, const std::string& key, const std::string& dflt)
{
  auto it = p.DataSet.Params.find(key);
  return it == p.DataSet.Params.end() ? dflt : it->second;
}
} // namespace

// FileSceneBase::GetDataSet (FileSceneBase.cxx:33-49): the file of this rank, "%s" standing
// for the number of ranks and "%r" for the rank; one device renders all bricks, so rank 0 of 1.
Result ReadFileDataSet(const Preset& preset, DataSet& dataSet, std::string& fieldName)
{
  std::string name = Param(preset, "file_pattern", "");
  if (name.empty())
    return Result::Failed("file dataset needs file_pattern");
  for (const auto& rep : { std::make_pair(std::string("%s"), std::string("1")),
                           std::make_pair(std::string("%r"), std::string("0")) })
    for (size_t p = name.find(rep.first); p != std::string::npos; p = name.find(rep.first, p))
      name.replace(p, rep.first.size(), rep.second);
  io::ReadResult r = io::DataSetUtils::Read(name);
  if (!r.Success)
    return Result::Failed(r.Error);
  const std::string want = Param(preset, "field", "");
  fieldName.clear();
  for (const auto& f : r.Value.Fields)
    if ((want.empty() && f.IsPointField()) || f.Name == want)
    {
      fieldName = f.Name;
      break;
    }
  if (fieldName.empty())
    return Result::Failed(name + ": no " + (want.empty() ? std::string("point field") : "field " + want));
  dataSet = std::move(r.Value);
  return Result::Succeeded();
}

Result DataSetFactory::CreateDataSetFromPreset(const Preset& preset, DataSet& dataSet,
                                               std::string& fieldName)
{
  if (preset.DataSet.Factory == "file")
    return ReadFileDataSet(preset, dataSet, fieldName);
  if (preset.DataSet.Factory != "blobs")
    return Result::Failed("unknown dataset factory: " + preset.DataSet.Factory);
  const int64_t n = std::atoll(Param(preset, "size", "64").c_str());
  if (n < 2)
    return Result::Failed("dataset size must be >= 2");
  beams_field_recipe r{};
  r.seed = std::strtoull(Param(preset, "seed", "1").c_str(), nullptr, 10);
  r.nblobs = std::atoi(Param(preset, "blobs", "16").c_str());
  r.noise_octaves = std::atoi(Param(preset, "noise_octaves", "0").c_str());
  r.noise_amp = std::atof(Param(preset, "noise_amp", "0").c_str());
  r.noise_freq = 4.0;
  r.time = std::atof(Param(preset, "time", "0").c_str());
  r.drift_speed = std::atof(Param(preset, "drift_speed", "0").c_str());
  r.noise_phase_rate = 0.0;
  const std::string dtype = Param(preset, "dtype", "f32");
  Field f;
  f.Name = "scalars";
  f.Assoc = Association::Points;
  const int64_t dims[3] = { n, n, n };
  const size_t count = static_cast<size_t>(n * n * n);
  int rc;
  if (dtype == "f32")
  {
    f.Type = DType::F32;
    f.Values.resize(count);
    rc = beams_gen_field(&r, dims, BEAMS_F32, f.Values.data(), BEAMS_MEM_HOST);
  }
  else if (dtype == "f16" || dtype == "bf16")
  {
    f.Type = dtype == "f16" ? DType::F16 : DType::BF16;
    f.Half.resize(count);
    rc = beams_gen_field(&r, dims, dtype == "f16" ? BEAMS_F16 : BEAMS_BF16, f.Half.data(),
                         BEAMS_MEM_HOST);
  }
  else
    return Result::Failed("unknown dtype: " + dtype);
  if (rc != BEAMS_OK)
    return Result::Failed("field generation failed: status " + std::to_string(rc));
  dataSet = DataSet();
  dataSet.CellSet.PointDims = { n, n, n };
  const double sp = 1.0 / static_cast<double>(n - 1);
  dataSet.Coords = CoordinateSystem::Uniform({ n, n, n }, { 0.0, 0.0, 0.0 }, { sp, sp, sp });
  dataSet.Fields.push_back(std::move(f));
  fieldName = "scalars";
  return Result::Succeeded();
}

namespace rendering
{
Result Scene::Ready()
{
  if (!this->CanvasPtr)
    return Result::Failed("scene has no canvas");
  this->CanvasPtr->Clear();
  this->Cam = Camera::Orbit(this->Data.Coords.GetBounds(), this->Azimuth, this->Elevation,
                            this->FieldOfView);
  this->Mapper.SetCompositeBackground(true);
  this->Mapper.SetCanvas(this->CanvasPtr);
  this->Mapper.SetActiveColorMap(this->ColorMap);
  this->Mapper.SetUseShadowMap(this->UseShadowMap);
  this->Mapper.SetShadowMapSize(this->ShadowMapSize);
  if (this->SampleDistance > 0.0f)
    this->Mapper.SetSampleDistance(this->SampleDistance);
  this->Mapper.GetRenderer().SetAlphaStop(this->AlphaStop);
  this->Mapper.ClearLights();
  if (this->SceneLight)
    this->Mapper.AddLight(this->SceneLight);
  return Result::Succeeded();
}

void Scene::RenderFrame()
{
  this->Mapper.RenderCells(this->Data.CellSet, this->Data.Coords,
                           this->Data.GetField(this->FieldName), this->Cam, this->ScalarRange);
}
} // namespace rendering

std::shared_ptr<rendering::Scene> SceneFactory::CreateFromPreset(const Preset& preset)
{
  auto scene = std::make_shared<rendering::Scene>();
  Result r = DataSetFactory::CreateDataSetFromPreset(preset, scene->Data, scene->FieldName);
  if (!r.Success)
    throw ErrorBadValue(r.Err);
  scene->Id = preset.Id;
  scene->ScalarRange = { preset.ColorTable.RangeMin, preset.ColorTable.RangeMax };
  if (preset.ColorTable.Name == "seeded")
  {
    const int ne = preset.ColorTable.NumEntries;
    std::vector<float> rgba(static_cast<size_t>(ne) * 4);
    if (beams_gen_colormap(preset.ColorTable.Seed, preset.ColorTable.NumPoints, ne, rgba.data(),
                           nullptr) != BEAMS_OK)
      throw ErrorBadValue("colour map generation failed");
    scene->ColorMap.resize(static_cast<size_t>(ne));
    for (int i = 0; i < ne; ++i)
      scene->ColorMap[static_cast<size_t>(i)] = { rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2],
                                                   rgba[4 * i + 3] };
  }
  else
  {
    // a named table with alpha points, as the reference's scenes build theirs
    rendering::ColorTable table(preset.ColorTable.Name);
    if (!table.Valid())
      throw ErrorBadValue("unknown colour table: " + preset.ColorTable.Name);
    for (const auto& pa : preset.ColorTable.PointAlphas)
      table.AddPointAlpha(pa[0], pa[1]);
    scene->ColorMap = table.SampleColorMap(1024);
  }
  const auto& pd = scene->Data.CellSet.PointDims;
  scene->ShadowMapSize = preset.OpacityMap.CalculateSize({ pd[0] - 1, pd[1] - 1, pd[2] - 1 });
  scene->UseShadowMap = preset.OpacityMap.Enabled;
  scene->Azimuth = preset.Camera.Azimuth;
  scene->Elevation = preset.Camera.Elevation;
  scene->FieldOfView = preset.Camera.FieldOfView;
  scene->SampleDistance = preset.SampleDistance;
  scene->AlphaStop = preset.AlphaStop;
  if (!preset.Lights.empty())
  {
    const LightOption& l = preset.Lights.front();
    if (l.Type == "directional")
      scene->SceneLight =
        std::make_shared<rendering::DirectionalLight<float>>(l.Position, l.Color, l.Intensity);
    else
      scene->SceneLight =
        std::make_shared<rendering::PointLight<float>>(l.Position, l.Color, l.Intensity);
  }
  return scene;
}
} // namespace beams
