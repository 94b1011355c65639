// DataSetFactory (beams/DataSetFactory.h:15), Scene (beams/rendering/Scene.h:25-46) and
// SceneFactory (beams/SceneFactory.h:13) of the Beams surface.  The dataset provider is
// the seeded blob + value-noise field of the benchmark configurations.
#ifndef beams_scenefactory_h
#define beams_scenefactory_h

#include "Config.h"
#include "DataModel.h"
#include "rendering/MapperLightedVolume.h"

#include <memory>

namespace beams
{
class DataSetFactory
{
public:
  static Result CreateDataSetFromPreset(const Preset& preset, DataSet& dataSet,
                                        std::string& fieldName);
};

namespace rendering
{
struct Scene
{
  virtual ~Scene() = default;
  // Scene::Ready (SpheresScene.cxx:181-211): camera, canvas, colour map, shadow map, light.
  virtual Result Ready();

  std::string Id;
  std::string FieldName;
  Range ScalarRange;
  std::vector<Vec4f> ColorMap;
  Camera Cam;
  MapperLightedVolume Mapper;
  DataSet Data;
  Canvas* CanvasPtr = nullptr;
  std::shared_ptr<Light> SceneLight;
  std::array<int64_t, 3> ShadowMapSize{ 16, 16, 16 };
  bool UseShadowMap = true;
  float Azimuth = 0.0f, Elevation = 0.0f, FieldOfView = 60.0f;
  float SampleDistance = -1.0f;
  float AlphaStop = 1.0f;

  // One frame: Mapper.RenderCells over the scene's dataset (Beams.cxx:252-257).
  void RenderFrame();
};
} // namespace rendering

struct SceneFactory
{
  static std::shared_ptr<rendering::Scene> CreateFromPreset(const Preset& preset);
};
} // namespace beams

#endif
