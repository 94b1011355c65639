// beams::rendering::LightedVolumeRenderer over the MI355X C ABI.
//
// Same method names and meaning as the reference class
// (beams/rendering/LightedVolumeRenderer.h:20-101): setters record state, Render()
// marches a Ray<Float32> batch (building the transmittance map first), and errors are
// thrown as ErrorBadValue where the reference throws vtkm::cont::ErrorBadValue.
#ifndef beams_rendering_lighted_volume_renderer_h
#define beams_rendering_lighted_volume_renderer_h

#include "../DataModel.h"
#include "Lights.h"

#include <limits>
#include <memory>
#include <vector>

struct beams_renderer;

namespace beams
{
namespace rendering
{
// The fields of vtkm::rendering::raytracing::Ray<Precision> the renderer reads and writes.
template <typename Precision>
struct Ray
{
  std::array<std::vector<Precision>, 3> Origin, Dir;
  std::vector<Precision> MinDistance, MaxDistance, Distance;
  std::vector<Precision> Buffer; // Buffers[0]: 4 interleaved channels per ray

  void Resize(size_t n)
  {
    for (int d = 0; d < 3; ++d)
    {
      Origin[d].assign(n, 0);
      Dir[d].assign(n, 0);
    }
    MinDistance.assign(n, 0);
    MaxDistance.assign(n, std::numeric_limits<Precision>::infinity());
    Distance.assign(n, 0);
    Buffer.assign(4 * n, 0);
  }
  size_t NumRays() const { return MinDistance.size(); }
};

class LightedVolumeRenderer
{
public:
  LightedVolumeRenderer();
  ~LightedVolumeRenderer();
  LightedVolumeRenderer(const LightedVolumeRenderer&) = delete;
  LightedVolumeRenderer& operator=(const LightedVolumeRenderer&) = delete;

  void SetColorMap(const std::vector<Vec4f>& colorMap);
  void SetData(const CoordinateSystem& coords,
               const Field& scalarField,
               const CellSetStructured& cellset,
               const Range& scalarRange);
  void Render(Ray<float>& rays);
  void SetSampleDistance(const float& distance);
  void AddLight(std::shared_ptr<Light> light);
  void ClearLights();
  void SetUseShadowMap(bool useShadowMap) { this->UseShadowMap = useShadowMap; }
  void SetShadowMapSize(std::array<int64_t, 3> size) { this->ShadowMapSize = size; }
  void SetShadowMapStorage(DType t) { this->MapStorage = t; }
  void SetAlphaStop(float a);

  // The frame of MapperLightedVolume::RenderCells: pinhole camera rays, map, march.
  void RenderImage(const double eye[3], const double lookAt[3], const double up[3], double fovY,
                   int width, int height, const float* background, float* rgba, float* depth,
                   bool deviceOutput);

  // Phase times of the last render (map, exchange, remap, march, composite) in ms.
  std::array<float, 5> GetTimings();
  std::vector<float> GetTransmittanceMap();

  float SampleDistance;
  Range ScalarRange;
  bool UseShadowMap;
  std::array<int64_t, 3> ShadowMapSize;

private:
  void Push(); // create the device handle on first use and upload changed state
  void Check(int status, const char* what);

  beams_renderer* Handle = nullptr;
  CoordinateSystem Coords;
  const Field* ScalarField = nullptr;
  CellSetStructured CellSet;
  std::vector<Vec4f> ColorMap;
  std::vector<std::shared_ptr<Light>> Lights;
  DType MapStorage = DType::F32;
  float AlphaStop = 1.0f;
  bool DataDirty = false, ColorDirty = false;
};
} // namespace rendering
} // namespace beams

#endif
