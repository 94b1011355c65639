// beams::rendering::MapperLightedVolume over the MI355X renderer
// (reference: beams/rendering/MapperLightedVolume.{h,cxx}).  RenderCells renders one
// frame of the block into a CanvasRayTracer-like canvas: camera rays, map build,
// march, optional background blend.
#ifndef beams_rendering_mapper_lighted_volume_h
#define beams_rendering_mapper_lighted_volume_h

#include "../io/ImageUtils.h"
#include "ColorTable.h"
#include "LightedVolumeRenderer.h"

namespace beams
{
namespace rendering
{
// Pinhole camera (the vtkm::rendering::Camera the reference passes to RenderCells).
struct Camera
{
  std::array<double, 3> Position{ 0, 0, 1 };
  std::array<double, 3> LookAt{ 0, 0, 0 };
  std::array<double, 3> ViewUp{ 0, 1, 0 };
  double FieldOfView = 60.0;

  // ResetToBounds + Azimuth + Elevation (SpheresScene.cxx:190-193): look at the centre of
  // the bounds from one diagonal away, up = +y.
  static Camera Orbit(const Bounds& b, double azimuthDeg, double elevationDeg, double fov = 60.0);
};

// RGBA + depth image (CanvasRayTracer).
struct Canvas
{
  Canvas(int w, int h)
    : Width(w)
    , Height(h)
    , Color(static_cast<size_t>(w) * h * 4, 0.0f)
    , Depth(static_cast<size_t>(w) * h, 0.0f)
  {
  }
  int GetWidth() const { return Width; }
  int GetHeight() const { return Height; }
  void SetBackgroundColor(const Vec4f& c) { Background = c; }
  void Clear()
  {
    std::fill(Color.begin(), Color.end(), 0.0f);
    std::fill(Depth.begin(), Depth.end(), 0.0f);
  }
  // Canvas::SaveAs as Beams.cxx:292-301 calls it: ".png" (RGBA) or ".ppm"/".pnm" (RGB), 8
  // bits per channel; an unknown extension or a failed write throws ErrorBadValue.
  void SaveAs(const std::string& fileName) const
  {
    const io::WriteResult r = io::SaveImage(Color, Width, Height, fileName);
    if (!r.Success)
      throw ErrorBadValue(r.Error);
  }
  int Width, Height;
  std::vector<float> Color;
  std::vector<float> Depth;
  Vec4f Background{ 1.0f, 1.0f, 1.0f, 1.0f };
};

class MapperLightedVolume
{
public:
  MapperLightedVolume();
  void SetCanvas(Canvas* canvas);
  Canvas* GetCanvas() const { return this->Canvas_; }
  void SetActiveColorMap(const std::vector<Vec4f>& colorMap) { this->ColorMap = colorMap; }
  // Mapper::SetActiveColorTable: the table sampled into 1024 8-bit RGBA entries / 255
  void SetActiveColorTable(const ColorTable& table) { this->ColorMap = table.SampleColorMap(1024); }
  void RenderCells(const CellSetStructured& cellset,
                   const CoordinateSystem& coords,
                   const Field& scalarField,
                   const Camera& camera,
                   const Range& scalarRange);
  void SetShadowMapSize(std::array<int64_t, 3> size);
  void SetUseShadowMap(bool useShadowMap);
  void SetSampleDistance(float sampleDistance);
  void SetCompositeBackground(bool compositeBackground);
  void AddLight(std::shared_ptr<Light> light);
  void ClearLights();
  void SetDensityScale(float) {} // stored but unused by the reference (LightedVolumeRenderer.h:71)
  LightedVolumeRenderer& GetRenderer() { return *this->Tracer; }

private:
  Canvas* Canvas_ = nullptr;
  std::shared_ptr<LightedVolumeRenderer> Tracer; // shared by copies, as Internals
  std::vector<Vec4f> ColorMap;
  float SampleDistance = -1.0f;
  bool CompositeBackground = true;
};
} // namespace rendering
} // namespace beams

#endif
