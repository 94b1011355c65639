#include "MapperLightedVolume.h"

#include <cmath>

namespace beams
{
namespace rendering
{

Camera Camera::Orbit(const Bounds& b, double azimuthDeg, double elevationDeg, double fov)
{
  Camera c;
  const auto ctr = b.Center();
  const double ex = b.X[1] - b.X[0], ey = b.Y[1] - b.Y[0], ez = b.Z[1] - b.Z[0];
  const double dist = std::sqrt(ex * ex + ey * ey + ez * ez);
  const double az = azimuthDeg * M_PI / 180.0, el = elevationDeg * M_PI / 180.0;
  const double dir[3] = { std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az) };
  for (int d = 0; d < 3; ++d)
  {
    c.Position[d] = ctr[d] + dist * dir[d];
    c.LookAt[d] = ctr[d];
  }
  c.FieldOfView = fov;
  return c;
}

MapperLightedVolume::MapperLightedVolume()
  : Tracer(std::make_shared<LightedVolumeRenderer>())
{
}

void MapperLightedVolume::SetCanvas(Canvas* canvas)
{
  this->Canvas_ = canvas;
}

void MapperLightedVolume::SetShadowMapSize(std::array<int64_t, 3> size)
{
  this->Tracer->SetShadowMapSize(size);
}

void MapperLightedVolume::SetUseShadowMap(bool useShadowMap)
{
  this->Tracer->SetUseShadowMap(useShadowMap);
}

void MapperLightedVolume::SetSampleDistance(float sampleDistance)
{
  this->SampleDistance = sampleDistance;
}

void MapperLightedVolume::SetCompositeBackground(bool compositeBackground)
{
  this->CompositeBackground = compositeBackground;
}

void MapperLightedVolume::AddLight(std::shared_ptr<Light> light)
{
  this->Tracer->AddLight(std::move(light));
}

void MapperLightedVolume::ClearLights()
{
  this->Tracer->ClearLights();
}

// MapperLightedVolume.cxx:114-193
void MapperLightedVolume::RenderCells(const CellSetStructured& cellset,
                                      const CoordinateSystem& coords,
                                      const Field& scalarField,
                                      const Camera& camera,
                                      const Range& scalarRange)
{
  if (!this->Canvas_)
    throw ErrorBadValue("Ray Tracer: no canvas set");
  auto& tracer = *this->Tracer;
  if (this->SampleDistance > 0.0f)
    tracer.SetSampleDistance(this->SampleDistance);
  tracer.SetData(coords, scalarField, cellset, scalarRange);
  tracer.SetColorMap(this->ColorMap);
  const float bg[4] = { this->Canvas_->Background.X, this->Canvas_->Background.Y,
                        this->Canvas_->Background.Z, this->Canvas_->Background.W };
  tracer.RenderImage(camera.Position.data(), camera.LookAt.data(), camera.ViewUp.data(),
                     camera.FieldOfView, this->Canvas_->Width, this->Canvas_->Height,
                     this->CompositeBackground ? bg : nullptr, this->Canvas_->Color.data(),
                     this->Canvas_->Depth.data(), false);
}

} // namespace rendering
} // namespace beams
