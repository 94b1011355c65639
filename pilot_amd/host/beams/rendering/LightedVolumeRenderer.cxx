#include "LightedVolumeRenderer.h"

#include "../../../../include/beams_mi355x.h"

#include <cmath>
#include <limits>
#include <string>

namespace beams
{
namespace rendering
{

// LightedVolumeRenderer.cxx:752-759 defaults
LightedVolumeRenderer::LightedVolumeRenderer()
  : SampleDistance(-1.0f)
  , UseShadowMap(true)
  , ShadowMapSize{ 16, 16, 16 }
{
}

LightedVolumeRenderer::~LightedVolumeRenderer()
{
  if (this->Handle)
    beams_destroy(this->Handle);
}

void LightedVolumeRenderer::Check(int status, const char* what)
{
  if (status == BEAMS_OK)
    return;
  std::string msg = std::string(what) + ": ";
  msg += this->Handle ? beams_last_error(this->Handle) : "no gfx950 device";
  throw ErrorBadValue(msg);
}

void LightedVolumeRenderer::SetColorMap(const std::vector<Vec4f>& colorMap)
{
  this->ColorMap = colorMap;
  this->ColorDirty = true;
}

// LightedVolumeRenderer.cxx:766-779: keeps a borrowed pointer to the field, as the reference.
void LightedVolumeRenderer::SetData(const CoordinateSystem& coords,
                                    const Field& scalarField,
                                    const CellSetStructured& cellset,
                                    const Range& scalarRange)
{
  this->Coords = coords;
  this->ScalarField = &scalarField;
  this->CellSet = cellset;
  this->ScalarRange = scalarRange;
  this->DataDirty = true;
}

// LightedVolumeRenderer.cxx:1374-1379
void LightedVolumeRenderer::SetSampleDistance(const float& distance)
{
  if (distance <= 0.f)
    throw ErrorBadValue("Sample distance must be positive.");
  this->SampleDistance = distance;
}

void LightedVolumeRenderer::SetAlphaStop(float a)
{
  if (!(a > 0.0f) || a > 1.0f)
    throw ErrorBadValue("alpha stop must be in (0, 1]");
  this->AlphaStop = a;
}

void LightedVolumeRenderer::AddLight(std::shared_ptr<Light> light)
{
  if (!light)
    throw ErrorBadValue("null light");
  this->Lights.push_back(std::move(light));
}

void LightedVolumeRenderer::ClearLights()
{
  this->Lights.clear();
}

void LightedVolumeRenderer::Push()
{
  if (!this->ScalarField)
    throw ErrorBadValue("SetData must be called before rendering");
  // LightedVolumeRenderer.cxx:809-813
  if (!this->ScalarField->IsCellField() && !this->ScalarField->IsPointField())
    throw ErrorBadValue("Field not accociated with cell set or points");
  if (!this->Handle)
    this->Check(beams_create(&this->Handle), "beams_create");
  if (this->DataDirty)
  {
    const float* axes[3] = { this->Coords.Axes[0].data(), this->Coords.Axes[1].data(),
                             this->Coords.Axes[2].data() };
    const int64_t dims[3] = { this->CellSet.PointDims[0], this->CellSet.PointDims[1],
                              this->CellSet.PointDims[2] };
    this->Check(beams_set_field(this->Handle, this->ScalarField->Data(),
                                static_cast<int32_t>(this->ScalarField->Type),
                                static_cast<int32_t>(this->ScalarField->Assoc), dims, axes, nullptr,
                                nullptr,
                                this->ScalarField->Device ? BEAMS_MEM_DEVICE : BEAMS_MEM_HOST),
                "SetData");
    this->Check(beams_set_scalar_range(this->Handle, this->ScalarRange.Min, this->ScalarRange.Max),
                "SetData(range)");
    this->DataDirty = false;
  }
  if (this->ColorDirty)
  {
    this->Check(beams_set_colormap(this->Handle, &this->ColorMap[0].X,
                                   static_cast<int64_t>(this->ColorMap.size())),
                "SetColorMap");
    this->ColorDirty = false;
  }
  // Only light 0 is used (TransmittanceMap.h:591-592).
  if (!this->Lights.empty())
  {
    const Light* l = this->Lights.front().get();
    if (auto* p = dynamic_cast<const PointLight<float>*>(l))
    {
      this->Check(beams_set_light(this->Handle, BEAMS_LIGHT_POINT, p->Position.data(),
                                  p->Color.data(), p->Intensity),
                  "AddLight");
    }
    else if (auto* d = dynamic_cast<const DirectionalLight<float>*>(l))
    {
      this->Check(beams_set_light(this->Handle, BEAMS_LIGHT_DIRECTIONAL, d->Direction.data(),
                                  d->Color.data(), d->Intensity),
                  "AddLight");
    }
    else
    {
      throw ErrorBadValue("unsupported light type");
    }
  }
  const int64_t cells[3] = { this->ShadowMapSize[0], this->ShadowMapSize[1],
                             this->ShadowMapSize[2] };
  this->Check(beams_set_shadow(this->Handle, this->UseShadowMap ? 1 : 0, cells,
                               BEAMS_SHADOW_REF_MARCH, static_cast<int32_t>(this->MapStorage)),
              "SetShadowMapSize");
  if (this->SampleDistance > 0.0f)
    this->Check(beams_set_sample_distance(this->Handle, this->SampleDistance), "SetSampleDistance");
  this->Check(beams_set_alpha_stop(this->Handle, this->AlphaStop), "SetAlphaStop");
}

void LightedVolumeRenderer::Render(Ray<float>& rays)
{
  this->Push();
  const float* o[3] = { rays.Origin[0].data(), rays.Origin[1].data(), rays.Origin[2].data() };
  const float* d[3] = { rays.Dir[0].data(), rays.Dir[1].data(), rays.Dir[2].data() };
  this->Check(beams_render_rays(this->Handle, static_cast<int64_t>(rays.NumRays()), o, d,
                                rays.MinDistance.data(), rays.MaxDistance.data(),
                                rays.Distance.data(), rays.Buffer.data(), BEAMS_MEM_HOST),
              "Render");
}

void LightedVolumeRenderer::RenderImage(const double eye[3], const double lookAt[3],
                                        const double up[3], double fovY, int width, int height,
                                        const float* background, float* rgba, float* depth,
                                        bool deviceOutput)
{
  this->Push();
  beams_camera cam{};
  for (int i = 0; i < 3; ++i)
  {
    cam.eye[i] = eye[i];
    cam.look_at[i] = lookAt[i];
    cam.up[i] = up[i];
  }
  cam.fov_y_deg = fovY;
  this->Check(beams_render_image(this->Handle, &cam, width, height, background, rgba, depth,
                                 deviceOutput ? BEAMS_MEM_DEVICE : BEAMS_MEM_HOST),
              "RenderImage");
}

std::array<float, 5> LightedVolumeRenderer::GetTimings()
{
  std::array<float, 5> t{};
  if (this->Handle)
    this->Check(beams_get_timings(this->Handle, t.data(), 5), "GetTimings");
  return t;
}

std::vector<float> LightedVolumeRenderer::GetTransmittanceMap()
{
  std::vector<float> m(static_cast<size_t>((this->ShadowMapSize[0] + 1) *
                                           (this->ShadowMapSize[1] + 1) *
                                           (this->ShadowMapSize[2] + 1)));
  if (!this->Handle)
    throw ErrorBadValue("nothing rendered yet");
  this->Check(beams_get_map(this->Handle, m.data(), BEAMS_MEM_HOST), "GetTransmittanceMap");
  return m;
}

} // namespace rendering
} // namespace beams
