// Light types of the renderer surface (Light.h, PointLight.h of the reference) plus the
// directional light the benchmark's plain configuration uses (an extension).
#ifndef beams_rendering_lights_h
#define beams_rendering_lights_h

#include <array>

namespace beams
{
namespace rendering
{
struct Light
{
  virtual ~Light() = default;
};

// PointLight.h:15-30: position, colour, intensity; the renderer uses Color * Intensity.
template <typename Precision>
struct PointLight : public Light
{
  using Vec3 = std::array<Precision, 3>;
  PointLight(Vec3 position, Vec3 color, Precision intensity)
    : Position(position)
    , Color(color)
    , Intensity(intensity)
  {
  }
  Vec3 Position;
  Vec3 Color;
  Precision Intensity;
};

// Parallel light travelling along Direction.
template <typename Precision>
struct DirectionalLight : public Light
{
  using Vec3 = std::array<Precision, 3>;
  DirectionalLight(Vec3 direction, Vec3 color, Precision intensity)
    : Direction(direction)
    , Color(color)
    , Intensity(intensity)
  {
  }
  Vec3 Direction;
  Vec3 Color;
  Precision Intensity;
};
} // namespace rendering
} // namespace beams

#endif
