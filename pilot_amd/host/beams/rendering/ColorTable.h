// Transfer functions as the reference's scenes build them: a vtkm::cont::ColorTable with a
// named colour preset and AddPointAlpha control points (SpheresScene.cxx:99-110,
// FileSceneBase.cxx:67-88, VortexPatchScene.cxx:95-110), handed to the mapper with
// SetActiveColorTable (SpheresScene.cxx:147), which samples it into the renderer's colour
// map of 1024 8-bit RGBA entries read back as value / 255.
//
// VTK-m is not available to this build, so the colour arithmetic follows the published
// method the "cool to warm" preset names: Moreland's diverging colour map (K. Moreland,
// "Diverging Color Maps for Scientific Visualization", ISVC 2009), interpolating in the Msh
// space derived from CIELAB (D65 white) between the endpoints (0.230, 0.299, 0.754) and
// (0.706, 0.016, 0.150), through white (M = 88) in the middle.  Alpha is piecewise linear
// between its points and constant beyond the first and last.  Channels are quantised to
// round(clamp(v, 0, 1) * 255).  VTK-m's exact preset table and rounding are not pinned.
#ifndef beams_rendering_colortable_h
#define beams_rendering_colortable_h

#include "../DataModel.h"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace beams
{
namespace rendering
{
class ColorTable
{
public:
  // "cool to warm" (also the default: an empty name) or "grayscale" (linear RGB, black to
  // white); any other name is rejected by Valid().
  explicit ColorTable(const std::string& presetName = "");

  bool Valid() const { return this->IsValid; }
  void AddPointAlpha(double x, float alpha);
  // The colour at x in [range] and the alpha there.
  std::array<double, 3> MapColor(double x) const;
  double MapAlpha(double x) const;
  // The range the colour and alpha points span.
  std::array<double, 2> GetRange() const;
  // n evenly spaced entries over the range (both ends included), 8 bits per channel.
  std::vector<std::array<uint8_t, 4>> Sample(int n) const;
  // Sample(n) as the renderer's colour map: each channel / 255.
  std::vector<Vec4f> SampleColorMap(int n = 1024) const;

private:
  bool IsValid = true;
  bool Diverging = true;
  std::array<double, 3> Low{ 0.0, 0.0, 0.0 }, High{ 1.0, 1.0, 1.0 };
  std::vector<std::array<double, 2>> Alpha; // (x, alpha), kept sorted by x
};

// Msh coordinates of an sRGB colour (exposed for the tests) and back.
std::array<double, 3> RgbToMsh(const std::array<double, 3>& rgb);
std::array<double, 3> MshToRgb(const std::array<double, 3>& msh);
// Moreland's interpolation between two colours at t in [0, 1].
std::array<double, 3> DivergingInterpolate(const std::array<double, 3>& rgb1,
                                           const std::array<double, 3>& rgb2, double t);
} // namespace rendering
} // namespace beams

#endif
