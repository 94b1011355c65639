#include "ColorTable.h"

#include <algorithm>
#include <cctype>
#include <cmath>

namespace beams
{
namespace rendering
{
namespace
{
constexpr double kPi = 3.14159265358979323846;
// D65 reference white of CIELAB
constexpr double kXn = 0.95047, kYn = 1.0, kZn = 1.08883;

double SrgbToLinear(double c)
{
  return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
}
double LinearToSrgb(double c)
{
  return c <= 0.0031308 ? 12.92 * c : 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
}
double LabF(double t)
{
  return t > 216.0 / 24389.0 ? std::cbrt(t) : (24389.0 / 27.0 * t + 16.0) / 116.0;
}
double LabFInv(double f)
{
  const double t = f * f * f;
  return t > 216.0 / 24389.0 ? t : (116.0 * f - 16.0) / (24389.0 / 27.0);
}

std::array<double, 3> RgbToLab(const std::array<double, 3>& rgb)
{
  const double r = SrgbToLinear(rgb[0]), g = SrgbToLinear(rgb[1]), b = SrgbToLinear(rgb[2]);
  const double x = 0.4124564 * r + 0.3575761 * g + 0.1804375 * b;
  const double y = 0.2126729 * r + 0.7151522 * g + 0.0721750 * b;
  const double z = 0.0193339 * r + 0.1191920 * g + 0.9503041 * b;
  const double fx = LabF(x / kXn), fy = LabF(y / kYn), fz = LabF(z / kZn);
  return { 116.0 * fy - 16.0, 500.0 * (fx - fy), 200.0 * (fy - fz) };
}

std::array<double, 3> LabToRgb(const std::array<double, 3>& lab)
{
  const double fy = (lab[0] + 16.0) / 116.0, fx = fy + lab[1] / 500.0, fz = fy - lab[2] / 200.0;
  const double x = kXn * LabFInv(fx), y = kYn * LabFInv(fy), z = kZn * LabFInv(fz);
  const double r = 3.2404542 * x - 1.5371385 * y - 0.4985314 * z;
  const double g = -0.9692660 * x + 1.8760108 * y + 0.0415560 * z;
  const double b = 0.0556434 * x - 0.2040259 * y + 1.0572252 * z;
  return { LinearToSrgb(r), LinearToSrgb(g), LinearToSrgb(b) };
}

// Hue of the unsaturated end when interpolating to a saturated colour (Moreland, section 3.3)
double AdjustHue(const std::array<double, 3>& sat, double unsatM)
{
  if (sat[0] >= unsatM)
    return sat[2];
  const double spin = sat[1] * std::sqrt(unsatM * unsatM - sat[0] * sat[0]) / (sat[0] * std::sin(sat[1]));
  return sat[2] > -kPi / 3.0 ? sat[2] + spin : sat[2] - spin;
}

double AngleDiff(double a, double b)
{
  double d = std::fabs(a - b);
  while (d > 2.0 * kPi)
    d -= 2.0 * kPi;
  return d > kPi ? 2.0 * kPi - d : d;
}
} // namespace

std::array<double, 3> RgbToMsh(const std::array<double, 3>& rgb)
{
  const auto lab = RgbToLab(rgb);
  const double m = std::sqrt(lab[0] * lab[0] + lab[1] * lab[1] + lab[2] * lab[2]);
  const double s = m > 0.0 ? std::acos(std::min(1.0, std::max(-1.0, lab[0] / m))) : 0.0;
  const double h = (s > 0.0) ? std::atan2(lab[2], lab[1]) : 0.0;
  return { m, s, h };
}

std::array<double, 3> MshToRgb(const std::array<double, 3>& msh)
{
  const double l = msh[0] * std::cos(msh[1]);
  const double a = msh[0] * std::sin(msh[1]) * std::cos(msh[2]);
  const double b = msh[0] * std::sin(msh[1]) * std::sin(msh[2]);
  return LabToRgb({ l, a, b });
}

std::array<double, 3> DivergingInterpolate(const std::array<double, 3>& rgb1,
                                           const std::array<double, 3>& rgb2, double t)
{
  auto c1 = RgbToMsh(rgb1), c2 = RgbToMsh(rgb2);
  // two saturated, distinct hues: through unsaturated white in the middle
  if (c1[1] > 0.05 && c2[1] > 0.05 && AngleDiff(c1[2], c2[2]) > kPi / 3.0)
  {
    const double mid = std::max({ c1[0], c2[0], 88.0 });
    if (t < 0.5)
    {
      c2 = { mid, 0.0, 0.0 };
      t = 2.0 * t;
    }
    else
    {
      c1 = { mid, 0.0, 0.0 };
      t = 2.0 * t - 1.0;
    }
  }
  if (c1[1] < 0.05 && c2[1] > 0.05)
    c1[2] = AdjustHue(c2, c1[0]);
  else if (c2[1] < 0.05 && c1[1] > 0.05)
    c2[2] = AdjustHue(c1, c2[0]);
  const std::array<double, 3> m = { (1.0 - t) * c1[0] + t * c2[0], (1.0 - t) * c1[1] + t * c2[1],
                                    (1.0 - t) * c1[2] + t * c2[2] };
  return MshToRgb(m);
}

ColorTable::ColorTable(const std::string& presetName)
{
  std::string name = presetName;
  std::transform(name.begin(), name.end(), name.begin(),
                 [](unsigned char ch) { return static_cast<char>(std::tolower(ch)); });
  if (name.empty() || name == "cool to warm")
  {
    this->Diverging = true;
    this->Low = { 0.230, 0.299, 0.754 };
    this->High = { 0.706, 0.016, 0.150 };
  }
  else if (name == "grayscale" || name == "gray")
  {
    this->Diverging = false;
    this->Low = { 0.0, 0.0, 0.0 };
    this->High = { 1.0, 1.0, 1.0 };
  }
  else
  {
    this->IsValid = false;
  }
}

void ColorTable::AddPointAlpha(double x, float alpha)
{
  const std::array<double, 2> p = { x, static_cast<double>(alpha) };
  auto it = std::lower_bound(this->Alpha.begin(), this->Alpha.end(), p,
                             [](const std::array<double, 2>& a, const std::array<double, 2>& b) { return a[0] < b[0]; });
  if (it != this->Alpha.end() && (*it)[0] == x)
    *it = p; // a point at the same x replaces the earlier one
  else
    this->Alpha.insert(it, p);
}

std::array<double, 2> ColorTable::GetRange() const
{
  double lo = 0.0, hi = 1.0; // the preset's colour points
  for (const auto& p : this->Alpha)
  {
    lo = std::min(lo, p[0]);
    hi = std::max(hi, p[0]);
  }
  return { lo, hi };
}

std::array<double, 3> ColorTable::MapColor(double x) const
{
  const double t = std::min(1.0, std::max(0.0, x));
  if (this->Diverging)
    return DivergingInterpolate(this->Low, this->High, t);
  std::array<double, 3> c;
  for (int k = 0; k < 3; ++k)
    c[k] = (1.0 - t) * this->Low[k] + t * this->High[k];
  return c;
}

double ColorTable::MapAlpha(double x) const
{
  if (this->Alpha.empty())
    return 1.0;
  if (x <= this->Alpha.front()[0])
    return this->Alpha.front()[1];
  if (x >= this->Alpha.back()[0])
    return this->Alpha.back()[1];
  for (size_t i = 1; i < this->Alpha.size(); ++i)
    if (x <= this->Alpha[i][0])
    {
      const auto& a = this->Alpha[i - 1];
      const auto& b = this->Alpha[i];
      const double t = (x - a[0]) / (b[0] - a[0]);
      return (1.0 - t) * a[1] + t * b[1];
    }
  return this->Alpha.back()[1];
}

std::vector<std::array<uint8_t, 4>> ColorTable::Sample(int n) const
{
  std::vector<std::array<uint8_t, 4>> out;
  if (n < 2 || !this->IsValid)
    return out;
  const auto range = this->GetRange();
  auto q = [](double v) {
    const double c = std::min(1.0, std::max(0.0, v));
    return static_cast<uint8_t>(std::lround(c * 255.0));
  };
  out.resize(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i)
  {
    const double x = range[0] + (range[1] - range[0]) * static_cast<double>(i) / static_cast<double>(n - 1);
    const auto c = this->MapColor(x);
    out[static_cast<size_t>(i)] = { q(c[0]), q(c[1]), q(c[2]), q(this->MapAlpha(x)) };
  }
  return out;
}

std::vector<Vec4f> ColorTable::SampleColorMap(int n) const
{
  std::vector<Vec4f> out;
  for (const auto& s : this->Sample(n))
    out.push_back({ s[0] / 255.0f, s[1] / 255.0f, s[2] / 255.0f, s[3] / 255.0f });
  return out;
}
} // namespace rendering
} // namespace beams
