#include "Config.h"

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

namespace beams
{
namespace
{
std::string Trim(const std::string& s)
{
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a])))
    ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1])))
    --b;
  return s.substr(a, b - a);
}

// Strips a trailing comment outside double quotes.
std::string StripComment(const std::string& s)
{
  bool quoted = false;
  for (size_t i = 0; i < s.size(); ++i)
  {
    if (s[i] == '"')
      quoted = !quoted;
    else if (s[i] == '#' && !quoted)
      return s.substr(0, i);
  }
  return s;
}

bool ParseNumbers(const std::string& v, std::vector<double>& out)
{
  std::string body = Trim(v);
  if (body.size() < 2 || body.front() != '[' || body.back() != ']')
    return false;
  body = body.substr(1, body.size() - 2);
  std::stringstream ss(body);
  std::string item;
  while (std::getline(ss, item, ','))
  {
    item = Trim(item);
    if (item.empty())
      continue;
    char* end = nullptr;
    const double d = std::strtod(item.c_str(), &end);
    if (end == item.c_str() || *end != '\0')
      return false;
    out.push_back(d);
  }
  return true;
}

Result Apply(Preset& p, const std::string& key, const std::string& raw, int line)
{
  auto fail = [&](const std::string& why) {
    return Result::Failed("line " + std::to_string(line) + ": " + key + ": " + why);
  };
  std::string v = Trim(raw);
  const bool isString = v.size() >= 2 && v.front() == '"' && v.back() == '"';
  const std::string str = isString ? v.substr(1, v.size() - 2) : v;
  std::vector<double> nums;
  const bool isArray = ParseNumbers(v, nums);
  auto num = [&](double& out) {
    char* end = nullptr;
    out = std::strtod(v.c_str(), &end);
    return end != v.c_str() && *end == '\0';
  };
  double d = 0.0;
  auto vec3 = [&](std::array<float, 3>& dst) {
    if (!isArray || nums.size() != 3)
      return false;
    for (int i = 0; i < 3; ++i)
      dst[i] = static_cast<float>(nums[i]);
    return true;
  };
  if (key == "dataset")
    p.DataSet.Factory = str;
  else if (key == "size" || key == "seed" || key == "blobs" || key == "noise_octaves" ||
           key == "noise_amp" || key == "dtype" || key == "time" || key == "drift_speed" ||
           key == "file_pattern" || key == "field")
    p.DataSet.Params[key] = str;
  else if (key == "map_enabled")
  {
    if (v != "true" && v != "false")
      return fail("expected true or false");
    p.OpacityMap.Enabled = (v == "true");
  }
  else if (key == "map_size")
  {
    if (!isArray || nums.size() != 3)
      return fail("expected [cx, cy, cz]");
    for (int i = 0; i < 3; ++i)
      p.OpacityMap.Size[i] = static_cast<int64_t>(nums[i]);
  }
  else if (key == "map_size_ratio")
  {
    if (!num(d))
      return fail("expected a number");
    p.OpacityMap.SizeRatio = static_cast<float>(d);
  }
  else if (key == "map_steps")
  {
    if (!num(d))
      return fail("expected a number");
    p.OpacityMap.NumSteps = static_cast<int>(d);
  }
  else if (key == "camera_azimuth" || key == "camera_elevation" || key == "camera_fov")
  {
    if (!num(d))
      return fail("expected a number");
    float& dst = key == "camera_azimuth" ? p.Camera.Azimuth
                                          : (key == "camera_elevation" ? p.Camera.Elevation
                                                                       : p.Camera.FieldOfView);
    dst = static_cast<float>(d);
  }
  else if (key == "light_type" || key == "light_position" || key == "light_color" ||
           key == "light_intensity")
  {
    if (p.Lights.empty())
      p.Lights.emplace_back();
    LightOption& l = p.Lights.front();
    if (key == "light_type")
    {
      if (str != "point" && str != "directional")
        return fail("light_type must be \"point\" or \"directional\"");
      l.Type = str;
    }
    else if (key == "light_position")
    {
      if (!vec3(l.Position))
        return fail("expected [x, y, z]");
    }
    else if (key == "light_color")
    {
      if (!vec3(l.Color))
        return fail("expected [r, g, b]");
    }
    else
    {
      if (!num(d))
        return fail("expected a number");
      l.Intensity = static_cast<float>(d);
    }
  }
  else if (key == "colormap_seed" || key == "colormap_points" || key == "colormap_entries")
  {
    if (!num(d))
      return fail("expected a number");
    if (key == "colormap_seed")
      p.ColorTable.Seed = static_cast<uint64_t>(d);
    else if (key == "colormap_points")
      p.ColorTable.NumPoints = static_cast<int>(d);
    else
      p.ColorTable.NumEntries = static_cast<int>(d);
  }
  else if (key == "color_table")
  {
    p.ColorTable.Name = str;
  }
  else if (key == "point_alphas")
  {
    if (!isArray || nums.size() % 2 != 0)
      return fail("expected [x0, alpha0, x1, alpha1, ...]");
    for (double v : nums)
      if (!std::isfinite(v))
        return fail("point_alphas must be finite");
    p.ColorTable.PointAlphas.clear();
    for (size_t i = 0; i < nums.size(); i += 2)
      p.ColorTable.PointAlphas.push_back({ static_cast<float>(nums[i]), static_cast<float>(nums[i + 1]) });
  }
  else if (key == "scalar_range")
  {
    if (!isArray || nums.size() != 2)
      return fail("expected [min, max]");
    p.ColorTable.RangeMin = static_cast<float>(nums[0]);
    p.ColorTable.RangeMax = static_cast<float>(nums[1]);
  }
  else if (key == "image")
  {
    if (!isArray || nums.size() != 2)
      return fail("expected [width, height]");
    p.ImageSize = { static_cast<int>(nums[0]), static_cast<int>(nums[1]) };
  }
  else if (key == "sample_distance" || key == "alpha_stop")
  {
    if (!num(d))
      return fail("expected a number");
    (key == "sample_distance" ? p.SampleDistance : p.AlphaStop) = static_cast<float>(d);
  }
  else
    return fail("unknown key");
  return Result::Succeeded();
}
} // namespace

Result Config::Deserialize(const std::string& tomlText)
{
  std::stringstream in(tomlText);
  std::string rawLine;
  Preset* cur = nullptr;
  int lineNo = 0;
  while (std::getline(in, rawLine))
  {
    ++lineNo;
    const std::string line = Trim(StripComment(rawLine));
    if (line.empty())
      continue;
    if (line.front() == '[')
    {
      if (line.back() != ']')
        return Result::Failed("line " + std::to_string(lineNo) + ": unterminated table header");
      const std::string name = Trim(line.substr(1, line.size() - 2));
      const std::string prefix = "presets.";
      if (name.compare(0, prefix.size(), prefix) != 0 || name.size() == prefix.size())
        return Result::Failed("line " + std::to_string(lineNo) + ": expected [presets.<id>]");
      const std::string id = name.substr(prefix.size());
      cur = &this->Presets[id];
      cur->Id = id;
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos)
      return Result::Failed("line " + std::to_string(lineNo) + ": expected key = value");
    const std::string key = Trim(line.substr(0, eq));
    const std::string value = Trim(line.substr(eq + 1));
    if (!cur)
    {
      if (key == "default_preset" && value.size() >= 2 && value.front() == '"')
      {
        this->DefaultPresetName = value.substr(1, value.size() - 2);
        continue;
      }
      return Result::Failed("line " + std::to_string(lineNo) + ": key outside a preset table");
    }
    Result r = Apply(*cur, key, value, lineNo);
    if (!r.Success)
      return r;
  }
  if (!this->DefaultPresetName.empty() && !this->Presets.count(this->DefaultPresetName))
    return Result::Failed("default_preset names no preset");
  return Result::Succeeded();
}

Result Config::DeserializeFile(const std::string& path)
{
  std::ifstream f(path);
  if (!f)
    return Result::Failed("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return this->Deserialize(ss.str());
}
} // namespace beams
