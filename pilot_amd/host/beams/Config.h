// Preset / Config surface of Beams (reference: beams/Config.h, whose preset schema is
// recovered from the commented JSON reader at beams/Config.cxx:140-266).  Presets are
// read from a small TOML subset: [presets.<id>] tables of `key = value` lines with
// strings, numbers, booleans and flat numeric arrays.
#ifndef beams_config_h
#define beams_config_h

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace beams
{
struct Result
{
  bool Success = false;
  std::string Err = "Unknown";
  static Result Failed(const std::string& err) { return Result{ false, err }; }
  static Result Succeeded() { return Result{ true, "" }; }
};

struct DataSetOptions
{
  std::string Factory = "blobs";                        // seeded Gaussian blobs + value noise
  std::unordered_map<std::string, std::string> Params;  // size, seed, blobs, noise_octaves, ...
};

struct OpacityMapOptions
{
  bool Enabled = true;
  std::array<int64_t, 3> Size{ -1, -1, -1 };  // cells; -1: SizeRatio * dataset cells
  float SizeRatio = 1.0f;
  int NumSteps = 128;  // parsed; the light-march step is |extent| / 128 (TransmittanceMap.h:584)

  std::array<int64_t, 3> CalculateSize(const std::array<int64_t, 3>& dataSetCells) const
  {
    if (Size[0] != -1)
      return Size;
    return { static_cast<int64_t>(dataSetCells[0] * SizeRatio),
             static_cast<int64_t>(dataSetCells[1] * SizeRatio),
             static_cast<int64_t>(dataSetCells[2] * SizeRatio) };
  }
};

struct CameraOptions
{
  std::string Type = "perspective";
  float Azimuth = 0.0f;
  float Elevation = 0.0f;
  float FieldOfView = 60.0f;
};

struct LightOption
{
  std::string Type = "point";  // "point" or "directional"
  std::array<float, 3> Position{ 2.0f, 2.0f, 2.0f };
  std::array<float, 3> Color{ 1.0f, 1.0f, 1.0f };
  float Intensity = 2.0f;
};

struct ColorTableOptions
{
  // "seeded": beams_gen_colormap's seeded table (Seed, NumPoints, NumEntries); otherwise a
  // ColorTable preset name ("cool to warm", "grayscale") with PointAlphas, sampled into
  // 1024 8-bit entries as Mapper::SetActiveColorTable does (FileSceneBase.cxx:67-88)
  std::string Name = "seeded";
  std::vector<std::array<float, 2>> PointAlphas;
  float RangeMin = 0.0f;
  float RangeMax = 1.0f;
  uint64_t Seed = 1;
  int NumPoints = 6;
  int NumEntries = 256;
};

struct Preset
{
  std::string Id;
  DataSetOptions DataSet;
  OpacityMapOptions OpacityMap;
  CameraOptions Camera;
  std::vector<LightOption> Lights;
  ColorTableOptions ColorTable;
  std::array<int, 2> ImageSize{ 512, 512 };
  float SampleDistance = -1.0f;  // <= 0: |extent| / 200
  float AlphaStop = 1.0f;
};

struct Config
{
  std::string DefaultPresetName;
  std::map<std::string, Preset> Presets;

  // Parses presets from TOML text; returns a failed Result with a line number on error.
  Result Deserialize(const std::string& tomlText);
  Result DeserializeFile(const std::string& path);
};
} // namespace beams

#endif
