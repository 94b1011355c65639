// Minimal data model of the Beams renderer surface: what the reference takes as
// vtkm::cont::DataSet / CoordinateSystem / Field / CellSetStructured<3> / Range,
// without VTK-m.  Arrays are x-fastest, as vtkm point and cell ordering.
#ifndef beams_datamodel_h
#define beams_datamodel_h

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace beams
{
// vtkm::cont::ErrorBadValue / ErrorBadType of the reference surface.
struct ErrorBadValue : std::runtime_error
{
  using std::runtime_error::runtime_error;
};

struct Range
{
  double Min = 0.0;
  double Max = 0.0;
};

struct Bounds
{
  double X[2] = { 0, 0 }, Y[2] = { 0, 0 }, Z[2] = { 0, 0 };
  std::array<double, 3> Center() const
  {
    return { (X[0] + X[1]) * 0.5, (Y[0] + Y[1]) * 0.5, (Z[0] + Z[1]) * 0.5 };
  }
};

struct Vec4f
{
  float X, Y, Z, W;
};

// Rectilinear coordinates (ArrayHandleCartesianProduct); a uniform grid is held as
// x_i = float(origin) + float(spacing) * i, the values ArrayHandleCounting holds.
struct CoordinateSystem
{
  std::array<std::vector<float>, 3> Axes;

  static CoordinateSystem Uniform(const std::array<int64_t, 3>& dims,
                                  const std::array<double, 3>& origin,
                                  const std::array<double, 3>& spacing)
  {
    CoordinateSystem c;
    for (int d = 0; d < 3; ++d)
    {
      const float o = static_cast<float>(origin[d]), s = static_cast<float>(spacing[d]);
      c.Axes[d].resize(static_cast<size_t>(dims[d]));
      for (int64_t i = 0; i < dims[d]; ++i)
        c.Axes[d][static_cast<size_t>(i)] = o + s * static_cast<float>(i);
    }
    return c;
  }

  Bounds GetBounds() const
  {
    Bounds b;
    double* dst[3] = { b.X, b.Y, b.Z };
    for (int d = 0; d < 3; ++d)
    {
      if (Axes[d].empty())
        continue;
      dst[d][0] = Axes[d].front();
      dst[d][1] = Axes[d].back();
    }
    return b;
  }
};

struct CellSetStructured
{
  std::array<int64_t, 3> PointDims = { 0, 0, 0 };
};

enum class Association
{
  Points = 0,
  Cells = 1
};

// Storage types of a scalar field (arithmetic is always FP32).
enum class DType
{
  F32 = 0,
  F16 = 1,
  BF16 = 2
};

struct Field
{
  std::string Name;
  Association Assoc = Association::Points;
  DType Type = DType::F32;
  std::vector<uint16_t> Half;   // F16 / BF16 bit patterns (host)
  std::vector<float> Values;    // F32 values (host)
  const void* Device = nullptr; // optional: values already resident in HBM

  bool IsPointField() const { return Assoc == Association::Points; }
  bool IsCellField() const { return Assoc == Association::Cells; }
  const void* Data() const
  {
    if (Device)
      return Device;
    return Type == DType::F32 ? static_cast<const void*>(Values.data())
                              : static_cast<const void*>(Half.data());
  }
};

struct DataSet
{
  CoordinateSystem Coords;
  CellSetStructured CellSet;
  std::vector<Field> Fields;

  const Field& GetField(const std::string& name) const
  {
    for (const auto& f : Fields)
      if (f.Name == name)
        return f;
    throw ErrorBadValue("no field named " + name);
  }
};
} // namespace beams

#endif
