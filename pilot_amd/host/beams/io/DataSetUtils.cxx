#include "DataSetUtils.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace beams
{
namespace io
{
namespace
{
// ---- half-precision storage decoded for writing ------------------------------------------
float HalfToFloat(uint16_t h)
{
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0)
  {
    if (man == 0)
      bits = sign;
    else
    { // subnormal: normalise
      int e = -1;
      uint32_t m = man;
      do
      {
        ++e;
        m <<= 1;
      } while ((m & 0x400u) == 0);
      bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((m & 0x3ffu) << 13);
    }
  }
  else if (exp == 0x1f)
    bits = sign | 0x7f800000u | (man << 13);
  else
    bits = sign | ((exp + 127 - 15) << 23) | (man << 13);
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

float DecodeValue(const Field& f, size_t i)
{
  if (f.Type == DType::F32)
    return f.Values[i];
  if (f.Type == DType::F16)
    return HalfToFloat(f.Half[i]);
  const uint32_t bits = (uint32_t)f.Half[i] << 16;
  float v;
  std::memcpy(&v, &bits, 4);
  return v;
}

// ---- reading ---------------------------------------------------------------------------------
struct Cursor
{
  std::string Data;
  size_t Pos = 0;
  bool Binary = false;

  bool AtEnd()
  {
    SkipSpace();
    return Pos >= Data.size();
  }
  void SkipSpace()
  {
    while (Pos < Data.size() && std::isspace(static_cast<unsigned char>(Data[Pos])))
      ++Pos;
  }
  std::string Token()
  {
    SkipSpace();
    const size_t b = Pos;
    while (Pos < Data.size() && !std::isspace(static_cast<unsigned char>(Data[Pos])))
      ++Pos;
    if (b == Pos)
      throw ErrorBadValue("unexpected end of file");
    return Data.substr(b, Pos - b);
  }
  std::string Line()
  {
    const size_t e = Data.find('\n', Pos);
    std::string l = Data.substr(Pos, e == std::string::npos ? std::string::npos : e - Pos);
    Pos = e == std::string::npos ? Data.size() : e + 1;
    if (!l.empty() && l.back() == '\r')
      l.pop_back();
    return l;
  }
  void SkipToLineEnd()
  {
    while (Pos < Data.size() && Data[Pos] != '\n')
      ++Pos;
    if (Pos < Data.size())
      ++Pos;
  }
  int64_t Int() { return std::stoll(Token()); }
  double Real() { return std::stod(Token()); }
};

size_t TypeSize(const std::string& t)
{
  if (t == "char" || t == "unsigned_char")
    return 1;
  if (t == "short" || t == "unsigned_short")
    return 2;
  if (t == "int" || t == "unsigned_int" || t == "float")
    return 4;
  if (t == "long" || t == "unsigned_long" || t == "double" || t == "vtktypeint64" || t == "vtktypeuint64")
    return 8;
  throw ErrorBadValue("unsupported VTK data type " + t);
}

// n values of VTK type t (big-endian when binary), as doubles.
std::vector<double> ReadValues(Cursor& c, int64_t n, const std::string& t)
{
  std::vector<double> out(static_cast<size_t>(n));
  if (!c.Binary)
  {
    for (auto& v : out)
      v = c.Real();
    return out;
  }
  const size_t sz = TypeSize(t);
  if (c.Pos + sz * out.size() > c.Data.size())
    throw ErrorBadValue("binary data runs past the end of the file");
  const unsigned char* p = reinterpret_cast<const unsigned char*>(c.Data.data() + c.Pos);
  for (size_t i = 0; i < out.size(); ++i, p += sz)
  {
    uint64_t u = 0;
    for (size_t b = 0; b < sz; ++b)
      u = (u << 8) | p[b];
    double v;
    if (t == "float")
    {
      const uint32_t w = static_cast<uint32_t>(u);
      float f;
      std::memcpy(&f, &w, 4);
      v = f;
    }
    else if (t == "double")
      std::memcpy(&v, &u, 8);
    else if (t == "char")
      v = static_cast<int8_t>(u);
    else if (t == "short")
      v = static_cast<int16_t>(u);
    else if (t == "int")
      v = static_cast<int32_t>(u);
    else if (t == "long" || t == "vtktypeint64")
      v = static_cast<double>(static_cast<int64_t>(u));
    else
      v = static_cast<double>(u);
    out[i] = v;
  }
  c.Pos += sz * out.size();
  return out;
}

// In binary files the data starts after the keyword line; an optional LOOKUP_TABLE line first.
void EndKeywordLine(Cursor& c) { c.SkipToLineEnd(); }

void SkipLookupTable(Cursor& c)
{
  size_t p = c.Pos;
  while (p < c.Data.size() && (c.Data[p] == ' ' || c.Data[p] == '\t' || c.Data[p] == '\r' || c.Data[p] == '\n'))
    ++p;
  if (c.Data.compare(p, 12, "LOOKUP_TABLE") == 0)
  {
    c.Pos = p;
    c.Token();
    c.Token();
    EndKeywordLine(c);
  }
}

bool IsUniform(const std::vector<float>& x, float& origin, float& spacing)
{
  origin = x.front();
  spacing = x.size() > 1 ? x[1] - x[0] : 1.0f;
  for (size_t i = 0; i < x.size(); ++i)
    if (x[i] != origin + spacing * static_cast<float>(i))
      return false;
  return true;
}

void PutValues(std::ostream& os, const float* v, size_t n, FileType type)
{
  if (type == FileType::ASCII)
  {
    char buf[32];
    for (size_t i = 0; i < n; ++i)
    {
      std::snprintf(buf, sizeof(buf), "%.9g", static_cast<double>(v[i]));
      os << buf << ((i % 9 == 8 || i + 1 == n) ? '\n' : ' ');
    }
    return;
  }
  std::vector<unsigned char> be(n * 4);
  for (size_t i = 0; i < n; ++i)
  {
    uint32_t u;
    std::memcpy(&u, &v[i], 4);
    be[4 * i] = static_cast<unsigned char>(u >> 24);
    be[4 * i + 1] = static_cast<unsigned char>(u >> 16);
    be[4 * i + 2] = static_cast<unsigned char>(u >> 8);
    be[4 * i + 3] = static_cast<unsigned char>(u);
  }
  os.write(reinterpret_cast<const char*>(be.data()), static_cast<std::streamsize>(be.size()));
  os << '\n';
}
} // namespace

ReadResult DataSetUtils::Read(const std::string& name)
{
  ReadResult r;
  try
  {
    std::ifstream in(name, std::ios::binary);
    if (!in)
      throw ErrorBadValue("cannot open " + name);
    Cursor c;
    c.Data.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    if (c.Line().rfind("# vtk DataFile", 0) != 0)
      throw ErrorBadValue(name + ": not a VTK legacy file");
    c.Line(); // title
    const std::string mode = c.Token();
    if (mode == "BINARY")
      c.Binary = true;
    else if (mode != "ASCII")
      throw ErrorBadValue(name + ": expected ASCII or BINARY, got " + mode);
    if (c.Token() != "DATASET")
      throw ErrorBadValue(name + ": expected DATASET");
    const std::string kind = c.Token();
    if (kind != "STRUCTURED_POINTS" && kind != "RECTILINEAR_GRID")
      throw ErrorBadValue(name + ": unsupported dataset " + kind +
                          " (STRUCTURED_POINTS or RECTILINEAR_GRID)");
    std::array<int64_t, 3> dims = { 0, 0, 0 };
    std::array<double, 3> origin = { 0, 0, 0 }, spacing = { 1, 1, 1 };
    std::array<std::vector<double>, 3> axes;
    Association assoc = Association::Points;
    int64_t count = 0;
    DataSet& ds = r.Value;
    while (!c.AtEnd())
    {
      const std::string kw = c.Token();
      if (kw == "DIMENSIONS")
      {
        for (auto& d : dims)
          d = c.Int();
      }
      else if (kw == "ORIGIN")
      {
        for (auto& o : origin)
          o = c.Real();
      }
      else if (kw == "SPACING" || kw == "ASPECT_RATIO")
      {
        for (auto& s : spacing)
          s = c.Real();
      }
      else if (kw == "X_COORDINATES" || kw == "Y_COORDINATES" || kw == "Z_COORDINATES")
      {
        const int a = kw[0] - 'X';
        const int64_t n = c.Int();
        const std::string t = c.Token();
        EndKeywordLine(c);
        axes[a] = ReadValues(c, n, t);
      }
      else if (kw == "POINT_DATA" || kw == "CELL_DATA")
      {
        assoc = kw == "POINT_DATA" ? Association::Points : Association::Cells;
        count = c.Int();
      }
      else if (kw == "SCALARS")
      {
        Field f;
        f.Name = c.Token();
        const std::string t = c.Token();
        // optional component count on the same line
        size_t p = c.Pos;
        while (p < c.Data.size() && (c.Data[p] == ' ' || c.Data[p] == '\t'))
          ++p;
        int64_t ncomp = 1;
        if (p < c.Data.size() && std::isdigit(static_cast<unsigned char>(c.Data[p])))
          ncomp = c.Int();
        if (ncomp != 1)
          throw ErrorBadValue(name + ": SCALARS " + f.Name + " has " + std::to_string(ncomp) +
                              " components; the renderer takes one");
        EndKeywordLine(c);
        SkipLookupTable(c);
        const auto v = ReadValues(c, count, t);
        f.Assoc = assoc;
        f.Type = DType::F32;
        f.Values.assign(v.begin(), v.end());
        ds.Fields.push_back(std::move(f));
      }
      else if (kw == "VECTORS" || kw == "NORMALS")
      {
        c.Token();
        const std::string t = c.Token();
        EndKeywordLine(c);
        ReadValues(c, 3 * count, t);
      }
      else if (kw == "FIELD")
      {
        c.Token();
        const int64_t narrays = c.Int();
        for (int64_t k = 0; k < narrays; ++k)
        {
          c.Token();
          const int64_t ncomp = c.Int(), ntup = c.Int();
          const std::string t = c.Token();
          EndKeywordLine(c);
          ReadValues(c, ncomp * ntup, t);
        }
      }
      else if (kw == "METADATA")
      {
        EndKeywordLine(c);
        while (c.Pos < c.Data.size() && !c.Line().empty())
        {
        }
      }
      else
        throw ErrorBadValue(name + ": unsupported keyword " + kw);
    }
    for (int d = 0; d < 3; ++d)
      if (dims[d] < 1)
        throw ErrorBadValue(name + ": missing DIMENSIONS");
    if (kind == "STRUCTURED_POINTS")
      ds.Coords = CoordinateSystem::Uniform(dims, origin, spacing);
    else
      for (int d = 0; d < 3; ++d)
      {
        if (static_cast<int64_t>(axes[d].size()) != dims[d])
          throw ErrorBadValue(name + ": coordinate count does not match DIMENSIONS");
        ds.Coords.Axes[d].assign(axes[d].begin(), axes[d].end());
      }
    ds.CellSet.PointDims = dims;
    const int64_t npts = dims[0] * dims[1] * dims[2];
    const int64_t ncells = std::max<int64_t>(dims[0] - 1, 1) * std::max<int64_t>(dims[1] - 1, 1) *
                           std::max<int64_t>(dims[2] - 1, 1);
    for (const auto& f : ds.Fields)
      if (static_cast<int64_t>(f.Values.size()) != (f.IsPointField() ? npts : ncells))
        throw ErrorBadValue(name + ": field " + f.Name + " has the wrong number of values");
    r.Success = true;
  }
  catch (const std::exception& e)
  {
    r = ReadResult();
    r.Error = e.what();
  }
  return r;
}

WriteResult DataSetUtils::Write(const DataSet& ds, const std::string& fileName, FileType fileType)
{
  WriteResult r;
  try
  {
    std::ofstream os(fileName, std::ios::binary);
    if (!os)
      throw ErrorBadValue("cannot create " + fileName);
    const auto& dims = ds.CellSet.PointDims;
    for (int d = 0; d < 3; ++d)
      if (static_cast<int64_t>(ds.Coords.Axes[d].size()) != dims[d])
        throw ErrorBadValue("coordinates do not match the point dimensions");
    os << "# vtk DataFile Version 3.0\nbeams dataset\n"
       << (fileType == FileType::BINARY ? "BINARY" : "ASCII") << "\n";
    float o[3], s[3];
    bool uniform = true;
    for (int d = 0; d < 3; ++d)
      uniform = IsUniform(ds.Coords.Axes[d], o[d], s[d]) && uniform;
    char buf[128];
    if (uniform)
    {
      os << "DATASET STRUCTURED_POINTS\n";
      os << "DIMENSIONS " << dims[0] << ' ' << dims[1] << ' ' << dims[2] << '\n';
      std::snprintf(buf, sizeof(buf), "ORIGIN %.9g %.9g %.9g\nSPACING %.9g %.9g %.9g\n", (double)o[0],
                    (double)o[1], (double)o[2], (double)s[0], (double)s[1], (double)s[2]);
      os << buf;
    }
    else
    {
      os << "DATASET RECTILINEAR_GRID\n";
      os << "DIMENSIONS " << dims[0] << ' ' << dims[1] << ' ' << dims[2] << '\n';
      const char* names[3] = { "X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES" };
      for (int d = 0; d < 3; ++d)
      {
        os << names[d] << ' ' << dims[d] << " float\n";
        PutValues(os, ds.Coords.Axes[d].data(), ds.Coords.Axes[d].size(), fileType);
      }
    }
    const int64_t npts = dims[0] * dims[1] * dims[2];
    const int64_t ncells = std::max<int64_t>(dims[0] - 1, 1) * std::max<int64_t>(dims[1] - 1, 1) *
                           std::max<int64_t>(dims[2] - 1, 1);
    for (int pass = 0; pass < 2; ++pass)
    {
      const bool points = pass == 0;
      bool header = false;
      for (const auto& f : ds.Fields)
      {
        if (f.IsPointField() != points)
          continue;
        if (f.Device && f.Values.empty() && f.Half.empty())
          throw ErrorBadValue("field " + f.Name + " lives in device memory only");
        const int64_t n = points ? npts : ncells;
        if (!header)
        {
          os << (points ? "POINT_DATA " : "CELL_DATA ") << n << '\n';
          header = true;
        }
        std::vector<float> v(static_cast<size_t>(n));
        for (size_t i = 0; i < v.size(); ++i)
          v[i] = DecodeValue(f, i);
        os << "SCALARS " << f.Name << " float 1\nLOOKUP_TABLE default\n";
        PutValues(os, v.data(), v.size(), fileType);
      }
    }
    if (!os)
      throw ErrorBadValue("write to " + fileName + " failed");
    r.Success = true;
  }
  catch (const std::exception& e)
  {
    r.Success = false;
    r.Error = e.what();
  }
  return r;
}

std::vector<std::array<int64_t, 6>> ChunkPointRanges(const std::array<int64_t, 3>& pointDims,
                                                     int chunksX, int chunksY, int chunksZ)
{
  const int chunks[3] = { chunksX, chunksY, chunksZ };
  std::array<std::vector<int64_t>, 3> cuts;
  for (int d = 0; d < 3; ++d)
  {
    const int64_t cells = pointDims[d] - 1;
    if (chunks[d] < 1 || chunks[d] > cells)
      throw ErrorBadValue("chunks per axis must be in [1, cells]");
    for (int i = 0; i <= chunks[d]; ++i)
      cuts[d].push_back(cells * i / chunks[d]);
  }
  std::vector<std::array<int64_t, 6>> out;
  for (int z = 0; z < chunksZ; ++z)
    for (int y = 0; y < chunksY; ++y)
      for (int x = 0; x < chunksX; ++x)
        out.push_back({ cuts[0][x], cuts[0][x + 1], cuts[1][y], cuts[1][y + 1], cuts[2][z], cuts[2][z + 1] });
  return out;
}

std::vector<DataSet> ChunkDataSet(const DataSet& ds, int chunksX, int chunksY, int chunksZ)
{
  const auto& dims = ds.CellSet.PointDims;
  std::vector<DataSet> out;
  for (const auto& r : ChunkPointRanges(dims, chunksX, chunksY, chunksZ))
  {
    DataSet c;
    std::array<int64_t, 3> cd;
    for (int d = 0; d < 3; ++d)
    {
      cd[d] = r[2 * d + 1] - r[2 * d] + 1;
      c.Coords.Axes[d].assign(ds.Coords.Axes[d].begin() + r[2 * d], ds.Coords.Axes[d].begin() + r[2 * d + 1] + 1);
    }
    c.CellSet.PointDims = cd;
    for (const auto& f : ds.Fields)
    {
      if (f.Device && f.Values.empty() && f.Half.empty())
        throw ErrorBadValue("field " + f.Name + " lives in device memory only");
      Field g;
      g.Name = f.Name;
      g.Assoc = f.Assoc;
      g.Type = f.Type;
      const bool pts = f.IsPointField();
      const int64_t sx = pts ? dims[0] : dims[0] - 1, sy = pts ? dims[1] : dims[1] - 1;
      const int64_t ex = pts ? cd[0] : cd[0] - 1, ey = pts ? cd[1] : cd[1] - 1, ez = pts ? cd[2] : cd[2] - 1;
      for (int64_t k = 0; k < ez; ++k)
        for (int64_t j = 0; j < ey; ++j)
          for (int64_t i = 0; i < ex; ++i)
          {
            const size_t src = static_cast<size_t>((r[4] + k) * sx * sy + (r[2] + j) * sx + (r[0] + i));
            if (f.Type == DType::F32)
              g.Values.push_back(f.Values[src]);
            else
              g.Half.push_back(f.Half[src]);
          }
      c.Fields.push_back(std::move(g));
    }
    out.push_back(std::move(c));
  }
  return out;
}

namespace
{
// The source cell holding x on one axis (the last cell for x on the upper bound) and the
// parametric coordinate in it; false when x lies outside beyond the tolerance.
bool LocateAxis(const std::vector<float>& ax, double x, double tol, int64_t& c, double& t)
{
  const double lo = ax.front(), hi = ax.back();
  if (x < lo - tol || x > hi + tol || !(x == x))
    return false;
  x = std::min(std::max(x, lo), hi);
  const int64_t n = static_cast<int64_t>(ax.size());
  if (n < 2)
  {
    c = 0;
    t = 0.0;
    return true;
  }
  const auto it = std::upper_bound(ax.begin(), ax.end(), static_cast<float>(x));
  c = std::min<int64_t>(std::max<int64_t>((it - ax.begin()) - 1, 0), n - 2);
  const double w = static_cast<double>(ax[c + 1]) - ax[c];
  t = w > 0.0 ? (x - ax[c]) / w : 0.0;
  return true;
}
} // namespace

std::vector<DataSet> ResampleChunks(const DataSet& ds, int chunksX, int chunksY, int chunksZ)
{
  const std::array<int, 3> chunks = { chunksX, chunksY, chunksZ };
  const auto& dims = ds.CellSet.PointDims;
  std::array<float, 3> origin, chunkSize, spacing;
  std::array<int64_t, 3> cdims, pdims;
  std::array<double, 3> tol;
  for (int d = 0; d < 3; ++d)
  {
    if (chunks[d] < 1 || dims[d] / chunks[d] < 1)
      throw ErrorBadValue("ResampleChunks: each axis needs 1 <= chunks <= points");
    const auto& ax = ds.Coords.Axes[d];
    origin[d] = ax.front();
    // convert/Chunk.cxx: bounds.X.Length() / chunksX in FP64 (vtkm::Bounds), rounded once
    chunkSize[d] = static_cast<float>((static_cast<double>(ax.back()) - static_cast<double>(ax.front())) /
                                      static_cast<double>(chunks[d]));
    cdims[d] = dims[d] / chunks[d];
    pdims[d] = cdims[d] + 1;
    spacing[d] = chunkSize[d] / static_cast<float>(cdims[d]);
    tol[d] = 1e-6 * std::max(1e-30, static_cast<double>(ax.back()) - ax.front());
  }
  for (const auto& f : ds.Fields)
    if (f.Device && f.Values.empty() && f.Half.empty())
      throw ErrorBadValue("field " + f.Name + " lives in device memory only");
  std::vector<DataSet> out;
  for (int z = 0; z < chunksZ; ++z)
    for (int y = 0; y < chunksY; ++y)
      for (int x = 0; x < chunksX; ++x)
      {
        const std::array<int, 3> idx = { x, y, z };
        DataSet c;
        c.CellSet.PointDims = pdims;
        for (int d = 0; d < 3; ++d)
        {
          const float o = origin[d] + static_cast<float>(idx[d]) * chunkSize[d];
          c.Coords.Axes[d].resize(static_cast<size_t>(pdims[d]));
          for (int64_t i = 0; i < pdims[d]; ++i)
            c.Coords.Axes[d][static_cast<size_t>(i)] = o + spacing[d] * static_cast<float>(i);
        }
        for (const auto& f : ds.Fields)
        {
          Field g;
          g.Name = f.Name;
          g.Assoc = Association::Points;
          g.Type = DType::F32;
          g.Values.resize(static_cast<size_t>(pdims[0] * pdims[1] * pdims[2]));
          const bool pts = f.IsPointField();
          const int64_t sx = pts ? dims[0] : dims[0] - 1, sy = pts ? dims[1] : dims[1] - 1;
          size_t o = 0;
          for (int64_t k = 0; k < pdims[2]; ++k)
            for (int64_t j = 0; j < pdims[1]; ++j)
              for (int64_t i = 0; i < pdims[0]; ++i, ++o)
              {
                const std::array<int64_t, 3> q = { i, j, k };
                std::array<int64_t, 3> cell;
                std::array<double, 3> t;
                bool found = true;
                for (int d = 0; d < 3 && found; ++d)
                  found = LocateAxis(ds.Coords.Axes[d], c.Coords.Axes[d][static_cast<size_t>(q[d])], tol[d],
                                     cell[d], t[d]);
                if (!found)
                {
                  g.Values[o] = std::numeric_limits<float>::quiet_NaN();
                  continue;
                }
                auto at = [&](int64_t a, int64_t b, int64_t e) {
                  return static_cast<double>(DecodeValue(f, static_cast<size_t>((e * sy + b) * sx + a)));
                };
                if (!pts)
                {
                  g.Values[o] = static_cast<float>(at(cell[0], cell[1], cell[2]));
                  continue;
                }
                double v = 0.0;
                for (int dz = 0; dz < 2; ++dz)
                  for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx)
                    {
                      const int64_t a = std::min(cell[0] + dx, dims[0] - 1);
                      const int64_t b = std::min(cell[1] + dy, dims[1] - 1);
                      const int64_t e = std::min(cell[2] + dz, dims[2] - 1);
                      const double w = (dx ? t[0] : 1.0 - t[0]) * (dy ? t[1] : 1.0 - t[1]) *
                                       (dz ? t[2] : 1.0 - t[2]);
                      v += w * at(a, b, e);
                    }
                g.Values[o] = static_cast<float>(v);
              }
          c.Fields.push_back(std::move(g));
        }
        out.push_back(std::move(c));
      }
  return out;
}

WriteResult SaveChunksToDisk(const std::vector<DataSet>& chunks, const std::string& prefix, FileType fileType)
{
  for (size_t i = 0; i < chunks.size(); ++i)
  {
    WriteResult r = DataSetUtils::Write(chunks[i], prefix + "_" + std::to_string(i) + ".vtk", fileType);
    if (!r.Success)
      return r;
  }
  WriteResult ok;
  ok.Success = true;
  return ok;
}
} // namespace io
} // namespace beams
