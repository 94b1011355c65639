// VTK legacy dataset files for the renderer's data model: the read/write of
// pilot::io::DataSetUtils (pilot/io/DataSetUtils.cxx:13-42, vtkm::io::VTKDataSetReader /
// VTKDataSetWriter) and the per-rank file scenes (FileSceneBase.cxx:33-49), restated for
// the structured grids the renderer takes.
//
// Read: "DATASET STRUCTURED_POINTS" (DIMENSIONS, ORIGIN, SPACING or ASPECT_RATIO) and
// "DATASET RECTILINEAR_GRID" (DIMENSIONS, X/Y/Z_COORDINATES), ASCII or BINARY (big-endian),
// with POINT_DATA / CELL_DATA sections of one-component SCALARS (char, unsigned_char, short,
// unsigned_short, int, unsigned_int, long, unsigned_long, float, double; any of them becomes
// an FP32 field).  VECTORS, NORMALS and FIELD arrays are skipped.
// Write: STRUCTURED_POINTS when every axis is uniform (to FP32 rounding), else
// RECTILINEAR_GRID; every field as float SCALARS.
#ifndef beams_io_datasetutils_h
#define beams_io_datasetutils_h

#include "../DataModel.h"

#include <array>
#include <string>
#include <vector>

namespace beams
{
namespace io
{
enum class FileType
{
  ASCII,
  BINARY
};

struct ReadResult
{
  bool Success = false;
  DataSet Value;
  std::string Error;
};

struct WriteResult
{
  bool Success = false;
  std::string Error;
};

struct DataSetUtils
{
  static ReadResult Read(const std::string& name);
  static WriteResult Write(const DataSet& dataSet, const std::string& fileName,
                           FileType fileType = FileType::BINARY);
};

// convert's chunking (convert/Chunk.cxx:48-152) for the brick path: chunks x chunks y chunks z
// blocks of a point grid as inclusive point ranges (i0,i1,j0,j1,k0,k1), neighbours sharing
// their boundary plane, sizes differing by at most one cell per axis.  These ranges cut the
// grid itself, so no value is interpolated; ResampleChunks below is the reference's own
// resampled chunking.
std::vector<std::array<int64_t, 6>> ChunkPointRanges(const std::array<int64_t, 3>& pointDims,
                                                     int chunksX, int chunksY, int chunksZ);

// The chunks as separate datasets (coordinates and point fields copied), as convert writes
// one file per rank; SaveChunksToDisk writes them as <prefix>_<index>.vtk.
std::vector<DataSet> ChunkDataSet(const DataSet& dataSet, int chunksX, int chunksY, int chunksZ);
WriteResult SaveChunksToDisk(const std::vector<DataSet>& chunks, const std::string& fileNamePrefix,
                             FileType fileType = FileType::BINARY);

// convert's own chunking (convert/Chunk.cxx:48-152): chunk (x, y, z), x fastest, gets a
// uniform grid of dims/chunks + 1 points per axis (dims = the source's point counts, integer
// division), spacing = (bounds length / chunks) / (dims / chunks), origin = bounds minimum +
// chunk index x bounds length / chunks, in FP32 as vtkm::Vec3f holds them.  Fields are probed
// onto it (vtkm::filter::resampling::Probe): a point field trilinearly from the source cell
// that holds the point, a cell field from that cell's value; FP32 output.  A point outside
// the source bounds by more than 1e-6 of the extent gets NaN (Probe's invalid value); closer
// ones are clamped onto the bounds.  VTK-m's locator tolerance and interpolation rounding are
// not reproduced.
std::vector<DataSet> ResampleChunks(const DataSet& dataSet, int chunksX, int chunksY, int chunksZ);
} // namespace io
} // namespace beams

#endif
