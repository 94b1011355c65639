// Image files of rendered frames: the reference saves each frame with the canvas's SaveAs
// (beams/Beams.cxx:292-301: <scene id>_<rank>_<nnn>.png, vtkm::rendering::Canvas::SaveAs).
// Written here from a canvas's RGBA float colours (row 0 at the top, as the renderer's
// images are laid out): 8 bits per channel, each value clamped to [0, 1] and rounded to the
// nearest of k/255.  ".png" gives an RGBA PNG (a zlib stream of stored deflate blocks, so no
// compression library is needed); ".ppm" or ".pnm" a binary RGB PPM (P6).  VTK-m's own
// conversion and orientation are not reproduced; this is the build's documented convention.
#ifndef beams_io_imageutils_h
#define beams_io_imageutils_h

#include "DataSetUtils.h"

#include <cstdint>
#include <string>
#include <vector>

namespace beams
{
namespace io
{
// 8-bit value of one colour channel: round(clamp(v, 0, 1) * 255).
uint8_t QuantizeChannel(float v);

WriteResult SaveImage(const std::vector<float>& rgba, int width, int height,
                      const std::string& fileName);

// The file name the reference gives frame `num` of a scene on a rank.
std::string FrameFileName(const std::string& sceneId, int rank, int num,
                          const std::string& extension = ".png");
} // namespace io
} // namespace beams

#endif
