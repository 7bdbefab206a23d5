// Point-based VTK PolyData writer in base64 binary form (.vtp), the output
// format of the reference solver (VTK::BINARY::writePoints,
// src/common/IO.h:414-507, called from LBMSolver::output,
// src/lbm/solver.cpp:323-382).
//
// File layout, kept identical so that files compare byte for byte:
//  * an XML header declaring LittleEndian and a UInt64 block header;
//  * Points: one Float32 array, 3 components per point (coordinates of 1D
//    and 2D grids padded with zeros, converted from double to float);
//  * Verts: Int64 connectivity 0..n-1 (binary), one offset n (ASCII);
//  * PointData: one Float64 array per field.
// Each binary array is the standard base64 encoding (RFC 4648 alphabet)
// of an 8-byte little-endian block header followed by the raw little-endian
// values.  The reference's header value is the element count times 8
// (src/common/util/base64.h:235), which is the byte size for 8-byte types and
// twice it for the Float32 points; it is reproduced here.  '=' padding
// follows the encoded length (one or two characters); the reference indexes
// past its padding table when no padding is due (src/common/IO.h:392-400),
// where this writer adds nothing.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lbmh::vtk {

// base64 of len bytes, without '=' padding
auto base64(const uint8_t* data, size_t len) -> std::string;

// One binary data array: block header (element count * 8, little endian) and
// the elements, base64-encoded and padded to a multiple of four characters.
template <typename T>
auto binaryArray(const T* data, int64_t n) -> std::string;

// A field value as the reference passes it to the writer: printed in fixed
// notation with 15 decimals and parsed back (toStringVector,
// include/common/util/string_helper.h:95-105; std::stod in IO.h:484).
auto textRoundTrip(double v) -> double;

struct Field {
  std::string name;
  std::vector<double> values;  // one per point, already filtered
};

// Write <fileBase>.vtp.  coords: n points of dim coordinates (dim 1..3).
// Returns false when the file cannot be opened.
auto writePoints(const std::string& fileBase, int dim, int64_t n, const double* coords,
                 const std::vector<Field>& fields) -> bool;

}  // namespace lbmh::vtk
