#include "vtk.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

namespace lbmh::vtk {

auto base64(const uint8_t* data, size_t len) -> std::string {
  static const char table[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((len + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 3 <= len; i += 3) {
    const uint32_t w = (uint32_t)data[i] << 16 | (uint32_t)data[i + 1] << 8 | data[i + 2];
    out += table[w >> 18];
    out += table[(w >> 12) & 63];
    out += table[(w >> 6) & 63];
    out += table[w & 63];
  }
  const size_t rest = len - i;
  if (rest == 1) {
    const uint32_t w = (uint32_t)data[i] << 16;
    out += table[w >> 18];
    out += table[(w >> 12) & 63];
  } else if (rest == 2) {
    const uint32_t w = (uint32_t)data[i] << 16 | (uint32_t)data[i + 1] << 8;
    out += table[w >> 18];
    out += table[(w >> 12) & 63];
    out += table[(w >> 6) & 63];
  }
  return out;
}

template <typename T>
auto binaryArray(const T* data, int64_t n) -> std::string {
  static_assert(sizeof(T) <= 8, "scalar element types only");
  std::vector<uint8_t> bytes(8 + sizeof(T) * (size_t)n);
  const uint64_t header = (uint64_t)n * 8;
  for (int b = 0; b < 8; ++b) bytes[b] = (uint8_t)(header >> (8 * b));
  // host byte order is little endian on every target of this build
  if (n > 0) std::memcpy(bytes.data() + 8, data, sizeof(T) * (size_t)n);
  std::string s = base64(bytes.data(), bytes.size());
  if (s.size() % 4 != 0) s.append(4 - s.size() % 4, '=');
  return s;
}
template auto binaryArray<float>(const float*, int64_t) -> std::string;
template auto binaryArray<double>(const double*, int64_t) -> std::string;
template auto binaryArray<int64_t>(const int64_t*, int64_t) -> std::string;
template auto binaryArray<int32_t>(const int32_t*, int64_t) -> std::string;

auto textRoundTrip(double v) -> double {
  // fixed notation, 15 decimals: at most 309 integer digits for finite values
  char buf[400];
  std::snprintf(buf, sizeof(buf), "%.15f", v);
  return std::strtod(buf, nullptr);
}

auto writePoints(const std::string& fileBase, int dim, int64_t n, const double* coords,
                 const std::vector<Field>& fields) -> bool {
  std::cerr << " Writing " << fileBase << ".vtp with #" << n << " cells" << std::endl;
  std::ofstream f(fileBase + ".vtp", std::ios::binary);
  if (!f) return false;
  f << "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
       "  <PolyData>\n";
  f << "<Piece NumberOfPoints=\"" << n
    << "\" NumberOfVerts=\"1\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\" > \n";
  f << "<Points>\n<DataArray type=\"Float32\" Name=\"Points\" NumberOfComponents=\"3\" format=\"binary\"> \n";
  {
    std::vector<float> xyz((size_t)n * 3, 0.0f);
    for (int64_t p = 0; p < n; ++p)
      for (int d = 0; d < dim; ++d) xyz[(size_t)p * 3 + d] = (float)coords[(size_t)p * dim + d];
    f << binaryArray(xyz.data(), n * 3);
  }
  f << "\n        </DataArray>\n      </Points>\n";
  f << "      <Verts>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"binary\"> \n";
  {
    std::vector<int64_t> ids((size_t)n);
    for (int64_t p = 0; p < n; ++p) ids[(size_t)p] = p;
    f << binaryArray(ids.data(), n);
  }
  f << "\n        </DataArray> \n";
  f << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\"> \n" << n << "\n        </DataArray> \n";
  f << "        </Verts> \n";
  f << "      <PointData> \n";
  for (const Field& fd : fields) {
    f << "<DataArray type=\"Float64\" Name=\"" << fd.name << "\" format=\"binary\">\n";
    f << binaryArray(fd.values.data(), (int64_t)fd.values.size());
    f << "\n        </DataArray> \n";
  }
  f << "      </PointData> \n";
  f << "    </Piece>\n  </PolyData>\n</VTKFile> \n";
  return (bool)f;
}

}  // namespace lbmh::vtk
