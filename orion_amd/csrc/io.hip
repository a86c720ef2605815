// io.hip -- host-only file output in the reference's formats.
//
//   orion_hip_write_vts  PostProcess::write_block_vts_binary
//                    (src/postprocess/PostProcess.cpp:126-252): a VTK XML
//                    StructuredGrid file, little endian, UInt32 headers, all
//                    arrays appended raw: float32 points, then the point
//                    fields Density, Pressure, Temperature, Mach, Velocity.
//   orion_hip_write_vtm  PostProcess::write_global_vtm (:254-283): the
//                    vtkMultiBlockDataSet index of one step.
//
// The bytes are the reference's: the same XML text, the same offsets
// (each appended array is a 4-byte byte count followed by its data) and the
// same float values (the double fields rounded to float).
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/orion_hip.h"

namespace orion_hip {
int fail(const char* fmt, ...);
}

namespace {

// one appended array: its byte count as UInt32, then the data
void put_array(std::ofstream& out, const float* data, size_t count) {
  const uint32_t bytes = static_cast<uint32_t>(count * sizeof(float));
  out.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
  if (count) out.write(reinterpret_cast<const char*>(data), bytes);
}

}  // namespace

int orion_hip_write_vts(const char* path, int idim, int jdim, int kdim, const double* x, const double* y,
                        const double* z, const float* fields) {
  if (!path) return orion_hip::fail("write_vts: no path");
  std::ofstream out(path, std::ios::out | std::ios::binary);
  if (!out.is_open()) return orion_hip::fail("write_vts: cannot open %s", path);
  if (idim <= 0 || jdim <= 0 || kdim <= 0) return 0;  // an empty file, as the reference leaves it
  if (!x || !y || !z || !fields) return orion_hip::fail("write_vts: null array");
  const size_t n = (size_t)idim * jdim * kdim;
  // byte offsets of the six appended arrays
  size_t off[6], at = 0;
  const size_t comps[6] = {3, 1, 1, 1, 1, 3};
  for (int a = 0; a < 6; ++a) {
    off[a] = at;
    at += sizeof(uint32_t) + n * comps[a] * sizeof(float);
  }
  const int ei = idim - 1, ej = jdim - 1, ek = kdim - 1;
  out << "<?xml version=\"1.0\"?>\n";
  out << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n";
  out << "  <StructuredGrid WholeExtent=\"0 " << ei << " 0 " << ej << " 0 " << ek << "\">\n";
  out << "    <Piece Extent=\"0 " << ei << " 0 " << ej << " 0 " << ek << "\">\n";
  out << "      <Points>\n";
  out << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << off[0]
      << "\"/>\n";
  out << "      </Points>\n";
  out << "      <PointData Scalars=\"Density\" Vectors=\"Velocity\">\n";
  const char* names[4] = {"Density", "Pressure", "Temperature", "Mach"};
  for (int a = 0; a < 4; ++a)
    out << "        <DataArray type=\"Float32\" Name=\"" << names[a] << "\" format=\"appended\" offset=\""
        << off[1 + a] << "\"/>\n";
  out << "        <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
      << off[5] << "\"/>\n";
  out << "      </PointData>\n";
  out << "    </Piece>\n";
  out << "  </StructuredGrid>\n";
  out << "  <AppendedData encoding=\"raw\">\n";
  out << "    _";
  {
    std::vector<float> pts(3 * n);
    for (size_t g = 0; g < n; ++g) {
      pts[3 * g] = (float)x[g];
      pts[3 * g + 1] = (float)y[g];
      pts[3 * g + 2] = (float)z[g];
    }
    put_array(out, pts.data(), pts.size());
  }
  for (int a = 0; a < 4; ++a) put_array(out, fields + a * n, n);
  put_array(out, fields + 4 * n, 3 * n);
  out << "\n  </AppendedData>\n";
  out << "</VTKFile>\n";
  out.close();
  if (!out) return orion_hip::fail("write_vts: write to %s failed", path);
  return 0;
}

int orion_hip_write_vtm(const char* output_dir, const char* stem, int nblocks, int step) {
  const std::string dir = output_dir ? output_dir : ".", base = stem ? stem : "";
  std::ostringstream name;
  name << dir << "/" << base << "_" << std::setfill('0') << std::setw(6) << step << ".vtm";
  std::ofstream out(name.str());
  if (!out.is_open()) return orion_hip::fail("write_vtm: cannot open %s", name.str().c_str());
  out << "<?xml version=\"1.0\"?>\n";
  out << "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" byte_order=\"LittleEndian\">\n";
  out << "  <vtkMultiBlockDataSet>\n";
  for (int b = 0; b < nblocks; ++b) {
    std::ostringstream vts;
    vts << base << "_" << std::setfill('0') << std::setw(6) << step << "_block_" << b << ".vts";
    out << "    <DataSet index=\"" << b << "\" file=\"" << vts.str() << "\"/>\n";
  }
  out << "  </vtkMultiBlockDataSet>\n";
  out << "</VTKFile>\n";
  out.close();
  if (!out) return orion_hip::fail("write_vtm: write to %s failed", name.str().c_str());
  return 0;
}
