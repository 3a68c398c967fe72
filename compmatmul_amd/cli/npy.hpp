// Minimal .npy reader/writer for the multiply tool (the format of NumPy's
// numpy.lib.format, versions 1.0-3.0): magic "\x93NUMPY", version bytes, a
// little-endian header length (2 bytes for 1.0, 4 for 2.0/3.0), then a
// Python dict literal with 'descr', 'fortran_order' and 'shape', padded
// with spaces and a newline, then the raw array data.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace cmnpy {

struct Array {
  std::string descr;           // e.g. "<f8"
  std::vector<size_t> shape;
  bool fortran_order = false;
  size_t word_size = 0;        // bytes per element, from descr
  std::vector<char> data;      // raw element bytes

  size_t count() const;
  template <class T>
  T* as() { return reinterpret_cast<T*>(data.data()); }
};

// Throws std::runtime_error on unreadable or malformed files.
Array load(const std::string& path);
// Writes a C-order little-endian float64 array of the given shape.
void save_f64(const std::string& path, const double* values, const std::vector<size_t>& shape);

}  // namespace cmnpy
