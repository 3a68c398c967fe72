#include "npy.hpp"

#include <cstring>
#include <fstream>
#include <stdexcept>

namespace cmnpy {

size_t Array::count() const {
  size_t n = 1;
  for (size_t s : shape) n *= s;
  return n;
}

namespace {

// value text of `key` in the header dict (up to the next top-level comma)
std::string dict_value(const std::string& h, const std::string& key) {
  const std::string k1 = "'" + key + "'", k2 = "\"" + key + "\"";
  size_t pos = h.find(k1);
  size_t klen = k1.size();
  if (pos == std::string::npos) {
    pos = h.find(k2);
    klen = k2.size();
  }
  if (pos == std::string::npos) throw std::runtime_error("npy header lacks '" + key + "'");
  pos = h.find(':', pos + klen);
  if (pos == std::string::npos) throw std::runtime_error("npy header: malformed '" + key + "'");
  ++pos;
  while (pos < h.size() && h[pos] == ' ') ++pos;
  size_t end = pos;
  int depth = 0;
  for (; end < h.size(); ++end) {
    const char c = h[end];
    if (c == '(' || c == '[') ++depth;
    if (c == ')' || c == ']') --depth;
    if (depth == 0 && (c == ',' || c == '}')) break;
  }
  return h.substr(pos, end - pos);
}

}  // namespace

Array load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open `" + path + "'");
  char magic[6];
  if (!f.read(magic, 6) || std::memcmp(magic, "\x93NUMPY", 6) != 0)
    throw std::runtime_error("`" + path + "' is not a .npy file");
  unsigned char ver[2];
  if (!f.read(reinterpret_cast<char*>(ver), 2)) throw std::runtime_error("truncated .npy header");
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    unsigned char b[2];
    if (!f.read(reinterpret_cast<char*>(b), 2)) throw std::runtime_error("truncated .npy header");
    hlen = uint32_t(b[0]) | (uint32_t(b[1]) << 8);
  } else if (ver[0] == 2 || ver[0] == 3) {
    unsigned char b[4];
    if (!f.read(reinterpret_cast<char*>(b), 4)) throw std::runtime_error("truncated .npy header");
    hlen = uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24);
  } else {
    throw std::runtime_error("unsupported .npy version " + std::to_string(ver[0]));
  }
  std::string h(hlen, '\0');
  if (!f.read(&h[0], hlen)) throw std::runtime_error("truncated .npy header");

  Array a;
  std::string d = dict_value(h, "descr");
  if (d.size() < 2 || (d.front() != '\'' && d.front() != '"'))
    throw std::runtime_error("npy header: malformed descr");
  a.descr = d.substr(1, d.size() - 2);
  if (a.descr.size() < 3) throw std::runtime_error("npy header: unsupported descr " + a.descr);
  a.word_size = size_t(std::stoul(a.descr.substr(2)));
  const std::string fo = dict_value(h, "fortran_order");
  if (fo.rfind("True", 0) == 0)
    a.fortran_order = true;
  else if (fo.rfind("False", 0) == 0)
    a.fortran_order = false;
  else
    throw std::runtime_error("npy header: malformed fortran_order");
  const std::string sh = dict_value(h, "shape");
  if (sh.empty() || sh.front() != '(') throw std::runtime_error("npy header: malformed shape");
  size_t i = 1;
  while (i < sh.size() && sh[i] != ')') {
    while (i < sh.size() && (sh[i] == ' ' || sh[i] == ',')) ++i;
    if (i >= sh.size() || sh[i] == ')') break;
    size_t j = i;
    while (j < sh.size() && sh[j] >= '0' && sh[j] <= '9') ++j;
    if (j == i) throw std::runtime_error("npy header: malformed shape");
    a.shape.push_back(size_t(std::stoull(sh.substr(i, j - i))));
    i = j;
  }
  const size_t bytes = a.count() * a.word_size;
  a.data.resize(bytes);
  if (bytes && !f.read(a.data.data(), std::streamsize(bytes)))
    throw std::runtime_error("`" + path + "': data shorter than its header says");
  return a;
}

void save_f64(const std::string& path, const double* values, const std::vector<size_t>& shape) {
  std::string dims;
  size_t n = 1;
  for (size_t k = 0; k < shape.size(); ++k) {
    if (k) dims += ", ";
    dims += std::to_string(shape[k]);
    n *= shape[k];
  }
  if (shape.size() == 1) dims += ",";
  std::string dict = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + dims + "), }";
  // pad so that magic + version + length + header is a multiple of 64
  size_t total = 10 + dict.size() + 1;
  const size_t pad = (64 - total % 64) % 64;
  dict.append(pad, ' ');
  dict += '\n';
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) throw std::runtime_error("cannot write `" + path + "'");
  f.write("\x93NUMPY\x01\x00", 8);
  const uint16_t hl = uint16_t(dict.size());
  const char lb[2] = {char(hl & 0xff), char(hl >> 8)};
  f.write(lb, 2);
  f.write(dict.data(), std::streamsize(dict.size()));
  f.write(reinterpret_cast<const char*>(values), std::streamsize(n * sizeof(double)));
  if (!f) throw std::runtime_error("writing `" + path + "' failed");
}

}  // namespace cmnpy
