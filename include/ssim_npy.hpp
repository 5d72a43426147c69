// ssim_npy.hpp -- header-only writer of NumPy .npy files (format version 1.0, little-endian float64, C order),
// the on-disk format of the reference drivers' per-step positions: tests/fem_cu/main.cu:86-92 writes
// deform_<step>.npy holding X + u as an [nv, 3] array through mathprim's io::numpy writer, and
// tests/fem_time_step/visualize.py:9-12 reads those frames back with numpy.load.
#ifndef SSIM_NPY_HPP
#define SSIM_NPY_HPP

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>

namespace ssim::hip::io {

// Writes rows × cols doubles (row-major) to `path` as a .npy array of shape (rows, cols).
inline void write_npy(const std::string& path, const double* data, std::size_t rows, std::size_t cols) {
  std::string dict = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                     std::to_string(cols) + "), }";
  // magic (6) + version (2) + header length (2) + dict + padding + '\n' is a multiple of 64 bytes
  const std::size_t unpadded = 10 + dict.size() + 1;
  dict.append((64 - unpadded % 64) % 64, ' ');
  dict.push_back('\n');
  if (dict.size() > 65535) throw std::invalid_argument("npy header too long");
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot open " + path);
  const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
  out.write(magic, 8);
  const std::uint16_t hlen = static_cast<std::uint16_t>(dict.size());
  const char len_le[2] = {static_cast<char>(hlen & 0xff), static_cast<char>(hlen >> 8)};
  out.write(len_le, 2);
  out.write(dict.data(), static_cast<std::streamsize>(dict.size()));
  static_assert(sizeof(double) == 8, "float64");
  static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the doubles are written as stored: '<f8' needs a little-endian host");
  out.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(sizeof(double) * rows * cols));
  if (!out) throw std::runtime_error("write failed: " + path);
}

}  // namespace ssim::hip::io

#endif  // SSIM_NPY_HPP
