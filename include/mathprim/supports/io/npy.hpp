// mathprim/supports/io/npy.hpp -- io::numpy<T, N>: writes a host view as a NumPy .npy array (format 1.0,
// little-endian, C order), the writer the reference drivers use for their deform_<step>.npy frames
// (tests/fem_cu/main.cu:86-92; read back by tests/fem_time_step/visualize.py:9-12).
#pragma once

#include <cstdint>
#include <ostream>
#include <string>

#include "mathprim/core/view.hpp"

namespace mathprim::io {

template <typename T>
struct npy_descr;
template <>
struct npy_descr<double> {
  static constexpr const char* value = "<f8";
};
template <>
struct npy_descr<float> {
  static constexpr const char* value = "<f4";
};
template <>
struct npy_descr<int> {
  static constexpr const char* value = "<i4";
};

template <typename T, index_t N>
class numpy {
 public:
  template <typename U, typename S>
  void write(std::ostream& out, const contiguous_view<U, S, device::cpu>& v) const {
    static_assert(std::is_same_v<std::remove_const_t<U>, T>, "element type");
    static_assert(S::ndim == N, "rank");
    static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the elements are written as stored");
    std::string dict = std::string("{'descr': '") + npy_descr<T>::value + "', 'fortran_order': False, 'shape': (";
    for (index_t i = 0; i < N; ++i) dict += std::to_string(v.shape(i)) + (N == 1 || i + 1 < N ? ", " : "");
    dict += "), }";
    const size_t unpadded = 10 + dict.size() + 1;  // magic, version, length, dict, '\n'
    dict.append((64 - unpadded % 64) % 64, ' ');
    dict.push_back('\n');
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    out.write(magic, 8);
    const std::uint16_t hlen = static_cast<std::uint16_t>(dict.size());
    const char len_le[2] = {static_cast<char>(hlen & 0xff), static_cast<char>(hlen >> 8)};
    out.write(len_le, 2);
    out.write(dict.data(), static_cast<std::streamsize>(dict.size()));
    out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(sizeof(T) * v.numel()));
  }
};

}  // namespace mathprim::io
