// mathprim/supports/stringify.hpp -- operator<< for shapes and views: a shape prints as (d0, d1, ...), a view as
// its device and shape (the contents of a device view are not read).
#pragma once

#include <ostream>

#include "mathprim/core/view.hpp"

namespace mathprim {

template <index_t... Dims>
std::ostream& operator<<(std::ostream& os, const shape_t<Dims...>& s) {
  os << '(';
  for (index_t i = 0; i < shape_t<Dims...>::ndim; ++i) os << (i ? ", " : "") << s(i);
  return os << ')';
}

template <typename T, typename S, typename D>
std::ostream& operator<<(std::ostream& os, const contiguous_view<T, S, D>& v) {
  return os << (std::is_same_v<D, device::cpu> ? "view<cpu>" : "view<device>") << v.shape();
}

}  // namespace mathprim
