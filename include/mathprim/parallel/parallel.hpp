// mathprim/parallel/parallel.hpp -- the parfor interface of ssim's ParImpl template argument:
// run(n, f) calls f(i) for i in [0, n); run(task) runs a task object; vmap(f, views...) calls f(v0[i], v1[i], ...)
// over the leading dimension of the views.  par::seq runs on the calling host thread.
#pragma once

#include "mathprim/core/functional.hpp"  // the functors vmap applies (affine_transform)
#include "mathprim/core/view.hpp"

namespace mathprim::par {

template <typename Derived>
class parfor {
 public:
  Derived& derived() noexcept { return static_cast<Derived&>(*this); }
  const Derived& derived() const noexcept { return static_cast<const Derived&>(*this); }

  template <typename F>
  void run(index_t n, F&& f) const {
    derived().run_impl(n, std::forward<F>(f));
  }
  // a task: any object with run_impl(par) (e.g. the visit task of a sparse view)
  template <typename Task, typename = decltype(std::declval<Task&>().run_impl(std::declval<const Derived&>()))>
  void run(Task&& task) const {
    task.run_impl(derived());
  }
  template <typename F, typename View0, typename... Views>
  void vmap(F&& f, const View0& v0, const Views&... vs) const {
    derived().vmap_impl(std::forward<F>(f), v0, vs...);
  }
};

class seq : public parfor<seq> {
 public:
  template <typename F>
  void run_impl(index_t n, F&& f) const {
    for (index_t i = 0; i < n; ++i) f(i);
  }
  template <typename F, typename View0, typename... Views>
  void vmap_impl(F&& f, const View0& v0, const Views&... vs) const {
    const index_t n = v0.shape(0);
    for (index_t i = 0; i < n; ++i) f(v0[i], vs[i]...);
  }
};

}  // namespace mathprim::par
