// pass_inst.hip - instantiates the enumeration kernels of one word class.
// Built once per class with -DKCL_PASS_CLASS=<c> (c = 0..8).
#include "pass_kernel.h"

#ifndef KCL_PASS_CLASS
#error "build with -DKCL_PASS_CLASS=<word class index>"
#endif

namespace kcl {

namespace {

template <int C, int K>
struct KTable {
  static void fill(PassFn *t, PassFn *t_all) {
    constexpr int W = (int)kClassWords[C];
    t[K] = &pass_kernel<W, K, (int)class_waves(C, K), false>;
    t_all[K] = &pass_kernel<W, K, (int)class_waves(C, K, true), true>;
    if constexpr (K < kMaxTemplK) KTable<C, K + 1>::fill(t, t_all);
  }
};

}  // namespace

#define KCL_CAT2(a, b) a##b
#define KCL_CAT(a, b) KCL_CAT2(a, b)

void KCL_CAT(fill_pass_class_, KCL_PASS_CLASS)(PassFn *fn, PassFn *fn_all, PassFn *deep,
                                               size_t *lds) {
  constexpr int C = KCL_PASS_CLASS;
  constexpr int W = (int)kClassWords[C];
  KTable<C, 2>::fill(fn, fn_all);
  deep[0] = &pass_kernel<W, 0, (int)deep_waves(C), false>;
  deep[1] = &pass_kernel<W, 0, (int)deep_waves(C), true>;
  *lds = sizeof(Lds<W>);
}

}  // namespace kcl
