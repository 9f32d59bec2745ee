// 3c2e / 2c2e class kernels with la = 1 (all lb <= 1, all lc).
#include "int3c2e_impl.h"

namespace dfhip {

namespace {
template <int LB, int LC>
struct Reg {
  static void run(Int3Table& t) {
    t[1][LB][LC] = make_int3<1, LB, LC>();
    Reg<LB, LC - 1>::run(t);
  }
};
template <int LB>
struct Reg<LB, -1> {
  static void run(Int3Table& t) {
    if constexpr (LB > 0) Reg<LB - 1, LMAX_AUX>::run(t);
  }
};
}  // namespace

void register_int3c2e_l1(Int3Table& t) { Reg<1, LMAX_AUX>::run(t); }

}  // namespace dfhip
