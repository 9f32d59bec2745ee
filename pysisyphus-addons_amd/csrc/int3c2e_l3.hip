// 3c2e / 2c2e class kernels with la = 3 (all lb <= 3, all lc).
#include "int3c2e_impl.h"

namespace dfhip {

namespace {
template <int LB, int LC>
struct Reg {
  static void run(Int3Table& t) {
    t[3][LB][LC] = make_int3<3, LB, LC>();
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

void register_int3c2e_l3(Int3Table& t) { Reg<3, LMAX_AUX>::run(t); }

}  // namespace dfhip
