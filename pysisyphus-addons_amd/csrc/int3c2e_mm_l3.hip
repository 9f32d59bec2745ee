// Matrix-core 3c2e class kernels with la = 3 (all lb <= la, lc <= LMAX_AUX) and their Et tables.
#include "int3c2e_mm_impl.h"

namespace dfhip {

namespace {
template <int LB, int LC>
struct RegMM {
  static void run(Int3MTable& t, EtTable& e) {
    t[3][LB][LC] = make_int3m<3, LB, LC>();
    if constexpr (LC == 0) e[3][LB] = make_et<3, LB>();
    if constexpr (LC > 0) RegMM<LB, LC - 1>::run(t, e);
    else if constexpr (LB > 0) RegMM<LB - 1, LMAX_AUX>::run(t, e);
  }
};
}  // namespace

void register_int3c2e_mm_l3(Int3MTable& t, EtTable& e) { RegMM<3, LMAX_AUX>::run(t, e); }

}  // namespace dfhip
