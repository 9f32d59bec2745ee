// Matrix-core 3c2e class kernels with la = 0 (all lb <= la, lc <= LMAX_AUX) and their Et tables.
#include "int3c2e_mm_impl.h"

namespace dfhip {

namespace {
template <int LB, int LC>
struct RegMM {
  static void run(Int3MTable& t, EtTable& e) {
    t[0][LB][LC] = make_int3m<0, LB, LC>();
    if constexpr (LC == 0) e[0][LB] = make_et<0, LB>();
    if constexpr (LC > 0) RegMM<LB, LC - 1>::run(t, e);
    else if constexpr (LB > 0) RegMM<LB - 1, LMAX_AUX>::run(t, e);
  }
};
}  // namespace

void register_int3c2e_mm_l0(Int3MTable& t, EtTable& e) { RegMM<0, LMAX_AUX>::run(t, e); }

}  // namespace dfhip
