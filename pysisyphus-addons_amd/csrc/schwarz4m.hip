// Matrix-core Schwarz class kernels (one per (la, lb) whose tables fit in LDS).
#include "schwarz4m_impl.h"

namespace dfhip {

namespace {
template <int LA, int LB>
struct RegM {
  static void run(Schwarz4MTable& t) {
    t[LA][LB] = make_schwarz4m<LA, LB>();
    if constexpr (LB > 0) RegM<LA, LB - 1>::run(t);
    else if constexpr (LA > 0) RegM<LA - 1, LA - 1>::run(t);
  }
};
}  // namespace

void register_schwarz4m(Schwarz4MTable& t) { RegM<LMAX_ORB, LMAX_ORB>::run(t); }

}  // namespace dfhip
