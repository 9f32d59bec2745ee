// 4c2e diagonal (ab|ab) class kernels for the Schwarz norms.
#include "schwarz4c_impl.h"

namespace dfhip {

namespace {
template <int LA, int LB>
struct Reg {
  static void run(Schwarz4Table& t) {
    t[LA][LB] = make_schwarz4<LA, LB>();
    if constexpr (LB > 0) Reg<LA, LB - 1>::run(t);
    else if constexpr (LA > 0) Reg<LA - 1, LA - 1>::run(t);
  }
};
}  // namespace

template <int LA, int LB>
struct RegC {
  static void run(Schwarz4CachedTable& t) {
    t[LA][LB] = make_schwarz4_cached<LA, LB>();
    if constexpr (LB > 0) RegC<LA, LB - 1>::run(t);
    else if constexpr (LA > 0) RegC<LA - 1, LA - 1>::run(t);
  }
};

void register_schwarz4c(Schwarz4Table& t) { Reg<LMAX_ORB, LMAX_ORB>::run(t); }
void register_schwarz4c_cached(Schwarz4CachedTable& t) { RegC<LMAX_ORB, LMAX_ORB>::run(t); }

}  // namespace dfhip
