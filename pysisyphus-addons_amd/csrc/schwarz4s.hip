// Streamed matrix-core Schwarz class kernels (tables in global scratch).
#include "schwarz4s_impl.h"

namespace dfhip {

namespace {
template <int LA, int LB>
struct RegS {
  static void run(Schwarz4STable& t) {
    t[LA][LB] = make_schwarz4s<LA, LB>();
    if constexpr (LB > 0) RegS<LA, LB - 1>::run(t);
    else if constexpr (LA > 0) RegS<LA - 1, LA - 1>::run(t);
  }
};
}  // namespace

void register_schwarz4s(Schwarz4STable& t) { RegS<LMAX_ORB, LMAX_ORB>::run(t); }

}  // namespace dfhip
