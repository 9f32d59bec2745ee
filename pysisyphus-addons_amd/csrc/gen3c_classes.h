// Generated by gen3c.py: registration of the register-resident 3c2e kernels.
#pragma once
#include "gen3c_common.h"

namespace dfhip {
static_assert(GEN3_LA_MAX == 2 && GEN3_LC_MAX == 3, "regenerate");
void register_gen3_00(Gen3Table& t);
void register_gen3_10(Gen3Table& t);
void register_gen3_11(Gen3Table& t);
void register_gen3_20(Gen3Table& t);
void register_gen3_21(Gen3Table& t);
void register_gen3_22(Gen3Table& t);
inline void register_gen3_all(Gen3Table& t) {
  register_gen3_00(t);
  register_gen3_10(t);
  register_gen3_11(t);
  register_gen3_20(t);
  register_gen3_21(t);
  register_gen3_22(t);
}
}  // namespace dfhip
