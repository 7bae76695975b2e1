// tables.h -- host-side slider table construction.
#pragma once
#include "sliders.h"

namespace pf {

u64 rook_mask(int sq);
u64 bishop_mask(int sq);
// Fills every slider table; 0 on success, negative if a table fails its
// Kogge-Stone check or the published magic set does not fit.
int build_slider_tables(SliderTables& t);

}  // namespace pf
