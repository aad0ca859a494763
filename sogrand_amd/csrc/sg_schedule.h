// Error-pattern schedule tables (host side).
#pragma once

#include <stdint.h>

#include "sg_common.h"

namespace sg {

// Fill out[0..kSchedStride) with the schedule for intercept `ic` (ic >= 55 gives
// the generic order): 65535 nonzero masks then a 0 end marker.
void build_schedule_table(int ic, uint16_t *out);

// All kSchedTables tables back to back (table t at out + t * kSchedStride).
void build_all_schedule_tables(uint16_t *out);

} // namespace sg
