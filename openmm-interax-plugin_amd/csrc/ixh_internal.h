// Hooks between the library's modules that are not part of the C ABI (used by the batch executor's tuner).
#ifndef IXH_INTERNAL_H
#define IXH_INTERNAL_H

#include "interax_hip.h"

namespace ixh {

// Largest number of workgroups of a pair tile launch (each wave then takes a contiguous range of tiles).
void pair_set_grid_limit(ixh_pair* pair, int workgroups);
int pair_num_tiles(const ixh_pair* pair);

}  // namespace ixh

#endif
