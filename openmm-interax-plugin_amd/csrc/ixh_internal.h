// Hooks between the library's modules that are not part of the C ABI (used by the batch executor's tuner).
#ifndef IXH_INTERNAL_H
#define IXH_INTERNAL_H

#include "interax_hip.h"

#include <hip/hip_runtime.h>

namespace ixh {

// Largest number of workgroups of a pair tile launch (each wave then takes a contiguous range of tiles).
void pair_set_grid_limit(ixh_pair* pair, int workgroups);
int pair_num_tiles(const ixh_pair* pair);
// ixh_pair_lambda_energies in two halves, for callers that cannot wait (the server): the launch, asynchronous on the
// buffers' stream (flags may hold IXH_DEFER_LIST_CHECK for the handle's own evaluation), and, once that stream has
// passed it, the energies at the count lambda values (without the long range correction).
void pair_lambda_launch(ixh_pair* pair, const ixh_buffers* buffers, const double* lambdas, int count, int flags);
void pair_lambda_collect(const ixh_pair* pair, int count, double* energies);
// Launch shape of a segmented three-body handle's triple kernel: wavefronts per work item (1, 2, 4, 8) and neighbour
// pairs per work item (a multiple of 64).  Waits for the device when the work items change.
void threebody_set_launch_shape(ixh_threebody* threebody, int waves, int chunkPairs);
// An event a segmented three-body handle records on its stream after each evaluation's neighbour search (NULL: none);
// the caller owns it.
void threebody_set_search_event(ixh_threebody* threebody, hipEvent_t event);
void threebody_launch_shape(const ixh_threebody* threebody, int* waves, int* chunkPairs);

}  // namespace ixh

#endif
