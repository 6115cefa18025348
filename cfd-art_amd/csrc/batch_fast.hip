// The batch's sweep kernels in the fast arithmetic (F = true). FMA contraction is allowed here:
// the fast arithmetic is held to the 1e-10 accuracy bar, not to the reference's bits, and the
// single-grid fast kernels (sweep_kernels_fast.hip) are compiled the same way, so a grid in a
// fast batch and the same grid alone still end on the same bits.
#pragma clang fp contract(fast)

#include "batch_kernels.h"

namespace cfd {

void launch_batch_phases_fast(int model, BatchGrid* grids, Tile* const* tiles, const size_t* cap, const BatchQueue* q,
                              hipStream_t st, int shape)
{
    if (model == EULER) launch_batch_phases<EULER, true>(grids, tiles, cap, q, st, shape);
    else launch_batch_phases<ALLAIRE, true>(grids, tiles, cap, q, st, shape);
}

}  // namespace cfd
