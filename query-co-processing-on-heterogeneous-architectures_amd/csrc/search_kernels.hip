// search_kernels.hip -- dispatch of the range-search kernels (the kernel
// templates live in search_kernels.h and are instantiated per (dims, object
// kind, layout family) by search_inst.hip), plus the queue reset kernel.
#include <hip/hip_runtime.h>

#include "search_kernels.h"

namespace qcp {

__global__ void k_reset_ctl(QueueCtl* ctl, uint32_t pending) {
  const uint32_t t = threadIdx.x;
  if (t < 8) {
    ctl->xhead[t] = 0;
    ctl->internal_entries[t] = 0;
    ctl->leaf_entries[t] = 0;
    ctl->total_hits[t] = 0;
    ctl->contained[t] = 0;
  }
  if (t == 0) {
    ctl->head = 0;
    ctl->tail = 0;
    ctl->pending = pending;
    ctl->done = pending == 0 ? 1u : 0u;
    ctl->donated = 0;
  }
}

size_t search_lds_bytes(bool expl, int mw, int out, unsigned block) {
  const size_t ids = out == kOutIds ? (size_t)(block / kWave) * 2 * kWave * sizeof(uint32_t) : 0;
  switch (mw) {
    case 1: return ids + (expl ? lds_for<1, true, 0>(block) : lds_for<1, false, 0>(block));
    case 2: return ids + (expl ? lds_for<2, true, 0>(block) : lds_for<2, false, 0>(block));
    case 3: return ids + (expl ? lds_for<3, true, 0>(block) : lds_for<3, false, 0>(block));
    default: return ids + (expl ? lds_for<4, true, 0>(block) : lds_for<4, false, 0>(block));
  }
}

template <int D, bool BOX>
static hipError_t launch_search_d(const SearchArgs& a, bool expl, int out, int u, unsigned grid,
                                  unsigned block, hipStream_t s) {
  return expl ? launch_search_dbe<D, BOX, true>(a, out, u, grid, block, s)
              : launch_search_dbe<D, BOX, false>(a, out, u, grid, block, s);
}

hipError_t launch_search(const SearchArgs& a, unsigned D, bool box, bool expl, int out, int u,
                         unsigned grid, unsigned block, hipStream_t s) {
  k_reset_ctl<<<1, 64, 0, s>>>(a.ctl, a.items ? a.n_items : a.nq);
  if (D == 2) return box ? launch_search_d<2, true>(a, expl, out, u, grid, block, s)
                         : launch_search_d<2, false>(a, expl, out, u, grid, block, s);
  if (D == 3) return box ? launch_search_d<3, true>(a, expl, out, u, grid, block, s)
                         : launch_search_d<3, false>(a, expl, out, u, grid, block, s);
  return hipErrorInvalidValue;
}

hipError_t launch_mprs(const SearchArgs& a, unsigned D, bool box, unsigned nblocks,
                       unsigned waves, hipStream_t s) {
  if (D == 2)
    return box ? launch_mprs_db<2, true>(a, nblocks, waves, s)
               : launch_mprs_db<2, false>(a, nblocks, waves, s);
  if (D == 3)
    return box ? launch_mprs_db<3, true>(a, nblocks, waves, s)
               : launch_mprs_db<3, false>(a, nblocks, waves, s);
  return hipErrorInvalidValue;
}

hipError_t launch_mprs_part(const SearchArgs& a, unsigned D, bool box, unsigned parts,
                            unsigned waves, hipStream_t s) {
  if (D == 2)
    return box ? launch_mprs_part_db<2, true>(a, parts, waves, s)
               : launch_mprs_part_db<2, false>(a, parts, waves, s);
  if (D == 3)
    return box ? launch_mprs_part_db<3, true>(a, parts, waves, s)
               : launch_mprs_part_db<3, false>(a, parts, waves, s);
  return hipErrorInvalidValue;
}

const void* search_kernel_ptr(unsigned D, bool box, bool expl, int out, int mw) {
  if (D == 2) {
    if (box) return expl ? search_ptr_dbe<2, true, true>(out, mw) : search_ptr_dbe<2, true, false>(out, mw);
    return expl ? search_ptr_dbe<2, false, true>(out, mw) : search_ptr_dbe<2, false, false>(out, mw);
  }
  if (D == 3) {
    if (box) return expl ? search_ptr_dbe<3, true, true>(out, mw) : search_ptr_dbe<3, true, false>(out, mw);
    return expl ? search_ptr_dbe<3, false, true>(out, mw) : search_ptr_dbe<3, false, false>(out, mw);
  }
  return nullptr;
}

}  // namespace qcp
