// search_inst.hip -- one (dims, object kind, layout family) of the range-search
// kernels: compiled once per combination with -DQCP_D=2|3 -DQCP_BOX=0|1
// -DQCP_EXPL=0|1 (the Makefile's sk_<D><p|b><0|1>.o objects).
#include "search_kernels.h"

#if !defined(QCP_D) || !defined(QCP_BOX) || !defined(QCP_EXPL)
#error "search_inst.hip needs QCP_D, QCP_BOX and QCP_EXPL"
#endif

namespace qcp {

template hipError_t launch_search_dbe<QCP_D, (bool)QCP_BOX, (bool)QCP_EXPL>(
    const SearchArgs&, int, int, unsigned, unsigned, hipStream_t);
template const void* search_ptr_dbe<QCP_D, (bool)QCP_BOX, (bool)QCP_EXPL>(int, int);
#if QCP_EXPL
template hipError_t launch_mprs_part_db<QCP_D, (bool)QCP_BOX>(const SearchArgs&, unsigned,
                                                              unsigned, hipStream_t);
#else
template hipError_t launch_mprs_db<QCP_D, (bool)QCP_BOX>(const SearchArgs&, unsigned, unsigned,
                                                         hipStream_t);
#endif

}  // namespace qcp
