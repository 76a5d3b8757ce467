// libacehip.so, force kernels for Lmax 9 (dispatched from ace_force.hip).
#include "ace_force_impl.hpp"

template int launch_force_l<9>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
