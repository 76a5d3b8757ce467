// libacehip.so, force kernels for Lmax 10 (dispatched from ace_force.hip).
#include "ace_force_impl.hpp"

template int launch_force_l<10>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
