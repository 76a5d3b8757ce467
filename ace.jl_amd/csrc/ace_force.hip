// libacehip.so, forces for Lmax 0-4, and the dispatch over Lmax (the other
// values are instantiated in ace_force_l5.hip ... ace_force_l10.hip).
#include "ace_force_impl.hpp"

template int launch_force_l<0>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
template int launch_force_l<1>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
template int launch_force_l<2>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
template int launch_force_l<3>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
template int launch_force_l<4>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*, int64_t,
                                 const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<5>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                        int64_t, const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<6>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                        int64_t, const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<7>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                        int64_t, const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<8>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                        int64_t, const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<9>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                        int64_t, const int8_t*, int64_t, const Workspace&, double*);
extern template int launch_force_l<10>(ace_ctx*, hipStream_t, int64_t, const int32_t*, const double*,
                                         int64_t, const int8_t*, int64_t, const Workspace&, double*);

int launch_force_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                     int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws, double* dE) {
    // the flat kernel reads the neighbour -> environment map of the pooling launch
    if (ctx->prop == ACE_PROP_INVARIANT && ctx->flat_force && (!ws.nbr_env || ws.nbr_cap < (size_t)nnbr))
        return set_err("force launch without the neighbour map of its pooling launch");
    if (!ws.Gbuf) return set_err("force launch without the adjoint buffer");
    switch (ctx->L) {
        case 0: return launch_force_l<0>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 1: return launch_force_l<1>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 2: return launch_force_l<2>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 3: return launch_force_l<3>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 4: return launch_force_l<4>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 5: return launch_force_l<5>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 6: return launch_force_l<6>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 7: return launch_force_l<7>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 8: return launch_force_l<8>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        case 9: return launch_force_l<9>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
        default: return launch_force_l<10>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
    }
}
