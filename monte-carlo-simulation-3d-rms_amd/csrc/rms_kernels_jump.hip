// rms_kernels_jump.hip -- the tangent variants of the walk kernel with the
// boundary terms of the D derivative (SIG bit 4).  A translation unit of
// their own so that the build can give them their own register
// allocation (see build.py).
#include "rms_walk.h"

namespace rms {

// Adjoint variants with the jump term (bit 4): exchange faces (bit 2), no
// PGSE table, with or without the b-table (bit 0), periodic edges (bit 3),
// brick-stored cell words (bit 5) and wrapped cell words (bit 6, only with
// bit 3).
template <typename R, int NP>
static const void *jump_fn(int sig)
{
    static const void *const fns[] = {
        (const void *)walk_kernel<R, 20, NP>, (const void *)walk_kernel<R, 21, NP>,
        (const void *)walk_kernel<R, 28, NP>, (const void *)walk_kernel<R, 29, NP>,
        (const void *)walk_kernel<R, 52, NP>, (const void *)walk_kernel<R, 53, NP>,
        (const void *)walk_kernel<R, 60, NP>, (const void *)walk_kernel<R, 61, NP>,
        nullptr, nullptr, (const void *)walk_kernel<R, 92, NP>, (const void *)walk_kernel<R, 93, NP>,
        nullptr, nullptr, (const void *)walk_kernel<R, 124, NP>, (const void *)walk_kernel<R, 125, NP>};
    return fns[(sig & 1) | ((sig >> 2) & 2) | ((sig >> 3) & 4) | ((sig >> 3) & 8)];
}

template <int NP>
static const void *jump_fn_prec(int prec_f64, int sig)
{
    return prec_f64 ? jump_fn<double, NP>(sig) : jump_fn<float, NP>(sig);
}

const void *jump_walk_fn(int prec_f64, int sig, int np)
{
    switch (np) {
    case 1: return jump_fn_prec<1>(prec_f64, sig);
    case 2: return jump_fn_prec<2>(prec_f64, sig);
    default: return jump_fn_prec<3>(prec_f64, sig);
    }
}

}  // namespace rms
