// rms_kernels_tangent.hip -- the tangent variants of the walk kernel for
// plans without a PGSE table and without the boundary terms (SIG bits 0, 2
// and 3 only): the pathwise D, analytic T2 and score-function kappa pass that
// the default adjoint runs.  A translation unit of its own so that the build
// compiles it beside the others.
#include "rms_walk.h"

namespace rms {

template <typename R, int NP>
static const void *tangent_plain_fn(int sig)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, 0, NP>,  (const void *)walk_kernel<R, 1, NP>,
                                      (const void *)walk_kernel<R, 4, NP>,  (const void *)walk_kernel<R, 5, NP>,
                                      (const void *)walk_kernel<R, 8, NP>,  (const void *)walk_kernel<R, 9, NP>,
                                      (const void *)walk_kernel<R, 12, NP>, (const void *)walk_kernel<R, 13, NP>};
    return (sig & 18) ? nullptr : fns[(sig & 1) | ((sig >> 1) & 6)];
}

template <int NP>
static const void *tangent_plain_prec(int prec_f64, int sig)
{
    return prec_f64 ? tangent_plain_fn<double, NP>(sig) : tangent_plain_fn<float, NP>(sig);
}

const void *tangent_walk_fn(int prec_f64, int sig, int np)
{
    switch (np) {
    case 1: return tangent_plain_prec<1>(prec_f64, sig);
    case 2: return tangent_plain_prec<2>(prec_f64, sig);
    default: return tangent_plain_prec<3>(prec_f64, sig);
    }
}

}  // namespace rms
