// rms_kernels_bricks.hip -- the walk kernel variants for plans with a
// geometry whose cell words are stored in bricks (SIG bit 5): the forward
// walk and the tangent passes without the boundary terms (those with them
// are in rms_kernels_jump.hip).  Plans without such a geometry run the
// variants of rms_kernels.hip and rms_kernels_tangent.hip, whose lookups
// read dense words only.  A translation unit of its own so that the build
// compiles it beside the others.
#include "rms_walk.h"

namespace rms {

template <typename R, int NP, int... S>
static const void *brick_fn_sig(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, 32 | S, NP>...};
    return fns[sig & 15];
}

template <int NP>
static const void *brick_fn_prec(int prec_f64, int sig)
{
    if (prec_f64) return brick_fn_sig<double, NP>(sig, std::make_integer_sequence<int, 16>{});
    return brick_fn_sig<float, NP>(sig, std::make_integer_sequence<int, 16>{});
}

const void *brick_walk_fn(int prec_f64, int sig, int np)
{
    if (sig & 16) return nullptr;  // the boundary terms: rms_kernels_jump.hip
    switch (np) {
    case 0: return brick_fn_prec<0>(prec_f64, sig);
    case 1: return brick_fn_prec<1>(prec_f64, sig);
    case 2: return brick_fn_prec<2>(prec_f64, sig);
    default: return brick_fn_prec<3>(prec_f64, sig);
    }
}

}  // namespace rms
