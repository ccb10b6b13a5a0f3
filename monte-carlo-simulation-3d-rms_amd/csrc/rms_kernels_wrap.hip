// rms_kernels_wrap.hip -- the walk kernel variants for plans with a geometry
// whose cell words name the voxels across its periodic sides (SIG bits 3
// and 6): the forward walk and the tangent passes without the boundary
// terms, with dense or brick-stored words (those with the boundary terms
// are in rms_kernels_jump.hip).  Their fast path takes the steps through a
// periodic side; the other periodic variants hand those steps to the march.
// A translation unit of its own so that the build compiles it beside the
// others.
#include "rms_walk.h"

namespace rms {

// Index bits 0-2: variant bits 0-2 (narrow-pulse, PGSE, exchange faces);
// index bit 3: variant bit 5 (brick-stored words).
template <typename R, int NP, int... S>
static const void *wrap_fn_sig(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, 72 | (S & 7) | ((S & 8) << 2), NP>...};
    return fns[(sig & 7) | ((sig >> 2) & 8)];
}

template <int NP>
static const void *wrap_fn_prec(int prec_f64, int sig)
{
    if (prec_f64) return wrap_fn_sig<double, NP>(sig, std::make_integer_sequence<int, 16>{});
    return wrap_fn_sig<float, NP>(sig, std::make_integer_sequence<int, 16>{});
}

const void *wrap_walk_fn(int prec_f64, int sig, int np)
{
    if ((sig & 16) || !(sig & 8) || !(sig & 64)) return nullptr;  // the host never asks for these here
    switch (np) {
    case 0: return wrap_fn_prec<0>(prec_f64, sig);
    case 1: return wrap_fn_prec<1>(prec_f64, sig);
    case 2: return wrap_fn_prec<2>(prec_f64, sig);
    default: return wrap_fn_prec<3>(prec_f64, sig);
    }
}

}  // namespace rms
