// rms_kernels_nocount.hip -- the forward walk variants without the fast
// path's diagnostic counters (SIG bit 7), for launches that do not report
// them (rms_run_params.skip_counters, the central differences' walks): the
// per-step sums of crossings, reflections, mirrors and draws are about 3 %
// of the step.  Dense, brick-stored and periodically wrapped cell words,
// each also with bit 8 (every T2 infinite and no exchange draw in the
// launch: no time-in-class sums, no draw bits, weight 1).  A
// translation unit of its own so that the build compiles it beside the
// others.
#include "rms_walk.h"

namespace rms {

// dense: variant bits 0-3 (narrow-pulse, PGSE, exchange faces, periodic edges)
template <typename R, int L, int... S>
static const void *nocount_dense(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, L | 128 | S, 0>...};
    return fns[sig & 15];
}

// brick-stored words: bit 5 with bits 0-3
template <typename R, int L, int... S>
static const void *nocount_bricks(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, L | 128 | 32 | S, 0>...};
    return fns[sig & 15];
}

// wrapped words (bits 3 and 6): index bits 0-2 = variant bits 0-2, index
// bit 3 = variant bit 5
template <typename R, int L, int... S>
static const void *nocount_wrap(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, L | 128 | 72 | (S & 7) | ((S & 8) << 2), 0>...};
    return fns[(sig & 7) | ((sig >> 2) & 8)];
}

template <typename R, int L>
static const void *nocount_fn(int sig)
{
    if (sig & 16) return nullptr;  // boundary terms: tangent passes only
    if (sig & 64) return (sig & 8) ? nocount_wrap<R, L>(sig, std::make_integer_sequence<int, 16>{}) : nullptr;
    if (sig & 32) return nocount_bricks<R, L>(sig, std::make_integer_sequence<int, 16>{});
    return nocount_dense<R, L>(sig, std::make_integer_sequence<int, 16>{});
}

const void *nocount_walk_fn(int prec_f64, int sig)
{
    if (sig & 256) return prec_f64 ? nocount_fn<double, 256>(sig) : nocount_fn<float, 256>(sig);
    return prec_f64 ? nocount_fn<double, 0>(sig) : nocount_fn<float, 0>(sig);
}

}  // namespace rms
