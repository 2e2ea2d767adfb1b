/* Companion of shim_check.cpp compiled without the reference headers. */
#include <cstring>
#include "radchan.h"
extern "C" int shim_check_values(int32_t *out) {
    rc_params q;
    rc_get_params(&q);
    const int32_t v[18] = {q.imax, q.jmax, q.kmax, q.p_col, q.photmax, q.grid_num, q.maxi[5], q.maxj[8],
                           q.maxk[8], q.maxs[8], q.nB, q.nT, q.nQ, q.bd[4], q.srt, (int32_t)(q.toll * 1e6f),
                           q.sampler, q.fp64};
    std::memcpy(out, v, sizeof v);
    return 0;
}
