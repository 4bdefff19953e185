// update_interp.h -- the stride-grid interpolation of T and w at one model
// cell, shared by the field update (update_mfma.hip) and the analysed
// observations (obs_update.hip).
#pragma once

#include "enkfg_internal.h"

namespace enkfg {

// one interpolated value of node data `base` (records of `rec` floats),
// update.c:224-271: j first, then i, FP32, no contraction
struct Bilin {
    long long o00, o10, o01, o11;
    bool jcopy, icopy;
    float wj1, wj2, wi1, wi2;
};

__device__ inline Bilin make_bilin(const UpdArgs& a, int j, int i, long long rec)
{
    Bilin b;
    const int jj = j / a.stride, stepj = j % a.stride;
    const int ii = i / a.stride, stepi = i % a.stride;
    const int jj2 = (jj < a.nj_s - 1) ? jj + 1 : jj;
    const int ii2 = (ii < a.ni_s - 1) ? ii + 1 : (a.periodic_i ? (ii + 1) % a.ni_s : ii);
    b.jcopy = (a.stride == 1 || stepj == 0);
    b.icopy = (a.stride == 1 || stepi == 0);
    b.wj2 = (float)stepj / (float)a.stride;
    b.wj1 = 1.0f - b.wj2;
    b.wi2 = (float)stepi / (float)a.stride;
    b.wi1 = 1.0f - b.wi2;
    // all four corners are valid nodes; the unused ones repeat (jj, ii)
    b.o00 = ((long long)jj * a.ni_s + ii) * rec;
    b.o10 = ((long long)(b.jcopy ? jj : jj2) * a.ni_s + ii) * rec;
    b.o01 = ((long long)jj * a.ni_s + (b.icopy ? ii : ii2)) * rec;
    b.o11 = ((long long)(b.jcopy ? jj : jj2) * a.ni_s + (b.icopy ? ii : ii2)) * rec;
    return b;
}

__device__ inline float bilin(const Bilin& b, float v00, float v10, float v01, float v11)
{
    const float ta = b.jcopy ? v00 : v00 * b.wj1 + v10 * b.wj2;
    const float tb = b.jcopy ? v01 : v01 * b.wj1 + v11 * b.wj2;
    return b.icopy ? ta : ta * b.wi1 + tb * b.wi2;
}

}  // namespace enkfg
