// Helpers of the row-form acoustic kernels (wave_row.hip, wave_row2.hip): one
// wave holds a z row, four cells per lane.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd {

namespace row {

template <class T> struct Row4;
template <> struct Row4<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct Row4<double> {
    typedef double type __attribute__((ext_vector_type(4)));
};

// DPP wave shifts (CDNA): wave_shr:1 gives lane i the value of lane i-1,
// wave_shl:1 the value of lane i+1; the lane whose source is outside the wave
// keeps `own` (bound_ctrl off).
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL> __device__ inline float lane_shift(float own, float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, own), __builtin_bit_cast(int, v), CTRL, 0xf, 0xf,
                                           false));
}
template <int CTRL> __device__ inline double lane_shift(double own, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, own);
    const unsigned long long s = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(int(unsigned(o)), int(unsigned(s)), CTRL, 0xf, 0xf, false));
    const unsigned hi =
        unsigned(__builtin_amdgcn_update_dpp(int(unsigned(o >> 32)), int(unsigned(s >> 32)), CTRL, 0xf, 0xf, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

template <class V4> __device__ inline void put(V4* dst, const V4& v, bool nt) {
    if (nt) __builtin_nontemporal_store(v, dst);
    else *dst = v;
}

}  // namespace row

}  // namespace cfd
