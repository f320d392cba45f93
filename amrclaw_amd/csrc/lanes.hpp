// lanes.hpp -- neighbour exchange between the lanes of a wave.
#pragma once
#include <hip/hip_runtime.h>

namespace amrgpu {

// Neighbour exchange between the lanes of a segment (segments of 16, 32 or 64
// lanes tile the 64-lane wave): lane_down gives lane i the value of lane i+1,
// lane_up that of lane i-1.  Both are one DPP move across the whole wave
// (wave_shl:1 / wave_shr:1) instead of an LDS permute; at a segment's last
// (first) lane they return the next (previous) segment's value, which the
// sweep never uses: the limiter and the transverse terms read a neighbour
// only from interfaces 1..n+1 of an n-cell slice, whose neighbours 0..n+2 sit
// in the same segment (segment width >= n + 3).
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<CTRL>((int)(b & 0xffffffffll)), hi = dpp_move<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ T lane_down(T v, int) { return dpp_move<DPP_WAVE_SHL1>(v); }
template <typename T>
__device__ __forceinline__ T lane_up(T v, int) { return dpp_move<DPP_WAVE_SHR1>(v); }

}  // namespace amrgpu
