// vs_fprep.hpp — the exact result of k repetitions of an FP32 addition
// a <- fl(a + d), in a few operations per binade instead of k additions.
//
// The raycast's march repeats the same step through regions that hold no
// data (position += S*rd, total += S), and the reference performs every one
// of those additions. A skipped run must end on the same FP32 values, so the
// run is evaluated exactly:
//
// Inside one binade [2^e, 2^(e+1)) of magnitudes the FP32 values are the
// multiples of u = 2^(e-23), and when the exact sum a + d stays in the binade,
// fl(a + d) = a + r with r = d rounded to a multiple of u. r does not depend
// on a, except when d lies exactly halfway between two multiples (then the
// result is the even multiple); after one such step inside the binade every
// value is even and r is the same for every later step. So once three
// consecutive values a1 -> a2 -> a3 lie in one binade, r = a3 - a2 (exact,
// Sterbenz) is the increment of every later step inside it, and
// a(3+m) = a3 + m*r exactly (m*r and the sum are multiples of u below
// 2^(e+1)), as long as every exact sum a(j) + d stays in the binade. Since
// |d| <= |r| + u/2, that holds when |a3| + m|r| < 2^(e+1) for a growing
// magnitude and |a3| - m|r| > 2^e for a shrinking one. A step that leaves
// the binade, and values near zero (|a| < 2^-63), are done one by one.
// r == 0 (a3 == a2) is a fixed point: every later step returns a3.
//
// Plain arithmetic, so the same code runs on the host, where it is checked
// against the step-by-step loop.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define VS_HD __host__ __device__ __forceinline__
#else
#define VS_HD inline
#endif

namespace vs {

constexpr int kRoomMax = 1 << 16;  // cap of one closed-form jump

VS_HD uint32_t f32_bits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}
VS_HD float f32_from(uint32_t u) {
    float x;
    memcpy(&x, &u, 4);
    return x;
}
// approximate 1/x (x > 0): the hardware reciprocal on the device (1 ulp),
// IEEE division on the host; callers allow for a relative error of 2^-22.
// VS_FPREP_RCP_SCALE lets a host check run with a deliberately high estimate.
#ifndef VS_FPREP_RCP_SCALE
#define VS_FPREP_RCP_SCALE 1.0f
#endif
VS_HD float approx_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return VS_FPREP_RCP_SCALE / x;
#endif
}

// How many further steps a3 -> a3 + m*r are exact, given three consecutive
// values of the run (0 when they are not in one binade).
VS_HD int binade_room(float a1, float a2, float a3) {
    const float r = a3 - a2;
    if (r == 0.0f) return kRoomMax;
    const uint32_t b1 = f32_bits(a1) >> 23, b2 = f32_bits(a2) >> 23, b3 = f32_bits(a3) >> 23;
    const uint32_t e = b3 & 0xFFu;  // biased exponent
    if (b1 != b3 || b2 != b3 || e < 64u || e > 253u) return 0;
    const float lo = f32_from(e << 23), mag = a3 < 0.0f ? -a3 : a3, ar = r < 0.0f ? -r : r;
    // distance to the edge the run moves towards (exact: multiples of u in [0, 2^e])
    const float gap = ((r > 0.0f) == (a3 > 0.0f)) ? 2.0f * lo - mag : mag - lo;
    float q = gap * approx_rcp(ar);
    q = q < (float)kRoomMax ? q : (float)kRoomMax;
    // floor(q) - 1 < gap/|r| for q <= 2^16 under a 2^-22 relative error of q
    const int m = (int)q - 1;
    return m > 0 ? m : 0;
}

// a after k >= 0 repetitions of a = fl(a + d).
VS_HD float repeat_add(float a, float d, int k) {
    while (k > 0) {
        const float a1 = a + d;
        if (k == 1) return a1;
        const float a2 = a1 + d;
        if (k == 2) return a2;
        const float a3 = a2 + d;
        int m = binade_room(a1, a2, a3);
        m = m < k - 3 ? m : k - 3;
        a = m > 0 ? a3 + (float)m * (a3 - a2) : a3;
        k -= 3 + m;
    }
    return a;
}

}  // namespace vs
