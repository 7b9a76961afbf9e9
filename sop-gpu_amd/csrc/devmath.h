// devmath.h -- device arithmetic shared by the SOP kernels.
//
// Hybrid Taus generator (src/Integrators/ht.cu:61-106 in the reference):
// three Tausworthe components plus one LCG, XOR-combined, mapped to (0,1]
// through the exponent-bit trick, then Box-Muller.  The reference takes its
// transcendental functions from the CUDA fast-math library, whose values are
// not pinned by anything in the reference; this build uses its own log and
// sincos (below), so the GPU values and the oracle's are the same bits.
//
// All of this compiles with -ffp-contract=off and correctly rounded
// division and square root, which is what lets every float expression here
// round exactly as written.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sop {

__device__ __forceinline__ uint32_t taus_step(uint32_t& z, int S1, int S2, int S3, uint32_t M) {
    uint32_t b = (((z << S1) ^ z) >> S2);
    z = (((z & M) << S3) ^ b);
    return z;
}

// One draw; advances all four state words.
__device__ __forceinline__ uint32_t hybrid_taus(uint4& s) {
    uint32_t a = taus_step(s.x, 13, 19, 12, 4294967294u);
    uint32_t b = taus_step(s.y, 2, 25, 4, 4294967288u);
    uint32_t c = taus_step(s.z, 3, 11, 17, 4294967280u);
    s.w = 1664525u * s.w + 1013904223u;
    return a ^ b ^ c ^ s.w;
}

// a / b correctly rounded, in fewer operations than the general IEEE
// division, for operands kept away from overflow and underflow: b in
// [2^-64, 2^62] and a = 0 or a in [2^-60, 2^60] (kDivFast* below).  The
// reciprocal y = RN(1/b) comes from the hardware estimate and one Newton
// step; q = RN(a y) is then within one ulp of a/b, the residual a - b q is
// exact in an FMA, and RN(q + (a - b q) y) = RN(a/b) by Markstein's theorem
// (P. Markstein, IBM J. Res. Develop. 34 (1990) 111-119).  That the Newton
// step yields RN(1/b) for every normal b is a property of the hardware
// estimate, checked over all of them by sop_check_division.
constexpr float kDivFastMinB = 0x1p-64f, kDivFastMaxB = 0x1p62f;
constexpr float kDivFastMinA = 0x1p-60f, kDivFastMaxA = 0x1p60f;
__device__ __forceinline__ float div_rn_fast(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float y = __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
    const float q = a * y;
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
}
// The IEEE division, kept on its own rare branch: the divisor passes
// through an empty asm, so the compiler cannot compute the quotient ahead of
// the branch for every lane.
__device__ __forceinline__ float div_ieee(float a, float b) {
    asm volatile("" : "+v"(b));
    return a / b;
}
// a can take the fast division (a wave-uniform test, made once per bead).
__device__ __forceinline__ bool div_fast_numerator(float a) {
    return a == 0.0f || (a >= kDivFastMinA && a <= kDivFastMaxA);
}

// Mantissa bits into [1,2), minus one; an exact zero becomes 1e-8.
__device__ __forceinline__ float uint_to_unit(uint32_t u) {
    float f = __uint_as_float(0x3f800000u | (0x007fffffu & u)) - 1.0f;
    return f == 0.0f ? 1.0e-8f : f;
}

// log(x) for positive normal x: x = 2^e m, m folded into [sqrt(1/2), sqrt(2)],
// log(m) = 2 atanh(s), s = (m-1)/(m+1), series to s^9; e*ln2 with ln2 split
// into a short leading part and a correction (Cody & Waite).
__device__ __forceinline__ float logf_series(float x) {
    uint32_t bits = __float_as_uint(x);
    int e = (int)((bits >> 23) & 0xffu) - 127;
    float m = __uint_as_float((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float p = 0.222222222f;
    p = p * z + 0.285714298f;
    p = p * z + 0.400000006f;
    p = p * z + 0.666666687f;
    p = p * z + 2.0f;
    p = p * s;
    float fe = (float)e;
    return fe * 0.693145751953125f + (p + fe * 1.42860677e-06f);
}

// sin and cos: reduction by q = rint(x*2/pi) with pi/2 in three parts
// (Cody & Waite), then Taylor polynomials on [-pi/4, pi/4].
__device__ __forceinline__ void sincosf_series(float x, float& sp, float& cp) {
    float qf = rintf(x * 0.636619772f);
    int q = (int)qf;
    float r = x - qf * 1.5703125f;
    r = r - qf * 4.83751297e-04f;
    r = r - qf * 7.54978995e-08f;
    float z = r * r;
    float ps = 2.75573192e-06f;
    ps = ps * z - 1.98412698e-04f;
    ps = ps * z + 8.33333377e-03f;
    ps = ps * z - 1.66666672e-01f;
    float sn = r + r * (z * ps);
    float pc = -2.75573188e-07f;
    pc = pc * z + 2.48015876e-05f;
    pc = pc * z - 1.38888892e-03f;
    pc = pc * z + 4.16666679e-02f;
    float cs = (1.0f - 0.5f * z) + (z * z) * pc;
    switch (q & 3) {
    case 0: sp = sn; cp = cs; break;
    case 1: sp = cs; cp = -sn; break;
    case 2: sp = -sn; cp = -cs; break;
    default: sp = -cs; cp = sn; break;
    }
}

// Three Gaussian deviates from four draws (the fourth Box-Muller output is
// unused by the integrators, as in the reference's rforce()).
__device__ __forceinline__ float3 gauss3(uint4& s) {
    float r = sqrtf(-2.0f * logf_series(uint_to_unit(hybrid_taus(s))));
    float th = (float)(6.283185307179586 * (double)uint_to_unit(hybrid_taus(s)));
    float sn, cs;
    sincosf_series(th, sn, cs);
    float3 g;
    g.x = r * sn;
    g.y = r * cs;
    r = sqrtf(-2.0f * logf_series(uint_to_unit(hybrid_taus(s))));
    th = (float)(6.283185307179586 * (double)uint_to_unit(hybrid_taus(s)));
    sincosf_series(th, sn, cs);
    g.z = r * sn;
    return g;
}

}  // namespace sop
