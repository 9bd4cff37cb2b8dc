// pfm_div.h -- correctly rounded fp32 division by a small positive integer
// (the diagonal of the Poisson stencil), usable from host and device code.
//
// Method: Markstein's correction step (P. Markstein, "Computation of
// elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34(1), 1990; J.-M. Muller et al., Handbook of Floating-Point
// Arithmetic, ch. "Newton-Raphson-based division with an FMA"): with
// y = RN(1/n) and q0 = RN(x y), the remainder r = x - q0 n is exact under an
// FMA and q = RN(q0 + r y) is RN(x / n).  Checked against IEEE division over
// every finite fp32 x for n = 1..6: the identity holds for all x with
// 2^-125 <= |x| <= FLT_MAX; zeros (the sign of -0), tiny x whose quotient is
// subnormal (n = 6), infinities and NaN take the IEEE division instead.
#pragma once

#if defined(__HIPCC__)
#define PFM_HD __host__ __device__ __forceinline__
#else
#define PFM_HD static inline
#endif

// RN(1/n) for n = 1..6; 0 otherwise (those n take the IEEE division).
PFM_HD float pfm_rcp_small(int n) {
    float y = 0.0f;
    y = n == 1 ? 1.0f : y;
    y = n == 2 ? 0.5f : y;
    y = n == 3 ? 0x1.555556p-2f : y;
    y = n == 4 ? 0.25f : y;
    y = n == 5 ? 0x1.99999ap-3f : y;
    y = n == 6 ? 0x1.555556p-3f : y;
    return y;
}

// RN(x / nf) where nf = (float)n and y = pfm_rcp_small(n).
PFM_HD float pfm_div_small(float x, float nf, float y) {
    const float q0 = x * y;
    const float r = __builtin_fmaf(-q0, nf, x);
    float q = __builtin_fmaf(r, y, q0);
    const float ax = __builtin_fabsf(x);
    if (__builtin_expect(!(ax >= 0x1p-125f && ax <= 0x1.fffffep127f && y != 0.0f), 0)) q = x / nf;
    return q;
}
