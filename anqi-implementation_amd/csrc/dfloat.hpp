// dfloat.hpp -- float-float sums with error-free transformations, and the
// certified rounding tests that let FP32 arithmetic reproduce a binary64
// result's rounding to binary32 bit for bit.
//
// The verifier (src/rknn_query.cpp:679-753) evaluates short sums in binary64
// and rounds them to float (coefficients, residuals, thresholds).  A sum of
// float terms and exact float products is carried here as hi + lo (two
// floats) using Knuth's TwoSum and the fma-based exact product, so the only
// rounding error is in the addition that folds the old low part into the
// new one; |that addition's result| is accumulated in `err`, so the estimate
// hi + lo differs from the exact sum by at most 2^-24 err.  The binary64
// evaluation of the same sum differs from the exact sum by at most dref (an
// a-priori bound the caller passes: 2^-53 per binary64 rounding times the
// magnitudes involved).  When every real number within |lo| + 2^-24 err +
// dref of hi rounds to the float hi, the binary64 result rounds to hi as
// well: the float answer is certified.  Otherwise the caller recomputes that
// element in soft binary64.  Compile without fp contraction: every product
// and sum below must round on its own.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FF_HD __host__ __device__ __forceinline__
#else
#define FF_HD static inline
#endif

FF_HD uint32_t ff_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
FF_HD float ff_from_bits(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

struct FFSum {
    float hi, lo;   // estimate hi + lo of the exact sum
    float err;      // sum of |folded low parts|: estimation error <= 2^-24 err
    float mag;      // sum of |terms| (for the binary64 error bound)
    int ok;         // 0 if a term was outside the range the analysis covers
};

FF_HD void ff_two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

FF_HD FFSum ff_zero() { return FFSum{0.0f, 0.0f, 0.0f, 0.0f, 1}; }

// Terms of magnitude in [2^-90, 2^100] (or zero) keep every TwoSum and exact
// product free of underflow and overflow.
FF_HD int ff_in_range(float t) {
    const float a = fabsf(t);
    return a == 0.0f || (a >= 0x1p-90f && a <= 0x1p100f);
}

// x += t (t a float)
FF_HD void ff_add(FFSum& x, float t) {
    x.ok &= ff_in_range(t);
    float s, e;
    ff_two_sum(x.hi, t, s, e);
    const float f = e + x.lo;        // the one rounding of the step
    x.err += fabsf(f);
    ff_two_sum(s, f, x.hi, x.lo);
    x.mag += fabsf(t);
}

// x += a * b, the product taken exactly as p + q (fma)
FF_HD void ff_add_prod(FFSum& x, float a, float b) {
    const float p = a * b;
    const float q = fmaf(a, b, -p);
    x.ok &= ff_in_range(p) && (p == 0.0f || fabsf(p) >= 0x1p-60f);   // q normal or exact
    ff_add(x, p);
    ff_add(x, q);
    x.mag -= fabsf(q);               // the product counts once in mag
}

// Short sums (a few terms): the running float s takes each term through
// TwoSum and every exact error (TwoSum's and the products' fma remainders)
// is added into one float w.  With ns TwoSum steps and nw additions into w,
// |w|'s partial sums stay below 2^-24 (ns + 1) mag, so the nw roundings of w
// cost at most nw (ns + 1) 2^-48 mag: an a-priori bound, no per-step
// tracking.  ff_short_finish returns hi + lo = s + w with err set to that
// bound over 2^-24 (so ff_radius applies unchanged).
struct FFShort {
    float s, w, mag;
    int ok;
};
FF_HD FFShort ffs_start(float a) { return FFShort{a, 0.0f, fabsf(a), ff_in_range(a)}; }
FF_HD void ffs_add(FFShort& x, float t) {
    float s, e;
    ff_two_sum(x.s, t, s, e);
    x.s = s;
    x.w += e;
    x.mag += fabsf(t);
    x.ok &= ff_in_range(t);
}
FF_HD void ffs_add_prod(FFShort& x, float a, float b) {
    const float p = a * b;
    const float q = fmaf(a, b, -p);
    x.ok &= p == 0.0f || (fabsf(p) >= 0x1p-60f && fabsf(p) <= 0x1p100f);   // q normal or exact
    ffs_add(x, p);
    x.w += q;
}
FF_HD FFSum ffs_finish(const FFShort& x, int ns, int nw) {
    FFSum r;
    ff_two_sum(x.s, x.w, r.hi, r.lo);
    r.mag = x.mag;
    r.err = (float)(nw * (ns + 1)) * 0x1p-24f * x.mag;
    r.ok = x.ok && ff_in_range(x.w);
    return r;
}

// The total distance from hi that the binary64 result may be: |lo| + the
// estimate's error + dref.  The 2^-24 err factor carries 2^-6 of slack for
// the float sum that formed err; the final factor covers the three roundings
// of this expression (each at most 2^-24 relative).
FF_HD float ff_radius(const FFSum& x, float dref) {
    return (fabsf(x.lo) + (x.err * 0x1.04p-24f + dref)) * (1.0f + 0x1p-22f);
}

// True if every real within ff_radius of hi rounds to the float hi (round to
// nearest even): the binary64 result, rounded to float, is hi.
FF_HD bool ff_round_certain(const FFSum& x, float dref) {
    if (!x.ok) return false;
    const float h = fabsf(x.hi);
    if (!(h >= 0x1p-100f && h <= 0x1p100f)) return false;   // zero, tiny, huge or NaN
    const uint32_t u = ff_bits(h);
    const float up = ff_from_bits(u + 1u) - h;            // gap above (exact)
    const float dn = h - ff_from_bits(u - 1u);            // gap below (exact)
    return ff_radius(x, dref) < 0.5f * fminf(up, dn);
}

// A float interval [lo_f, hi_f] that holds the binary64 result rounded to
// float, whatever that rounding is (false if the sum is outside the range).
FF_HD bool ff_float_interval(const FFSum& x, float dref, float& lo_f, float& hi_f) {
    if (!x.ok || !(fabsf(x.hi) <= 0x1p100f)) return false;
    const float r = ff_radius(x, dref);
    const float a = x.hi - r, b = x.hi + r;
    // one step outward covers the rounding of a and b
    lo_f = a > 0.0f ? ff_from_bits(ff_bits(a) - 1u)
                    : (a < 0.0f ? ff_from_bits(ff_bits(a) + 1u) : -0x1p-149f);
    hi_f = b > 0.0f ? ff_from_bits(ff_bits(b) + 1u)
                    : (b < 0.0f ? ff_from_bits(ff_bits(b) - 1u) : 0x1p-149f);
    return true;
}
