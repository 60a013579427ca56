// softdouble.hpp -- correctly rounded IEEE-754 binary64 fma/add/mul and the
// binary64 -> binary32 conversion, computed with integer ALU operations only.
//
// The reference verifier (src/rknn_query.cpp:679-753) accumulates in double
// with GCC's contracted fma.  The GPU kernels reproduce those values bit for
// bit without issuing FP64 instructions: every double is carried as its bit
// pattern and each operation rounds once, to nearest-even, exactly like the
// hardware instruction it stands for.  Inputs are finite; results that
// overflow become infinities.  Shared by host and device code.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SD_HD __host__ __device__ __forceinline__
#else
#define SD_HD static inline
#endif

typedef unsigned __int128 sd_u128;

struct sd_num {        // value = (-1)^neg * m * 2^e, m == 0 means zero
    int neg;
    uint64_t m;
    int e;
};

SD_HD int sd_clz128(sd_u128 x) {
    uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    if (hi) return __builtin_clzll(hi);
    if (lo) return 64 + __builtin_clzll(lo);
    return 128;
}

SD_HD sd_num sd_unpack(uint64_t b) {
    sd_num r;
    r.neg = (int)(b >> 63);
    int E = (int)((b >> 52) & 0x7ff);
    uint64_t F = b & 0xfffffffffffffull;
    if (E == 0) { r.m = F; r.e = -1074; }
    else { r.m = F | (1ull << 52); r.e = E - 1075; }
    return r;
}

SD_HD uint64_t sd_from_float(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    uint64_t s = (uint64_t)(u >> 31) << 63;
    int E = (int)((u >> 23) & 0xff);
    uint32_t F = u & 0x7fffff;
    if (E == 255) return s | (0x7ffull << 52) | ((uint64_t)F << 29);
    if (E == 0) {
        if (F == 0) return s;
        int sh = __builtin_clz(F) - 8;        // normalise to bit 23
        F <<= sh;
        E = 1 - sh;
        F &= 0x7fffff;
    }
    return s | ((uint64_t)(E - 127 + 1023) << 52) | ((uint64_t)F << 29);
}

// Round (neg, m * 2^e) to prec bits with minimum exponent emin (subnormals),
// ties to even.  Returns mantissa (< 2^prec) and exponent through refs.
SD_HD void sd_round(sd_u128 m, int e, int prec, int emin, uint64_t& mo, int& eo) {
    if (m == 0) { mo = 0; eo = emin; return; }
    int p = 127 - sd_clz128(m);               // msb index
    int s = p - (prec - 1);
    if (e + s < emin) s = emin - e;
    if (s <= 0) {                             // exact (shift left or none)
        mo = (uint64_t)(m << (-s));
        eo = e + s;
        return;
    }
    sd_u128 q = s >= 128 ? 0 : (m >> s);
    sd_u128 rem = s >= 128 ? m : (m & ((((sd_u128)1) << s) - 1));
    sd_u128 half = s >= 129 ? 0 : (((sd_u128)1) << (s - 1));
    if (s >= 129) {                           // everything below half an ulp
        mo = 0;
        eo = e + s;
        return;
    }
    if (rem > half || (rem == half && (q & 1))) q += 1;
    if (q >> prec) { q >>= 1; s += 1; }
    mo = (uint64_t)q;
    eo = e + s;
}

SD_HD uint64_t sd_pack(int neg, sd_u128 m, int e) {
    uint64_t mo;
    int eo;
    sd_round(m, e, 53, -1074, mo, eo);
    uint64_t s = (uint64_t)neg << 63;
    if (mo == 0) return s;
    if (mo < (1ull << 52)) return s | mo;    // subnormal (eo == -1074)
    int E = eo + 1075;
    if (E >= 2047) return s | (0x7ffull << 52);
    return s | ((uint64_t)E << 52) | (mo & 0xfffffffffffffull);
}

// fma(a, b, c) with a single rounding.
SD_HD uint64_t sd_fma(uint64_t a, uint64_t b, uint64_t c) {
    sd_num A = sd_unpack(a), B = sd_unpack(b), C = sd_unpack(c);
    int pneg = A.neg ^ B.neg;
    sd_u128 pm = (sd_u128)A.m * (sd_u128)B.m;
    int pe = A.e + B.e;
    if (pm == 0) {
        if (C.m == 0) return ((uint64_t)(pneg & C.neg)) << 63;
        return c;
    }
    if (C.m == 0) return sd_pack(pneg, pm, pe);
    // normalise both to msb at bit 125 (two spare bits for the carry)
    int kp = sd_clz128(pm) - 2;
    sd_u128 xm = pm << kp;
    int xe = pe - kp;
    sd_u128 cm = (sd_u128)C.m;
    int kc = sd_clz128(cm) - 2;
    sd_u128 ym = cm << kc;
    int ye = C.e - kc;
    int xneg = pneg, yneg = C.neg;
    if (ye > xe || (ye == xe && ym > xm)) {   // x := larger magnitude
        sd_u128 tm = xm; xm = ym; ym = tm;
        int te = xe; xe = ye; ye = te;
        int tn = xneg; xneg = yneg; yneg = tn;
    }
    int diff = xe - ye;
    if (diff > 0) {                           // align with a jammed sticky bit
        if (diff >= 127) ym = (ym != 0) ? 1 : 0;
        else {
            sd_u128 lost = ym & ((((sd_u128)1) << diff) - 1);
            ym = (ym >> diff) | (lost != 0 ? 1 : 0);
        }
    }
    sd_u128 r = (xneg == yneg) ? xm + ym : xm - ym;
    if (r == 0) return 0;                     // exact cancellation: +0
    return sd_pack(xneg, r, xe);
}

SD_HD uint64_t sd_add(uint64_t a, uint64_t c) {
    return sd_fma(a, 0x3ff0000000000000ull, c);
}
SD_HD uint64_t sd_sub(uint64_t a, uint64_t c) {
    return sd_fma(a, 0x3ff0000000000000ull, c ^ (1ull << 63));
}
SD_HD uint64_t sd_mul(uint64_t a, uint64_t b) {
    sd_num A = sd_unpack(a), B = sd_unpack(b);
    int neg = A.neg ^ B.neg;
    sd_u128 pm = (sd_u128)A.m * (sd_u128)B.m;
    if (pm == 0) return (uint64_t)neg << 63;
    return sd_pack(neg, pm, A.e + B.e);
}

// binary64 -> binary32, round to nearest even.
SD_HD float sd_to_float(uint64_t b) {
    sd_num A = sd_unpack(b);
    uint32_t s = (uint32_t)A.neg << 31, u;
    int E = (int)((b >> 52) & 0x7ff);
    if (E == 0x7ff) {
        u = s | 0x7f800000u | (uint32_t)((b & 0xfffffffffffffull) != 0 ? 0x400000u : 0);
    } else if (A.m == 0) {
        u = s;
    } else {
        uint64_t mo;
        int eo;
        sd_round((sd_u128)A.m, A.e, 24, -149, mo, eo);
        if (mo == 0) u = s;
        else if (mo < (1ull << 23)) u = s | (uint32_t)mo;
        else {
            int FE = eo + 150;
            if (FE >= 255) u = s | 0x7f800000u;
            else u = s | ((uint32_t)FE << 23) | ((uint32_t)mo & 0x7fffffu);
        }
    }
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// sqrt, correctly rounded (b >= 0 finite; a negative b gives the default NaN).
// The root of m 2^e (e even, m normalised to 53 or 54 bits) is the integer
// root of M = m 2^58 (56 bits, M in [2^110, 2^112)) times 2^(e/2 - 29); the
// 56-bit floor root plus a sticky bit for a non-zero remainder rounds to the
// same 53 bits as the exact root (a square root is never exactly halfway).
SD_HD uint64_t sd_sqrt(uint64_t b) {
    if ((b & 0x7fffffffffffffffull) == 0) return b;                 // +-0
    if (b >> 63) return 0x7ff8000000000000ull;
    if (((b >> 52) & 0x7ff) == 0x7ff) return b;                      // inf, NaN
    sd_num A = sd_unpack(b);
    uint64_t m = A.m;
    int e = A.e;
    const int lz = __builtin_clzll(m) - 11;                          // msb to bit 52
    m <<= lz;
    e -= lz;
    if (e & 1) { m <<= 1; e -= 1; }
    sd_u128 op = (sd_u128)m << 58, res = 0, one = ((sd_u128)1) << 126;
    while (one > op) one >>= 2;
    while (one != 0) {
        if (op >= res + one) { op -= res + one; res = (res >> 1) + one; }
        else res >>= 1;
        one >>= 2;
    }
    return sd_pack(0, (res << 1) | (op != 0 ? 1 : 0), e / 2 - 30);
}

SD_HD int sd_lt_zero_or_zero(uint64_t b) {     // value <= 0 (for max(0, x))
    return (b >> 63) || (b & 0x7fffffffffffffffull) == 0;
}
