// exp_ff.h -- e^x rounded to f32 from single-precision arithmetic only.
//
// Tang's table method carried in float-float (pairs of f32):
//   x = n ln2/64 + r,  |r| <= ln2/128,  n = 64 m + j,
//   e^x = 2^m * 2^(j/64) * e^r,
// with ln2/64 split in three f32 parts (the first 10 bits wide, so n times it
// is exact for |n| < 2^14), 2^(j/64) tabulated as an f32 pair hi + lo (its
// value to about 2^-48), and e^r - 1 as its Taylor polynomial to r^6 with
// the r and r^2 terms kept as pairs.  The parts are summed exactly (TwoSum)
// and rounded once, so the result is the correctly rounded f32 unless e^x
// lies within about 2^-45 (relative) of a rounding boundary; results in the
// subnormal range are rounded twice (to the f32 significand, then by the
// scaling).  Compiled for the device by hipcc and for the host by g++ (the
// CPU test holds it to the double-precision exponential).
#pragma once

#if defined(__HIPCC__)
#define IXG_EXP_FN __device__ __forceinline__
#define IXG_EXP_TABLE __constant__ const
#else
#define IXG_EXP_FN static inline
#define IXG_EXP_TABLE static const
#endif

namespace ixg_exp {

// 2^(j/64), j = 0..63: the f32 nearest, and the f32 nearest the remainder.
IXG_EXP_TABLE float kExp2Hi[64] = {
    0x1.0000000000000p+0f,
    0x1.02c9a40000000p+0f,
    0x1.059b0e0000000p+0f,
    0x1.0874520000000p+0f,
    0x1.0b55860000000p+0f,
    0x1.0e3ec40000000p+0f,
    0x1.11301e0000000p+0f,
    0x1.1429aa0000000p+0f,
    0x1.172b840000000p+0f,
    0x1.1a35be0000000p+0f,
    0x1.1d48740000000p+0f,
    0x1.2063b80000000p+0f,
    0x1.2387a60000000p+0f,
    0x1.26b4560000000p+0f,
    0x1.29e9e00000000p+0f,
    0x1.2d285a0000000p+0f,
    0x1.306fe00000000p+0f,
    0x1.33c08c0000000p+0f,
    0x1.371a740000000p+0f,
    0x1.3a7db40000000p+0f,
    0x1.3dea640000000p+0f,
    0x1.4160a20000000p+0f,
    0x1.44e0860000000p+0f,
    0x1.486a2c0000000p+0f,
    0x1.4bfdae0000000p+0f,
    0x1.4f9b280000000p+0f,
    0x1.5342b60000000p+0f,
    0x1.56f4740000000p+0f,
    0x1.5ab07e0000000p+0f,
    0x1.5e76f20000000p+0f,
    0x1.6247ec0000000p+0f,
    0x1.6623880000000p+0f,
    0x1.6a09e60000000p+0f,
    0x1.6dfb240000000p+0f,
    0x1.71f75e0000000p+0f,
    0x1.75feb60000000p+0f,
    0x1.7a11480000000p+0f,
    0x1.7e2f340000000p+0f,
    0x1.82589a0000000p+0f,
    0x1.868d9a0000000p+0f,
    0x1.8ace540000000p+0f,
    0x1.8f1aea0000000p+0f,
    0x1.93737c0000000p+0f,
    0x1.97d82a0000000p+0f,
    0x1.9c49180000000p+0f,
    0x1.a0c6680000000p+0f,
    0x1.a5503c0000000p+0f,
    0x1.a9e6b60000000p+0f,
    0x1.ae89fa0000000p+0f,
    0x1.b33a2c0000000p+0f,
    0x1.b7f7700000000p+0f,
    0x1.bcc1ea0000000p+0f,
    0x1.c199be0000000p+0f,
    0x1.c67f120000000p+0f,
    0x1.cb720e0000000p+0f,
    0x1.d072d40000000p+0f,
    0x1.d5818e0000000p+0f,
    0x1.da9e600000000p+0f,
    0x1.dfc9740000000p+0f,
    0x1.e502ee0000000p+0f,
    0x1.ea4afa0000000p+0f,
    0x1.efa1be0000000p+0f,
    0x1.f507660000000p+0f,
    0x1.fa7c180000000p+0f
};
IXG_EXP_TABLE float kExp2Lo[64] = {
    0x0.0p+0f,
    -0x1.887fa00000000p-28f,
    -0x1.9d4f520000000p-25f,
    -0x1.e2990e0000000p-26f,
    0x1.9f31220000000p-25f,
    -0x1.a585cc0000000p-25f,
    -0x1.fdb4960000000p-25f,
    0x1.d525bc0000000p-25f,
    -0x1.c157420000000p-27f,
    0x1.6df96e0000000p-25f,
    -0x1.d2e8ca0000000p-25f,
    0x1.0c519a0000000p-25f,
    0x1.ceac480000000p-25f,
    0x1.789f380000000p-26f,
    -0x1.5c04240000000p-25f,
    0x1.b900c20000000p-26f,
    0x1.4636e20000000p-25f,
    -0x1.b37d200000000p-25f,
    -0x1.18aac60000000p-25f,
    -0x1.634c020000000p-25f,
    0x1.8246840000000p-25f,
    0x1.f72e2a0000000p-28f,
    0x1.8624b40000000p-30f,
    -0x1.47d8660000000p-25f,
    -0x1.593abc0000000p-25f,
    -0x1.2c5a6c0000000p-25f,
    -0x1.2c56100000000p-25f,
    -0x1.295b040000000p-25f,
    -0x1.5bd5ec0000000p-27f,
    -0x1.4a5bd60000000p-25f,
    -0x1.f8b5500000000p-25f,
    0x1.2a91120000000p-27f,
    0x1.9fcef40000000p-26f,
    -0x1.cd72e80000000p-27f,
    0x1.1d8bee0000000p-25f,
    -0x1.37b3060000000p-25f,
    -0x1.829fd00000000p-25f,
    -0x1.2616340000000p-25f,
    -0x1.accc7c0000000p-26f,
    -0x1.2edb440000000p-26f,
    0x1.15506e0000000p-27f,
    -0x1.baa2320000000p-26f,
    -0x1.e647440000000p-25f,
    -0x1.0d8d840000000p-31f,
    0x1.51f8480000000p-27f,
    -0x1.2886a60000000p-26f,
    -0x1.b83b540000000p-25f,
    -0x1.50c0480000000p-25f,
    -0x1.a94b140000000p-26f,
    -0x1.ec3a820000000p-26f,
    -0x1.a094380000000p-25f,
    -0x1.f687c60000000p-25f,
    -0x1.3d56b20000000p-27f,
    0x1.cafa2a0000000p-25f,
    -0x1.8837cc0000000p-27f,
    0x1.40f1300000000p-25f,
    -0x1.822dbc0000000p-27f,
    0x1.ed99420000000p-27f,
    -0x1.908c940000000p-25f,
    0x1.e2cffe0000000p-26f,
    0x1.52486c0000000p-27f,
    0x1.cc2b440000000p-25f,
    -0x1.246eb00000000p-26f,
    0x1.9e90d80000000p-28f
};

IXG_EXP_FN float exp_ff(float x) {
    if (!(x < 88.72283935546875f)) return x != x ? x : __builtin_inff();  // overflow, or NaN
    if (x < -103.97208404541015625f) return 0.0f;                         // below half the least subnormal
    const float kInvL = 0x1.715476p+6f;   // 64 / ln2
    const float kL1 = 0x1.628p-7f;        // ln2 / 64 = kL1 + kL2 + kL3
    const float kL2 = 0x1.90bfbep-17f;
    const float kL3 = 0x1.1cf79ap-42f;
    const float n = __builtin_rintf(x * kInvL);
    // r = x - n ln2/64 as r_hi + r_lo
    const float a = __builtin_fmaf(-n, kL1, x);  // exact: n kL1 has at most 24 bits, a is near x
    const float p2 = n * kL2;
    const float p2e = __builtin_fmaf(n, kL2, -p2);  // n kL2 - p2, exactly
    const float r_hi = a - p2;
    const float bb = r_hi - a;
    const float r_lo = ((a - (r_hi - bb)) + (-p2 - bb)) - __builtin_fmaf(n, kL3, p2e);
    // e^r - 1 = r + r^2/2 + r^3/6 + ... with the leading two terms as pairs
    const float sq = r_hi * r_hi;
    const float sq_lo = __builtin_fmaf(r_hi, r_hi, -sq) + 2.0f * r_hi * r_lo;
    const float tail =
        sq * r_hi *
        __builtin_fmaf(r_hi, __builtin_fmaf(r_hi, __builtin_fmaf(r_hi, 1.0f / 720.0f, 1.0f / 120.0f), 1.0f / 24.0f),
                       1.0f / 6.0f);
    const float h = 0.5f * sq;
    const float p_hi = r_hi + h;
    const float pb = p_hi - r_hi;
    const float p_lo = ((r_hi - (p_hi - pb)) + (h - pb)) + ((r_lo + 0.5f * sq_lo) + tail);
    // 2^(j/64) (1 + p) = t_hi + [t_hi p_hi + t_hi p_lo + t_lo + t_lo p_hi]
    const int ni = (int)n;
    const int j = ni & 63;
    const float t_hi = kExp2Hi[j], t_lo = kExp2Lo[j];
    const float m_hi = t_hi * p_hi;
    const float m_lo = __builtin_fmaf(t_hi, p_hi, -m_hi);
    const float c = m_lo + __builtin_fmaf(t_hi, p_lo, __builtin_fmaf(t_lo, p_hi, t_lo));
    const float s_hi = t_hi + m_hi;
    const float sb = s_hi - t_hi;
    const float s_lo = (t_hi - (s_hi - sb)) + (m_hi - sb);
    const float y = s_hi + (s_lo + c);
    return __builtin_ldexpf(y, ni >> 6);  // ni >> 6 = floor(n / 64)
}

}  // namespace ixg_exp
