// Shared host/device definitions for the (16,8)^3 SOGRAND decoder.
//
// Compiled by hipcc for both the host and the gfx950 device passes, always
// with -ffp-contract=off: no multiply-add below is ever fused, so host and
// device evaluate every double expression with identical rounding.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SG_HD __host__ __device__
#else
#define SG_HD
#endif

namespace sg {

constexpr int kN = 16;         // code length per dimension
constexpr int kK = 8;          // info bits per dimension
constexpr int kCube = 4096;    // LLRs per codeword
constexpr int kInfoBytes = 64; // packed info bits per codeword
constexpr int kMaxL = 3;       // largest list size the kernels hold in registers

// ---------------------------------------------------------------------------
// (16,8) systematic CRC code, generator polynomial from Koopman 0xeb:
// g(x) = (0xeb << 1) | 1 = x^8 + x^7 + x^6 + x^4 + x^2 + x + 1
// (SOGRAND_C/cubic_decoder_final.c:454-519; cubic_encoder.c:187-243).
// Row i of P is x^(15-i) mod g(x); bit (7-j) of that remainder is P[i][j].
// G = [I | P], H = [P^T | I].
// ---------------------------------------------------------------------------
constexpr unsigned kGPoly = 0x1d7u;

constexpr unsigned crc_remainder(int i)
{
    unsigned v = 1u << (15 - i);
    for (int d = 15; d >= 8; --d)
        if (v & (1u << d)) v ^= kGPoly << (d - 8);
    return v;
}

// Column c of H as an 8-bit syndrome (bit j = H[j][c]). For c < 8 this is
// also row c of P with bit j = P[c][j]: the parity bits that info bit c feeds.
constexpr uint8_t hcol(int c)
{
    if (c >= 8) return (uint8_t)(1u << (c - 8));
    unsigned r = crc_remainder(c), s = 0;
    for (int j = 0; j < 8; ++j) s |= ((r >> (7 - j)) & 1u) << j;
    return (uint8_t)s;
}

// All 16 H columns packed: bytes 0..7 in kHcolLo, 8..15 in kHcolHi.
constexpr uint64_t pack_hcols(int off)
{
    uint64_t v = 0;
    for (int c = 0; c < 8; ++c) v |= (uint64_t)hcol(off + c) << (8 * c);
    return v;
}
constexpr uint64_t kHcolLo = pack_hcols(0);
constexpr uint64_t kHcolHi = pack_hcols(8);

SG_HD inline uint8_t hcol_of(unsigned c)
{
    uint64_t w = (c & 8u) ? kHcolHi : kHcolLo;
    return (uint8_t)(w >> (8u * (c & 7u)));
}

// Syndrome of a 16-bit word (bit i = code bit i).
SG_HD inline uint32_t syndrome16(uint32_t word)
{
    uint32_t s = 0;
    for (uint32_t r = word & 0xffffu; r; r &= r - 1u) s ^= hcol_of((unsigned)__builtin_ctz(r));
    return s;
}

// Component codeword of an 8-bit message (bit a = message bit a): message in
// positions 0..7, parity in 8..15.
SG_HD inline uint32_t encode_row(uint32_t m)
{
    uint32_t p = 0;
    for (uint32_t r = m & 0xffu; r; r &= r - 1u) p ^= hcol_of((unsigned)__builtin_ctz(r));
    return (m & 0xffu) | (p << 8);
}

// ---------------------------------------------------------------------------
// Error-pattern schedule geometry. A pattern is a 16-bit mask over sorted
// positions (bit b = b-th least reliable bit); w = popcount, S = sum of (b+1),
// logistic weight wt = IC*w + S. The reference visits wt = IC+1, IC+2, ...,
// inside each wt the Hamming weights w in ascending order, and inside each
// (w, S) level the position tuples in lexicographic order
// (cubic_decoder_final.c:642-744, 775-789). For IC >= 55 the order no longer
// depends on IC (levels are then sorted by (w, S)), so 56 tables cover all IC.
// ---------------------------------------------------------------------------
constexpr int kSchedLen = 65535;    // nonzero patterns
constexpr int kSchedStride = 65536; // table stride (entry 65535 = 0, end marker)
constexpr int kSchedTables = 57;    // IC = 0..55 and one generic table for IC >= 56
constexpr int kSchedGenericIc = 56;

// The same schedules as position words, one u32 per entry: bits 1 + 5i ..
// 5 + 5i hold b + 1 for the i-th lowest set bit b of the mask (i = 0..5), 0
// where the mask has fewer bits; bit 31 is set when it has more than six (the
// rest are then read from the mask) and bit 0 when the entry starts a new
// weight level of its table (the reference's budget guard runs there, :652).
// The end marker (mask 0) is the word 0. A field is a record row of the line's
// LDS records (row 0 the null record), so a pattern's positions come out with
// one bit-field extract each.
//
// Weight levels: wt = IC w + S. For IC >= 56 two entries have equal wt only if
// they have equal (w, S) (IC (w1 - w2) = S2 - S1 needs IC <= 55), so one
// generic table serves every IC >= 56, order and levels alike.
constexpr uint32_t kPosFields = 6;
constexpr uint32_t kPosLevel = 1u;
constexpr uint32_t kPosMore = 1u << 31;
inline constexpr uint32_t pos_word(uint32_t m, bool level_start)
{
    uint32_t w = 0;
    for (uint32_t i = 0; i < kPosFields && m; ++i) {
        uint32_t b = 0;
        while (!((m >> b) & 1u)) ++b;
        w |= (b + 1u) << (1u + 5u * i);
        m &= m - 1u;
    }
    if (m) w |= kPosMore;
    return w && level_start ? (w | kPosLevel) : w;
}
static_assert(pos_word(0x1u, false) == 2u && pos_word(0x8001u, true) == (1u | 2u | (16u << 6)),
              "position word");
static_assert(pos_word(0x7fu, false) ==
                  (kPosMore | 1u << 1 | 2u << 6 | 3u << 11 | 4u << 16 | 5u << 21 | 6u << 26),
              "position word overflow");
static_assert(pos_word(0u, true) == 0u, "end marker");
// Weight level of a mask under intercept ic (getPM's landslide weight, :646-653).
inline constexpr int64_t pattern_wt(int64_t ic, uint32_t m)
{
    int64_t w = 0, sum = 0;
    for (uint32_t b = 0; b < 16; ++b)
        if ((m >> b) & 1u) {
            w += 1;
            sum += b + 1;
        }
    return ic * w + sum;
}
// Device layout of the schedule buffer: the mask tables, 8 zero masks of
// padding, then the position-word tables and 8 zero words of padding (the
// search loops prefetch one chunk past an end marker).
constexpr size_t kSchedMaskEntries = (size_t)kSchedTables * kSchedStride + 8;
constexpr size_t kSchedPosOffsetBytes = kSchedMaskEntries * 2;
constexpr size_t kSchedBufferBytes = kSchedPosOffsetBytes + ((size_t)kSchedTables * kSchedStride + 8) * 4;
static_assert(kSchedPosOffsetBytes % 16 == 0, "position tables are 16-byte aligned");

// ---------------------------------------------------------------------------
// Seeded synthetic data. Counter-based hashing (splitmix64 finaliser) so that
// codeword c depends only on (seed, c); a portable log and sin/cos built from
// IEEE basic operations only, so host and device produce identical doubles.
// ---------------------------------------------------------------------------
SG_HD inline uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z;
}

SG_HD inline uint64_t hash4(uint64_t seed, uint64_t domain, uint64_t a, uint64_t b)
{
    uint64_t k = mix64(seed + 0x9e3779b97f4a7c15ull * (domain + 1u));
    return mix64(k ^ mix64(a * 0xd1b54a32d192ed03ull + b + 0x632be59bd9b4e019ull));
}

enum : uint64_t { kDomInfo = 1, kDomU1 = 2, kDomU2 = 3 };

SG_HD inline double bits_to_double(uint64_t u)
{
    union {
        uint64_t u;
        double d;
    } c;
    c.u = u;
    return c.d;
}
SG_HD inline uint64_t double_to_bits(double d)
{
    union {
        uint64_t u;
        double d;
    } c;
    c.d = d;
    return c.u;
}

// Natural log of a positive normal double. x = m * 2^e with m in [sqrt(1/2),
// sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, series to s^23.
SG_HD inline double portable_log(double x)
{
    uint64_t b = double_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    double lnm = 2.0 * s * p;
    const double ln2_hi = 6.93147180369123816490e-01; // upper bits of ln 2 (exact in e*ln2_hi)
    const double ln2_lo = 1.90821492927058770002e-10;
    double de = (double)e;
    return de * ln2_hi + (de * ln2_lo + lnm);
}

// sin and cos of phi in [0, pi/4] by Taylor series (terms to phi^19 / phi^18).
SG_HD inline void portable_sincos_small(double phi, double *s, double *c)
{
    double z = phi * phi;
    double ps = -1.0 / 121645100408832000.0; // -1/19!
    ps = ps * z + 1.0 / 355687428096000.0;   // 1/17!
    ps = ps * z - 1.0 / 1307674368000.0;     // 1/15!
    ps = ps * z + 1.0 / 6227020800.0;        // 1/13!
    ps = ps * z - 1.0 / 39916800.0;          // 1/11!
    ps = ps * z + 1.0 / 362880.0;            // 1/9!
    ps = ps * z - 1.0 / 5040.0;              // 1/7!
    ps = ps * z + 1.0 / 120.0;               // 1/5!
    ps = ps * z - 1.0 / 6.0;                 // 1/3!
    *s = phi + phi * (z * ps);
    double pc = 1.0 / 6402373705728000.0; // 1/18!
    pc = pc * z - 1.0 / 20922789888000.0; // 1/16!
    pc = pc * z + 1.0 / 87178291200.0;    // 1/14!
    pc = pc * z - 1.0 / 479001600.0;      // 1/12!
    pc = pc * z + 1.0 / 3628800.0;        // 1/10!
    pc = pc * z - 1.0 / 40320.0;          // 1/8!
    pc = pc * z + 1.0 / 720.0;            // 1/6!
    pc = pc * z - 1.0 / 24.0;             // 1/4!
    pc = pc * z + 0.5;                    // 1/2!
    *c = 1.0 - z * pc;
}

// cos and sin of 2*pi*u for u = j * 2^-53 in [0, 1): octant reduction is exact
// because 8u and 1 - frac(8u) are exact in double.
SG_HD inline void portable_cossin_2pi(double u, double *cs, double *sn)
{
    const double quarter_pi = 0.78539816339744830962;
    double v = u * 8.0;
    int k = (int)v;
    double f = v - (double)k;
    double s, c;
    if (k & 1) {
        portable_sincos_small((1.0 - f) * quarter_pi, &s, &c);
        double t = s;
        s = c;
        c = t;
    } else {
        portable_sincos_small(f * quarter_pi, &s, &c);
    }
    switch ((k >> 1) & 3) {
    case 0: *cs = c; *sn = s; break;
    case 1: *cs = -s; *sn = c; break;
    case 2: *cs = -c; *sn = -s; break;
    default: *cs = s; *sn = -c; break;
    }
}

// Standard normal pair k of codeword c (Box-Muller; u1 in (0, 1] never 0).
SG_HD inline void normal_pair(uint64_t seed, uint64_t c, uint64_t k, double *z0, double *z1)
{
    uint64_t h1 = hash4(seed, kDomU1, c, k), h2 = hash4(seed, kDomU2, c, k);
    double u1 = (double)((h1 >> 11) + 1u) * 1.1102230246251565404e-16; // 2^-53
    double u2 = (double)(h2 >> 11) * 1.1102230246251565404e-16;
    double r = __builtin_sqrt(-2.0 * portable_log(u1));
    double cs, sn;
    portable_cossin_2pi(u2, &cs, &sn);
    *z0 = r * cs;
    *z1 = r * sn;
}

// Info bytes of codeword c: the 64 little-endian bytes of 8 hashed words.
SG_HD inline uint64_t info_word(uint64_t seed, uint64_t c, int j)
{
    return hash4(seed, kDomInfo, c, (uint64_t)j);
}

SG_HD inline uint32_t rev8(uint32_t x)
{
    x = ((x & 0xf0u) >> 4) | ((x & 0x0fu) << 4);
    x = ((x & 0xccu) >> 2) | ((x & 0x33u) << 2);
    x = ((x & 0xaau) >> 1) | ((x & 0x55u) << 1);
    return x;
}

// Product encoding into 256 words (word a1 + 16*a2, bit a0 = code bit
// a0 + 16*a1 + 256*a2) from 64 info bytes (bit b = a0 + 8*a1 + 64*a2 in byte
// b/8, MSB first). Same codeword as the reference encoder (cubic_encoder.c:23-100).
SG_HD inline void encode_cube_words(const uint8_t *info, uint32_t *W)
{
    for (int i2 = 0; i2 < 8; ++i2)
        for (int i1 = 0; i1 < 8; ++i1) W[i1 + 16 * i2] = encode_row(rev8(info[i1 + 8 * i2]));
    for (int i2 = 0; i2 < 8; ++i2)
        for (int a1 = 8; a1 < 16; ++a1) {
            uint32_t acc = 0;
            for (int i1 = 0; i1 < 8; ++i1)
                if ((hcol_of((unsigned)i1) >> (a1 - 8)) & 1u) acc ^= W[i1 + 16 * i2];
            W[a1 + 16 * i2] = acc;
        }
    for (int a2 = 8; a2 < 16; ++a2)
        for (int a1 = 0; a1 < 16; ++a1) {
            uint32_t acc = 0;
            for (int i2 = 0; i2 < 8; ++i2)
                if ((hcol_of((unsigned)i2) >> (a2 - 8)) & 1u) acc ^= W[a1 + 16 * i2];
            W[a1 + 16 * a2] = acc;
        }
}

} // namespace sg
