// Host construction of the error-pattern schedule tables.
//
// The reference generates patterns on the fly with its landslide loop
// (SOGRAND_C/cubic_decoder_final.c:642-744, mountain_build :775-789): for each
// logistic weight wt = IC+1, IC+2, ..., for each Hamming weight w (ascending)
// with a nonempty level, every w-subset of sorted positions whose position
// sum S = wt - IC*w, in lexicographic order of the ascending position tuple.
// Here the same order is materialised once per intercept: the decoder streams
// masks from a table instead of running a data-dependent successor function.
#include "sg_schedule.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace sg {

namespace {

int popc(unsigned m) { return __builtin_popcount(m); }

int possum(unsigned m)
{
    int s = 0;
    for (unsigned r = m; r; r &= r - 1u) s += __builtin_ctz(r) + 1;
    return s;
}

unsigned rev16(unsigned m)
{
    unsigned r = 0;
    for (int b = 0; b < 16; ++b)
        if (m & (1u << b)) r |= 1u << (15 - b);
    return r;
}

long long smin(int w) { return (long long)w * (w + 1) / 2; }
long long smax(int w) { return (long long)w * (33 - w) / 2; }

struct Levels {
    std::vector<uint16_t> masks;  // grouped by (w, S), lexicographic inside a group
    int start[17][138];           // [w][S] -> first index
    int count[17][137];
    Levels() : masks(kSchedLen)
    {
        std::memset(count, 0, sizeof count);
        for (unsigned m = 1; m < 65536u; ++m) count[popc(m)][possum(m)]++;
        int off = 0;
        for (int w = 0; w <= 16; ++w)
            for (int s = 0; s <= 137; ++s) {
                start[w][s] = off;
                if (s <= 136) off += count[w][s];
            }
        int fill[17][137];
        std::memset(fill, 0, sizeof fill);
        // Lexicographic order of ascending position tuples is descending order
        // of the bit-reversed mask: the lowest differing position decides.
        for (int r = 65535; r >= 1; --r) {
            unsigned m = rev16((unsigned)r);
            int w = popc(m), s = possum(m);
            masks[start[w][s] + fill[w][s]++] = (uint16_t)m;
        }
    }
};

} // namespace

void build_schedule_table(int ic, uint16_t *out)
{
    static const Levels lv;
    int n = 0;
    long long IC = ic;
    // Walk wt upward; skip weights with no pattern at all in one step.
    long long wt = IC + 1;
    for (;;) {
        bool any = false;
        for (int w = 1; w <= 16; ++w) {
            long long S = wt - IC * w;
            if (S < smin(w) || S > smax(w)) continue;
            any = true;
            int a = lv.start[w][S], c = lv.count[w][S];
            std::memcpy(out + n, lv.masks.data() + a, sizeof(uint16_t) * (size_t)c);
            n += c;
        }
        if (n >= kSchedLen) break;
        long long next = -1;
        for (int w = 1; w <= 16; ++w) {
            long long lo = IC * w + smin(w), hi = IC * w + smax(w);
            long long cand = std::max(wt + 1, lo);
            if (cand <= hi && (next < 0 || cand < next)) next = cand;
        }
        (void)any;
        if (next < 0) break;
        wt = next;
    }
    for (int i = n; i < kSchedStride; ++i) out[i] = 0; // end marker(s)
}

void build_all_schedule_tables(uint16_t *out)
{
    for (int t = 0; t < kSchedTables; ++t) build_schedule_table(t, out + (size_t)t * kSchedStride);
}

} // namespace sg
