// Host check of the schedule's position words (sogrand_amd/csrc/sg_common.h,
// built as sogrand_amd/csrc/sg_api.cpp builds them from sg_schedule.cpp's
// masks): every word decodes to its mask's lowest six set bits in ascending
// order, carries the overflow bit exactly when the mask has more, and carries
// the level bit exactly where the weight level IC w + S changes. For the
// generic table, the level bits are also checked against intercepts 56 ..
// 2000 and a few large ones: its levels must not depend on IC there.
// Prints "ok" or the first failure.
#include <cstdio>
#include <vector>

#include "sg_common.h"
#include "sg_schedule.h"

using namespace sg;

static bool level_start(int64_t ic, const uint16_t *m, size_t j)
{
    return j == 0 || pattern_wt(ic, m[j]) != pattern_wt(ic, m[j - 1]);
}

int main()
{
    std::vector<uint16_t> t((size_t)kSchedStride);
    for (int tbl = 0; tbl < kSchedTables; ++tbl) {
        build_schedule_table(tbl, t.data());
        for (size_t j = 0; j < (size_t)kSchedStride; ++j) {
            const uint32_t m = t[j];
            const uint32_t w = pos_word(m, level_start(tbl, t.data(), j));
            if (m == 0) {
                if (w != 0) { printf("table %d entry %zu: end marker word %08x\n", tbl, j, w); return 1; }
                continue;
            }
            uint32_t r = m;
            for (uint32_t i = 0; i < kPosFields; ++i) {
                const uint32_t f = (w >> (1 + 5 * i)) & 31u;
                const uint32_t want = r ? (uint32_t)__builtin_ctz(r) + 1u : 0u;
                if (f != want) { printf("table %d entry %zu field %u: %u != %u\n", tbl, j, i, f, want); return 1; }
                if (r) r &= r - 1u;
            }
            if (((w & kPosMore) != 0) != (r != 0)) { printf("table %d entry %zu: overflow bit\n", tbl, j); return 1; }
            if (((w & kPosLevel) != 0) != level_start(tbl, t.data(), j)) {
                printf("table %d entry %zu: level bit\n", tbl, j);
                return 1;
            }
        }
    }
    // the generic table (built for IC = kSchedGenericIc) serves every IC >= it
    build_schedule_table(kSchedGenericIc, t.data());
    std::vector<int64_t> ics;
    for (int64_t ic = kSchedGenericIc; ic <= 2000; ++ic) ics.push_back(ic);
    for (int64_t ic : {100000ll, 1ll << 24, 1ll << 40}) ics.push_back(ic);
    // weight level IC w + S from each entry's (w, S), computed once
    std::vector<int64_t> hw((size_t)kSchedLen), sum((size_t)kSchedLen);
    std::vector<char> generic((size_t)kSchedLen);
    for (size_t j = 0; j < (size_t)kSchedLen; ++j) {
        hw[j] = __builtin_popcount(t[j]);
        sum[j] = pattern_wt(0, t[j]);
        generic[j] = level_start(kSchedGenericIc, t.data(), j);
    }
    for (int64_t ic : ics)
        for (size_t j = 1; j < (size_t)kSchedLen; ++j) {
            const bool here = ic * hw[j] + sum[j] != ic * hw[j - 1] + sum[j - 1];
            if (here != (generic[j] != 0)) {
                printf("IC %lld entry %zu: levels differ from the generic table's\n", (long long)ic, j);
                return 1;
            }
        }
    printf("ok\n");
    return 0;
}
