// .card planner-state files (executable/prepare.cpp:65-71).  Layout: the per-bag DP blobs of
// SubsetStructure::saveToSteam (plan/subset_structure.cpp:455-560), then the subsetToCard map as
// uint64 n followed by n x (uint64 subset mask, double cardinality) (plan/estimator.cpp:981-992),
// then int32 planning time x 10000.  The bag count is not stored, so the map is located from the
// end of the file: the unique n for which the tail parses as a map of that length.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host.h"

namespace {

bool read_file(const char *path, std::vector<unsigned char> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(f);
        return false;
    }
    buf.resize((size_t)sz);
    size_t got = sz ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return got == (size_t)sz;
}

}  // namespace

extern "C" {

int asdm_card_read(const char *path, uint64_t *masks, double *cards, uint32_t cap, uint32_t *n, double *dp_time) try {
    if (!path || !n) return ASDM_E_INVALID;
    std::vector<unsigned char> b;
    if (!read_file(path, b)) return ASDM_E_IO;
    const size_t S = b.size();
    if (S < 12) return ASDM_E_IO;
    int64_t found = -1;
    for (uint64_t k = (S - 12) / 16 + 1; k-- > 0;) {
        const size_t p = S - 4 - 16 * k - 8;
        uint64_t stored;
        std::memcpy(&stored, b.data() + p, 8);
        if (stored != k) continue;
        bool ok = true;
        for (uint64_t i = 0; i < k && ok; ++i) {
            uint64_t key;
            double v;
            std::memcpy(&key, b.data() + p + 8 + 16 * i, 8);
            std::memcpy(&v, b.data() + p + 16 + 16 * i, 8);
            ok = std::isfinite(v) && v >= 0 && key < (1ull << 32);
        }
        if (ok) {
            found = (int64_t)k;
            break;
        }
    }
    if (found < 0) return ASDM_E_IO;
    const size_t p = S - 4 - 16 * (size_t)found - 8;
    *n = (uint32_t)found;
    int32_t t;
    std::memcpy(&t, b.data() + S - 4, 4);
    if (dp_time) *dp_time = (double)t / 10000.0;
    if (!masks || !cards) return ASDM_OK;
    if (cap < (uint32_t)found) return ASDM_E_INVALID;
    for (int64_t i = 0; i < found; ++i) {
        std::memcpy(&masks[i], b.data() + p + 8 + 16 * i, 8);
        std::memcpy(&cards[i], b.data() + p + 16 + 16 * i, 8);
    }
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_card_write(const char *path, const uint64_t *masks, const double *cards, uint32_t n, double dp_time) try {
    if (!path || (n && (!masks || !cards))) return ASDM_E_INVALID;
    FILE *f = std::fopen(path, "wb");
    if (!f) return ASDM_E_IO;
    uint64_t m = n;
    bool ok = std::fwrite(&m, 8, 1, f) == 1;
    for (uint32_t i = 0; i < n && ok; ++i) {
        double scaled = (double)(int64_t)(cards[i] * 100.0) / 100.0;  // two decimals, as the reference
        ok = std::fwrite(&masks[i], 8, 1, f) == 1 && std::fwrite(&scaled, 8, 1, f) == 1;
    }
    int32_t t = (int32_t)(dp_time * 10000);
    ok = ok && std::fwrite(&t, 4, 1, f) == 1;
    return std::fclose(f) == 0 && ok ? ASDM_OK : ASDM_E_IO;
} ASDM_HOST_CATCH

}  // extern "C"
