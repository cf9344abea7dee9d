// Pattern class tables: quick code -> class id.
//
// Reference: QuickMapping (src/main/QuickMapping.cpp:5-43) loads these
// tables from dictionaries/k.csv, which were produced offline with nauty
// (src/motifs_DM_DFS/Manager.cpp:1067-1156).  Here they are regenerated on the
// host by exhaustive permutation: every graph on k vertices is grouped into
// its isomorphism class (orbit under all k! relabellings).  The class ids
// must be the reference's: a class gets id i when the reference's canonical
// code of class i (the cg_global column of dictionaries/k.csv, kept below as
// data) lies in its orbit.  For k = 7 the reference ships no table; ids there
// are this library's own (classes ranked by their largest code) and match no
// reference file.
//
// Quick code layout (src/motifs_DM_DFS/Manager.cpp:1069-1097): vertices in DFS
// order, edge (0,1) implicit, bit (s-2)(s+1)/2 + t for edge (s,t), s >= 2,
// t < s.  Canonical code layout (Manager.cpp:1134-1144): bit i(i-1)/2 + j for
// edge (i,j), i > j.
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>
#include <algorithm>

#include "internal.h"

namespace dm {

// cg_global column of the reference dictionaries (dictionaries/3.csv..6.csv,
// section "#cg_local,cg_global"), indexed by cg_local.
static const uint64_t kClassCodes3[3] = {
    4, 6, 7,
};
static const uint64_t kClassCodes4[10] = {
    32, 33, 48, 50, 51, 52, 56, 60, 62, 63,
};
static const uint64_t kClassCodes5[33] = {
    224, 424, 496, 504, 512, 516, 736, 768, 784, 787, 788, 800,
    801, 884, 896, 928, 929, 936, 937, 944, 947, 948, 960, 992,
    993, 1008, 1010, 1011, 1012, 1016, 1020, 1022, 1023,
};
static const uint64_t kClassCodes6[155] = {
    3872, 6656, 6720, 7680, 7952, 8104, 10592, 12032, 12928, 12931, 14560, 14728,
    15104, 15168, 15169, 15200, 15201, 15232, 15240, 15242, 15243, 15584, 15856, 16096,
    16128, 16256, 16288, 16289, 16320, 16352, 16353, 16370, 16384, 16416, 16417, 23040,
    23104, 23137, 23320, 24064, 24336, 24576, 24832, 24833, 24852, 24864, 24865, 25088,
    25092, 25095, 25872, 26116, 26976, 27050, 27424, 27425, 28072, 28322, 28416, 28448,
    28449, 28672, 28896, 28897, 29184, 29188, 29252, 29312, 29315, 29316, 29320, 29324,
    29327, 29440, 29441, 29460, 29461, 29472, 29473, 30212, 30436, 30480, 30484, 30548,
    30644, 30720, 30944, 31216, 31224, 31232, 31236, 31394, 31488, 31489, 31504, 31507,
    31508, 31520, 31521, 31552, 31553, 31568, 31570, 31571, 31576, 31580, 31583, 31584,
    31585, 31616, 31624, 31627, 31648, 31649, 31656, 31657, 31658, 31659, 31664, 31667,
    31668, 31744, 31968, 32168, 32240, 32248, 32256, 32260, 32512, 32528, 32531, 32532,
    32544, 32545, 32640, 32648, 32672, 32673, 32680, 32681, 32688, 32690, 32691, 32692,
    32704, 32736, 32737, 32752, 32754, 32755, 32756, 32760, 32764, 32766, 32767,
};
static inline int tri(int i, int j) { return i * (i - 1) / 2 + j; }  // i > j

static uint64_t permute_code(uint64_t code, int k, const int *p) {
    uint64_t out = 0;
    for (int i = 1; i < k; i++)
        for (int j = 0; j < i; j++)
            if ((code >> tri(i, j)) & 1) {
                int a = p[i], b = p[j];
                if (a < b) std::swap(a, b);
                out |= 1ull << tri(a, b);
            }
    return out;
}

static bool build_table(int k, PatternTable &t) {
    const int E = k * (k - 1) / 2;
    const uint64_t ncodes = 1ull << E;
    std::vector<int> perm(k);
    std::vector<std::vector<int>> perms;
    for (int i = 0; i < k; i++) perm[i] = i;
    do perms.push_back(perm);
    while (std::next_permutation(perm.begin(), perm.end()));

    // orbit id of every code (code 0, the empty graph, is never a pattern)
    std::vector<int32_t> orbit(ncodes, -1);
    std::vector<uint64_t> orbit_max;
    int norb = 0;
    for (uint64_t c = 1; c < ncodes; c++) {
        if (orbit[c] >= 0) continue;
        uint64_t mx = 0;
        for (const auto &p : perms) {
            uint64_t d = permute_code(c, k, p.data());
            orbit[d] = norb;
            mx = std::max(mx, d);
        }
        orbit_max.push_back(mx);
        norb++;
    }
    std::vector<int32_t> orbit_to_cls(norb, -1);
    t.cls_code.assign(norb, 0);
    const uint64_t *codes = nullptr;
    size_t ncodes_ref = 0;
    switch (k) {
        case 3: codes = kClassCodes3; ncodes_ref = 3; break;
        case 4: codes = kClassCodes4; ncodes_ref = 10; break;
        case 5: codes = kClassCodes5; ncodes_ref = 33; break;
        case 6: codes = kClassCodes6; ncodes_ref = 155; break;
        default: break;
    }
    if (codes) {
        if ((int)ncodes_ref != norb) return false;
        for (size_t i = 0; i < ncodes_ref; i++) {
            int o = orbit[codes[i]];
            if (o < 0 || orbit_to_cls[o] >= 0) return false;  // one code per class
            orbit_to_cls[o] = (int32_t)i;
            t.cls_code[i] = codes[i];
        }
    } else {
        std::vector<int> ord(norb);
        for (int i = 0; i < norb; i++) ord[i] = i;
        std::sort(ord.begin(), ord.end(),
                  [&](int a, int b) { return orbit_max[a] < orbit_max[b]; });
        for (int i = 0; i < norb; i++) {
            orbit_to_cls[ord[i]] = i;
            t.cls_code[i] = orbit_max[ord[i]];
        }
    }
    const int bits = (k + 1) * (k - 2) / 2;
    t.k = k;
    t.ncls = norb;
    t.q2c.assign(1ull << bits, 0);
    for (uint64_t q = 0; q < (1ull << bits); q++) {
        uint64_t code = 1ull << tri(1, 0);
        int b = 0;
        for (int s = 2; s < k; s++)
            for (int tt = 0; tt < s; tt++, b++)
                if ((q >> b) & 1) code |= 1ull << tri(s, tt);
        t.q2c[q] = (uint16_t)orbit_to_cls[orbit[code]];
    }
    return true;
}

const PatternTable *pattern_table(int k) {
    static std::mutex mu;
    static PatternTable tables[8];
    static bool built[8] = {false};
    if (k < 3 || k > 7) {
        set_error(DM_ERR_UNSUPPORTED, "motif size k=%d has no class table (3..7)", k);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!built[k]) {
        if (!build_table(k, tables[k])) {
            set_error(DM_ERR_UNSUPPORTED, "class table for k=%d does not match the reference ids", k);
            return nullptr;
        }
        built[k] = true;
    }
    return &tables[k];
}

}  // namespace dm

using namespace dm;

extern "C" int dm_motif_num_classes(int k, size_t *ncls) {
    if (!ncls) return set_error(DM_ERR_ARG, "ncls is NULL");
    const PatternTable *t = pattern_table(k);
    if (!t) return last_code();
    *ncls = (size_t)t->ncls;
    return DM_OK;
}

extern "C" int dm_pattern_table(int k, uint16_t *q2c, size_t nquick, uint64_t *codes,
                                size_t ncls) {
    const PatternTable *t = pattern_table(k);
    if (!t) return last_code();
    if (q2c) {
        if (nquick < t->q2c.size()) return set_error(DM_ERR_ARG, "quick table needs %zu entries", t->q2c.size());
        memcpy(q2c, t->q2c.data(), t->q2c.size() * sizeof(uint16_t));
    }
    if (codes) {
        if (ncls < (size_t)t->ncls) return set_error(DM_ERR_ARG, "class table needs %d entries", t->ncls);
        memcpy(codes, t->cls_code.data(), t->ncls * sizeof(uint64_t));
    }
    return DM_OK;
}
