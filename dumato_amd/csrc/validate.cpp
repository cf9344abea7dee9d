// Result validation against Peregrine output files, the reference's
// DuMatoCPU::validateAggregateCounter / validateAggregatePattern
// (src/main/DuMatoCPU.cu:564-661).
//
// File name: "<graph path>.<k>.aggregateCounter" or ".aggregatePattern"
// (src/main/DuMatoCPU.cu:565-569, 599-603).  Only lines that contain '[' are
// read; the value is the text after the first ':' up to the next ':' or the
// end of the line, read as a decimal integer (src/main/DuMatoCPU.cu:580-588,
// 612-620).  Counter files: the values are summed and compared with the
// count.  Pattern files: positive values form a multiset; the check compares
// the totals and the sorted multisets of the positive histogram slots, so
// class ids need not agree between the two programs (Peregrine numbers its
// patterns its own way).
//
// One difference on purpose: values are read as 64-bit integers, where the
// reference's counter check uses atoi (32-bit) and wraps above 2^31-1.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dumato.h"
#include "internal.h"

using dm::set_error;

namespace {

// Reads every '['-line value of path into vals.  Returns DM_OK or DM_ERR_IO.
int read_values(const std::string &path, std::vector<long long> &vals) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return set_error(DM_ERR_IO, "cannot open validation file %s", path.c_str());
    std::string line;
    int c;
    bool eof = false;
    while (!eof) {
        line.clear();
        while ((c = fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        if (c == EOF) eof = true;
        if (line.find('[') == std::string::npos) continue;
        size_t colon = line.find(':');
        long long v = 0;
        if (colon != std::string::npos) {
            size_t end = line.find(':', colon + 1);
            std::string field = line.substr(colon + 1, end == std::string::npos ? std::string::npos : end - colon - 1);
            v = strtoll(field.c_str(), nullptr, 10);  // leading blanks skipped, junk -> 0 (as atoi)
        }
        vals.push_back(v);
    }
    fclose(f);
    return DM_OK;
}

std::string file_for(const char *graph_path, int k, const char *suffix) {
    return std::string(graph_path) + "." + std::to_string(k) + "." + suffix;
}

}  // namespace

extern "C" int dm_validate_counter(const char *graph_path, int k, uint64_t result, int *match) {
    if (!graph_path || !match) return set_error(DM_ERR_ARG, "dm_validate_counter: NULL argument");
    std::vector<long long> vals;
    int rc = read_values(file_for(graph_path, k, "aggregateCounter"), vals);
    if (rc != DM_OK) return rc;
    unsigned long long sum = 0;
    for (long long v : vals) sum += (unsigned long long)v;
    *match = sum == result ? 1 : 0;
    return DM_OK;
}

extern "C" int dm_validate_pattern(const char *graph_path, int k, const uint64_t *hist, size_t hist_len,
                                   int *count_match, int *pattern_match, uint64_t *file_total,
                                   size_t *file_patterns) {
    if (!graph_path || (!hist && hist_len) || !count_match || !pattern_match)
        return set_error(DM_ERR_ARG, "dm_validate_pattern: NULL argument");
    std::vector<long long> vals;
    int rc = read_values(file_for(graph_path, k, "aggregatePattern"), vals);
    if (rc != DM_OK) return rc;
    std::vector<unsigned long long> theirs, ours;
    unsigned long long tt = 0, to = 0;
    for (long long v : vals)
        if (v > 0) {
            theirs.push_back((unsigned long long)v);
            tt += (unsigned long long)v;
        }
    for (size_t i = 0; i < hist_len; i++)
        if (hist[i] > 0) {
            ours.push_back(hist[i]);
            to += hist[i];
        }
    std::sort(theirs.begin(), theirs.end());
    std::sort(ours.begin(), ours.end());
    *count_match = tt == to ? 1 : 0;
    *pattern_match = theirs == ours ? 1 : 0;
    if (file_total) *file_total = tt;
    if (file_patterns) *file_patterns = theirs.size();
    return DM_OK;
}
