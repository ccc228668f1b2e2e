/*
 * tal_bench_sw.cpp -- bench-smith-waterman on the MI355X engine.
 *
 * The command line, input format and output of TAL's
 * benchmarks/bench-smith-waterman.cpp (parseCmdLine :73-115, loadPairs
 * :117-157, main :160-248): a pair file holds two lines per pair (seq1, then
 * seq2, raw bytes); each sequence goes to its own MAX_SEQ_LEN slot of one
 * buffer, pair k at 2k and 2k+1; all pairs are scored with one getScores
 * call of the chosen precision through the PairWiseSW shim, which is the
 * reference benchmark's call pattern compiled against this engine.  Scores are
 * printed one per line as the reference prints them, then the cell count and
 * GCUPS (device time of the call).
 *
 *   tal-bench-sw -pairs FILE [-match 1] [-mismatch -1] [-gapo 1] [-gape 1]
 *                [-precision 32]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "PairWiseSW_tal.hpp"

int main(int argc, char **argv)
{
    int32_t w_match = 1, w_mismatch = -1, w_open = 1, w_extend = 1, precision = 32;
    const char *path = nullptr;
    for (int i = 1; i + 1 < argc; i += 2) { /* option/value pairs, as :80 */
        if (!strcmp(argv[i], "-match")) w_match = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "-mismatch")) w_mismatch = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "-gapo")) w_open = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "-gape")) w_extend = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "-precision")) precision = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "-pairs")) path = argv[i + 1];
    }
    if (!path) {
        printf("ERROR! pairFileName not specified.\n");
        return 2;
    }
    const int32_t S = tal_sw_slot(precision);
    if (S == 0) {
        printf("Error! precision value of %d not supported.\n", precision);
        return 2;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "Could not open file: %s\n", path);
        return 1;
    }
    /* read every line; pair k = lines 2k, 2k+1 */
    std::vector<std::string> lines;
    {
        std::string cur;
        int c;
        bool any = false;
        while ((c = fgetc(f)) != EOF) {
            any = true;
            if (c == '\n') {
                lines.push_back(cur);
                cur.clear();
                any = false;
            } else {
                cur.push_back((char)c);
            }
        }
        if (any) lines.push_back(cur);
        fclose(f);
    }
    if (lines.size() % 2) printf("ERROR! Odd number of sequences in %s\n", path);
    const int32_t n = (int32_t)(lines.size() / 2);
    std::vector<SeqPair> pairs(n > 0 ? n : 1);
    std::vector<uint8_t> buf((size_t)2 * S * (n > 0 ? n : 1), 0);
    int64_t cells = 0;
    for (int32_t k = 0; k < n; k++) {
        const std::string &a = lines[2 * (size_t)k], &b = lines[2 * (size_t)k + 1];
        const size_t limit = precision == 8 ? 127 : (size_t)S;
        if (a.size() > limit || b.size() > limit) {
            fprintf(stderr, "pair %d: a sequence is longer than %zu\n", k, limit);
            return 1;
        }
        memcpy(&buf[2 * (size_t)k * S], a.data(), a.size());
        memcpy(&buf[(2 * (size_t)k + 1) * S], b.data(), b.size());
        pairs[k].id = k;
        pairs[k].len1 = (int16_t)a.size();
        pairs[k].len2 = (int16_t)b.size();
        pairs[k].score = 0;
        cells += (int64_t)a.size() * (int64_t)b.size();
    }
    CycleBreakup ticks = {0, 0, 0, 0};
    double wall = 0;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        if (precision == 32) {
            PairWiseSW_32 sw(w_match, w_mismatch, w_open, w_extend);
            sw.getScores(pairs.data(), buf.data(), n, 1);
            ticks = sw.getTicks();
        } else if (precision == 16) {
            PairWiseSW_16 sw(w_match, w_mismatch, w_open, w_extend);
            sw.getScores(pairs.data(), buf.data(), n, 1);
            ticks = sw.getTicks();
        } else {
            PairWiseSW_8 sw(w_match, w_mismatch, w_open, w_extend);
            sw.getScores(pairs.data(), buf.data(), n, 1);
            ticks = sw.getTicks();
        }
        wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    for (int32_t k = 0; k < n; k++) printf("%d\n", pairs[k].score);
    const double dev_s = ticks.swTicks / 1e9;
    printf("SW device time = %.6f s (call wall time %.6f s)\n", dev_s, wall);
    printf("SW cells  = %lld\n", (long long)cells);
    printf("SW GCUPS  = %lf\n", dev_s > 0 ? cells / dev_s / 1e9 : 0.0);
    return 0;
}
