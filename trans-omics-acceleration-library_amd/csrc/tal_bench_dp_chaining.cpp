/*
 * tal_bench_dp_chaining.cpp -- bench-dp-chaining on the MI355X engine.
 *
 * The input format and output of TAL's benchmarks/bench-dp-chaining.cpp
 * (main :43-116): an anchors file holds sets one after another, each a count
 * followed by that many "r q l" triples (whitespace-separated; r unsigned
 * 64-bit, q and l 32-bit); every set is chained with dp_chain(5000, 5000,
 * 500, 25, 5000, 1.0, 0, 1) and dr = dq = 5000, as :67 and :103 do.  The
 * reference chains set by set; here all sets are one tal_chain_dp batch.
 *
 * With -debug, prints "f p" per anchor, set after set, as the reference's
 * DEBUG build does (:105-109).  stderr gets the set and anchor counts and the
 * device time (the reference prints rdtsc ticks there).  -parse-only reads
 * the file and prints the counts without touching the GPU.
 *
 *   tal-bench-dp-chaining ANCHORS_FILE [-debug] [-parse-only]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "tal_chain.h"

/* The next whitespace-separated token as an unsigned 64-bit (r) or signed
 * 32-bit (count, q, l) integer; false at the end of the file. */
static bool next_u64(FILE *f, unsigned long long *out) { return fscanf(f, "%llu", out) == 1; }
static bool next_i32(FILE *f, int *out) { return fscanf(f, "%d", out) == 1; }

int main(int argc, char **argv)
{
    const char *path = nullptr;
    bool debug = false, parse_only = false;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-debug")) debug = true;
        else if (!strcmp(argv[i], "-parse-only")) parse_only = true;
        else path = argv[i];
    }
    if (!path) {
        fprintf(stderr, "Input parameters: 1. Anchors_file_name\n");
        return 2;
    }
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "Could not open file: %s\n", path);
        return 1;
    }
    std::vector<tal_chain_anchor> anchors;
    std::vector<int64_t> offsets(1, 0);
    int count = 0;
    while (next_i32(f, &count)) { /* while (fin >> id) */
        if (count < 0) {
            fprintf(stderr, "bad set size %d in %s\n", count, path);
            return 1;
        }
        for (int k = 0; k < count; k++) {
            unsigned long long r;
            int q, l;
            if (!next_u64(f, &r) || !next_i32(f, &q) || !next_i32(f, &l)) {
                fprintf(stderr, "truncated set %zu in %s\n", offsets.size() - 1, path);
                return 1;
            }
            anchors.push_back(tal_chain_anchor{(uint64_t)r, (int32_t)q, (int32_t)l});
        }
        offsets.push_back((int64_t)anchors.size());
    }
    fclose(f);
    const int32_t nsets = (int32_t)(offsets.size() - 1);
    const size_t n = anchors.size();
    fprintf(stderr, "%d sets, %zu anchors\n", nsets, n);
    if (parse_only) return 0;

    tal_chain_params prm = {5000, 5000, 500, 25, 5000, 1.0f, 0, 1}; /* :67 */
    tal_chain *h = nullptr;
    int rc = tal_chain_create(&prm, 0, &h);
    if (rc != TAL_OK) {
        fprintf(stderr, "tal_chain_create failed: %d\n", rc);
        return 1;
    }
    std::vector<uint32_t> fo(n > 0 ? n : 1);
    std::vector<int32_t> po(n > 0 ? n : 1), vo(n > 0 ? n : 1);
    const auto t0 = std::chrono::steady_clock::now();
    rc = tal_chain_dp(h, anchors.data(), offsets.data(), nsets, 5000, 5000, fo.data(), po.data(),
                      vo.data());
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != TAL_OK) {
        fprintf(stderr, "tal_chain_dp failed: %s\n", tal_chain_last_error(h));
        tal_chain_destroy(h);
        return 1;
    }
    if (debug)
        for (size_t i = 0; i < n; i++) printf("%u %d\n", fo[i], po[i]);
    tal_chain_stats st;
    tal_chain_get_stats(h, &st);
    fprintf(stderr, "device %.3f ms (kernel %.3f ms), call %.3f ms; %.1f M anchors/s on the device\n",
            st.ms_total, st.ms_kernel, wall * 1e3,
            st.ms_total > 0 ? n / (st.ms_total * 1e3) : 0.0);
    tal_chain_destroy(h);
    return 0;
}
