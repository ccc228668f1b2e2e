/*
 * tal_bench_bsw.cpp -- bench-bsw on the MI355X engine.
 *
 * The command line, input format and chunking of TAL's benchmarks/
 * bench-bsw.cpp (parseCmdLine :66-115, loadPairs :117-169, the loop at
 * :232-263): pairs are read from a text pair file in chunks of 20,000
 * (MAX_NUM_PAIRS_ALLOC) into 256/128-byte slots and scored chunk by chunk on
 * one BandedPairWiseSW object -- here the engine's C++ shim, so this file is
 * also the reference benchmark's call pattern compiled against it.  With
 * -maxi each pair's outputs go to stderr in the format of the reference's
 * MAXI prints (bandedSWA.cpp:2676-2686): "score (tle qle) gscore max_off gtle".
 *
 *   tal-bench-bsw -pairs FILE [-match 1] [-mismatch 4] [-ambig -1] [-gapo 6]
 *                 [-gape 1] [-w 100] [-zdrop 100] [-end_bonus 5] [-auto]
 *                 [-scalar] [-maxi] [-ref_slot 256] [-qer_slot 128]
 *
 * -auto scores each chunk with the 8/16-bit dispatcher instead of
 * getScores16; -scalar with scalarBandedSWAWrapper, bench-bsw's branch for
 * builds without AVX2 (:256-258), whose matrix comes from bwa_fill_scmat of
 * -match/-mismatch/-ambig (:56-64, :213); -ref_slot/-qer_slot widen the
 * per-pair slots (a slot of S
 * bytes holds S-2 bases) for inputs longer than bench-bsw's.  A summary
 * (pairs, seconds, band cells, GCUPS) goes to stdout.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "bandedSWA_tal.hpp"

namespace {

constexpr int REF_SLOT = 256;    /* MAX_SEQ_LEN_REF, bandedSWA.h:47 */
constexpr int QER_SLOT = 128;    /* MAX_SEQ_LEN_QER, bandedSWA.h:48 */
constexpr int CHUNK = 20000;     /* MAX_NUM_PAIRS_ALLOC, bandedSWA.h:51 */
constexpr int MAX_SLOT = 32769;  /* a 32767-base sequence, its newline and NUL */

/* band cells of one pair, sum_i max(0, min(len2, i+b+1) - max(0, i-b)), with
 * b the pair's band as the 16-bit wrapper narrows it (bandedSWA.cpp:2882-2904);
 * the work unit of the GCUPS figure */
long long band_cells(int len1, int len2, int w, int match, int mismatch, int o_ins, int e_ins,
                     int o_del, int e_del, int end_bonus)
{
    int sm = 0;
    if ((int8_t)match > sm) sm = (int8_t)match;
    if ((int8_t)(-(int8_t)mismatch) > sm) sm = (int8_t)(-(int8_t)mismatch);
    const int ql = (len2 * sm) & 0xFFFF;
    int b = w;
    const int gaps[2][2] = {{o_ins, e_ins}, {o_del, e_del}};
    for (const auto &g : gaps) {
        const int t = (ql + end_bonus - g[0]) & 0xFFFF;
        int bb = t / g[1] + 1;
        if (bb < 1) bb = 1;
        b = (b < bb ? b : bb) & 0xFFFF;
    }
    long long c = 0;
    for (int i = 0; i < len1; i++) {
        const int lo = i - b > 0 ? i - b : 0;
        const int hi = i + b + 1 < len2 ? i + b + 1 : len2;
        if (hi > lo) c += hi - lo;
    }
    return c;
}

/* bench-bsw's scoring matrix (bwa_fill_scmat, bench-bsw.cpp:56-64): match on
 * the diagonal of codes 0-3, -mismatch off it, ambig in row and column 4 */
void fill_scmat(int a, int b, int ambig, int8_t mat[25])
{
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++)
            mat[i * 5 + j] = (int8_t)((i == 4 || j == 4) ? ambig : (i == j ? a : -b));
}

} // namespace

int main(int argc, char **argv)
{
    const char *path = nullptr;
    int match = 1, mismatch = 4, ambig = -1, gapo = 6, gape = 1, w = 100, zdrop = 100,
        end_bonus = 5;
    int ref_slot = REF_SLOT, qer_slot = QER_SLOT;
    bool maxi = false, autodispatch = false, scalar = false;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const bool has = i + 1 < argc;
        if (!strcmp(a, "-pairs") && has) path = argv[++i];
        else if (!strcmp(a, "-match") && has) match = atoi(argv[++i]);
        else if (!strcmp(a, "-mismatch") && has) mismatch = atoi(argv[++i]);
        else if (!strcmp(a, "-gapo") && has) gapo = atoi(argv[++i]);
        else if (!strcmp(a, "-gape") && has) gape = atoi(argv[++i]);
        else if (!strcmp(a, "-w") && has) w = atoi(argv[++i]);
        else if (!strcmp(a, "-zdrop") && has) zdrop = atoi(argv[++i]);
        else if (!strcmp(a, "-end_bonus") && has) end_bonus = atoi(argv[++i]);
        else if (!strcmp(a, "-ref_slot") && has) ref_slot = atoi(argv[++i]);
        else if (!strcmp(a, "-qer_slot") && has) qer_slot = atoi(argv[++i]);
        else if (!strcmp(a, "-maxi")) maxi = true;
        else if (!strcmp(a, "-auto")) autodispatch = true;
        else if (!strcmp(a, "-scalar")) scalar = true;
        else if (!strcmp(a, "-ambig") && has) ambig = atoi(argv[++i]);
        else if (!strcmp(a, "-h0")) i++; /* accepted, unused as in bench-bsw */
        else {
            fprintf(stderr, "unknown argument %s\n", a);
            return 2;
        }
    }
    if (!path) {
        printf("usage: %s -pairs <InSeqFile> [-match m] [-mismatch x] [-gapo o] [-gape e] "
               "[-w w] [-auto] [-scalar] [-maxi]\n", argv[0]);
        return 2;
    }
    if (ref_slot < 3 || ref_slot > MAX_SLOT || qer_slot < 3 || qer_slot > MAX_SLOT) {
        fprintf(stderr, "slots must be 3..%d bytes\n", MAX_SLOT);
        return 2;
    }
    tal_bsw_pairfile *pf = nullptr;
    if (tal_bsw_pairfile_open(path, &pf) != TAL_OK) {
        fprintf(stderr, "Could not open file: %s\n", path);
        return 1;
    }
    std::vector<SeqPair> pairs(CHUNK);
    std::vector<uint8_t> ref((size_t)ref_slot * CHUNK), qer((size_t)qer_slot * CHUNK);
    long long total = 0, cells = 0;
    double secs = 0;
    int rc = 0;
    try {
        int8_t mat[25];
        fill_scmat(match, mismatch, ambig, mat);
        BandedPairWiseSW bsw(gapo, gape, gapo, gape, zdrop, end_bonus, mat, (int8_t)match,
                             (int8_t)mismatch, 1);
        for (;;) {
            int32_t n = 0;
            const int r = tal_bsw_pairfile_read(pf, pairs.data(), ref.data(), ref_slot, qer.data(),
                                                qer_slot, CHUNK, &n);
            if (r != TAL_OK) {
                fprintf(stderr, "error reading %s (status %d)\n", path, r);
                rc = 1;
                break;
            }
            if (n == 0) break;
            const auto t0 = std::chrono::steady_clock::now();
            if (scalar)
                bsw.scalarBandedSWAWrapper(pairs.data(), ref.data(), qer.data(), n, 1, w);
            else if (autodispatch)
                bsw.getScoresAuto(pairs.data(), ref.data(), qer.data(), n, w);
            else
                bsw.getScores16(pairs.data(), ref.data(), qer.data(), n, 1, w);
            secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (int32_t k = 0; k < n; k++) {
                const SeqPair &p = pairs[k];
                cells += band_cells(p.len1, p.len2, w, match, mismatch, gapo, gape, gapo, gape,
                                    end_bonus);
                if (maxi)
                    fprintf(stderr, "%d (%d %d) %d %d %d\n", p.score, p.tle, p.qle, p.gscore,
                            p.max_off, p.gtle);
            }
            total += n;
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        rc = 1;
    }
    tal_bsw_pairfile_close(pf);
    printf("Total pairs processed: %lld\n", total);
    printf("Engine time (host-buffer calls) = %.6f s\n", secs);
    printf("SW band cells = %lld\n", cells);
    printf("SW GCUPS = %.3f\n", secs > 0 ? cells / secs / 1e9 : 0.0);
    return rc;
}
