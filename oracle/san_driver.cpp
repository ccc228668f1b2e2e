/*
 * san_driver.cpp -- runs the CPU code of this repository under AddressSanitizer
 * and UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`): the
 * banded-SW restatement (getScores16/8, the scalar routine, the band DP), its
 * vectorised port, the full-SW and chaining restatements, the synthetic
 * workload generator, and the pair-file and packed-file readers and writers
 * of the product library (csrc/bsw_pairfile.cpp, host code only).
 *
 * TEST INFRASTRUCTURE ONLY (tests/test_sanitize_cpu.py runs it).  Exit status
 * 0 means every check passed and no sanitizer fired (the build makes every
 * sanitizer report fatal).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "bsw_oracle.h"
#include "bsw_synth.h"
#include "chain_oracle.h"
#include "sw_oracle.h"
#include "tal_bsw.h"

extern "C" int bsw_simd_run_mt(int o_del, int e_del, int o_ins, int e_ins, int zdrop,
                               int end_bonus, int8_t w_match, int8_t w_mismatch, int bits,
                               oracle_seqpair *pairs, const uint8_t *ref, const uint8_t *qer,
                               int32_t n, int32_t w, int32_t chunk, int nthreads);

static int failures = 0;

#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                                  \
            fprintf(stderr, __VA_ARGS__);                                                          \
            fprintf(stderr, "\n");                                                                 \
            failures++;                                                                            \
        }                                                                                          \
    } while (0)

struct Batch {
    std::vector<oracle_seqpair> pairs;
    std::vector<uint8_t> ref, qer;
    std::vector<int32_t> w;
};

/* Pairs [first, first + n) of synthetic workload cfg, laid out back to back
 * like the Python generator does. */
static Batch make(int cfg, uint64_t seed, int64_t first, int32_t n)
{
    Batch b;
    std::vector<int32_t> len1(n), len2(n), h0(n);
    b.w.resize(n);
    CHECK(tal_synth_shape(cfg, seed, first, n, len1.data(), len2.data(), h0.data(),
                          b.w.data()) == 0,
          "shape of workload %d", cfg);
    std::vector<int64_t> roff(n), qoff(n);
    int64_t r = 0, q = 0;
    for (int32_t k = 0; k < n; k++) {
        roff[k] = r;
        qoff[k] = q;
        r += len1[k];
        q += len2[k];
    }
    b.ref.assign(r > 0 ? r : 1, 0);
    b.qer.assign(q > 0 ? q : 1, 0);
    b.pairs.assign(n, oracle_seqpair{});
    CHECK(tal_synth_fill(cfg, seed, first, n, roff.data(), qoff.data(), b.pairs.data(),
                         b.ref.data(), b.qer.data(), 2) == 0,
          "fill of workload %d", cfg);
    return b;
}

static bool same_outputs(const oracle_seqpair &a, const oracle_seqpair &b)
{
    return a.score == b.score && a.tle == b.tle && a.gtle == b.gtle && a.qle == b.qle &&
           a.gscore == b.gscore && a.max_off == b.max_off;
}

/* getScores16 on the restatement and on its vectorised port agree pair for
 * pair; ragged sizes exercise the padded last group. */
static void banded(int cfg, int32_t n, int32_t w)
{
    Batch b = make(cfg, 0xB5A0000000000000ull + cfg, 0, n);
    bsw_oracle *o = bsw_oracle_create(6, 1, 6, 1, 100, 5, 1, 4);
    CHECK(o != nullptr, "oracle create");
    std::vector<oracle_seqpair> a = b.pairs, v = b.pairs;
    bsw_oracle_get_scores16(o, a.data(), b.ref.data(), b.qer.data(), n, w);
    CHECK(bsw_simd_run_mt(6, 1, 6, 1, 100, 5, 1, 4, 16, v.data(), b.ref.data(), b.qer.data(), n,
                          w, 1000, 2) == 0,
          "simd run");
    for (int32_t k = 0; k < n; k++)
        CHECK(same_outputs(a[k], v[k]), "workload %d pair %d: port differs from restatement",
              cfg, k);
    /* 8-bit on the pairs short enough for it */
    std::vector<oracle_seqpair> p8;
    for (const oracle_seqpair &p : b.pairs)
        if (p.len1 <= 127 && p.len2 <= 127) p8.push_back(p);
    if (!p8.empty()) {
        std::vector<oracle_seqpair> s8 = p8;
        bsw_oracle_get_scores8(o, p8.data(), b.ref.data(), b.qer.data(), (int32_t)p8.size(), w);
        CHECK(bsw_simd_run_mt(6, 1, 6, 1, 100, 5, 1, 4, 8, s8.data(), b.ref.data(),
                              b.qer.data(), (int32_t)s8.size(), w, 1000, 2) == 0,
              "simd run 8");
        for (size_t k = 0; k < p8.size(); k++)
            CHECK(same_outputs(p8[k], s8[k]), "workload %d 8-bit pair %zu differs", cfg, k);
    }
    /* the scalar routine and the band DP on the same inputs */
    std::vector<oracle_seqpair> sc = b.pairs;
    bool codes_ok = true;
    for (uint8_t c : b.ref) codes_ok = codes_ok && c <= 4;
    for (uint8_t c : b.qer) codes_ok = codes_ok && c <= 4;
    if (codes_ok) {
        bsw_oracle_scalar(o, sc.data(), b.ref.data(), b.qer.data(), n, w);
        int8_t mat[25];
        for (int i = 0; i < 25; i++) mat[i] = (i / 5 == 4 || i % 5 == 4) ? -1 : (i / 5 == i % 5 ? 1 : -4);
        std::vector<oracle_seqpair> sm = b.pairs;
        bsw_oracle_scalar_mat(o, mat, sm.data(), b.ref.data(), b.qer.data(), n, w);
        for (int32_t k = 0; k < n; k++)
            CHECK(same_outputs(sc[k], sm[k]), "scalar with the default matrix differs at %d", k);
    }
    bsw_oracle_destroy(o);
}

static void full_sw()
{
    const int slot = 1024;
    const int32_t n = 37;
    std::vector<uint8_t> buf((size_t)2 * n * slot, 0);
    std::vector<sw_oracle_pair> p(n), q(n);
    uint64_t x = 12345;
    for (int32_t k = 0; k < n; k++) {
        p[k].id = k;
        p[k].len1 = (int16_t)(50 + k * 7 % 300);
        p[k].len2 = (int16_t)(40 + k * 11 % 300);
        for (int i = 0; i < slot; i++) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            buf[(size_t)2 * k * slot + i] = i < p[k].len1 ? "ACGT"[x >> 62] : 0;
            buf[(size_t)(2 * k + 1) * slot + i] = i < p[k].len2 ? "ACGT"[(x >> 60) & 3] : 0;
        }
    }
    q = p;
    CHECK(sw_oracle_get_scores(16, 1, -4, -6, -1, p.data(), buf.data(), n, 2) == 0, "sw 16");
    sw_oracle_scalar(1, -4, -6, -1, q.data(), buf.data(), n, slot);
    for (int32_t k = 0; k < n; k++)
        CHECK(p[k].score == q[k].score, "full SW pair %d: %d vs %d", k, p[k].score, q[k].score);
}

static void chaining()
{
    const int32_t nsets = 5;
    std::vector<int64_t> off(nsets + 1, 0);
    std::vector<chain_oracle_anchor> a;
    uint64_t x = 99;
    for (int32_t s = 0; s < nsets; s++) {
        const int m = 1 + s * 40;
        uint64_t r = 1000;
        int32_t q = 0;
        for (int i = 0; i < m; i++) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            r += 1 + (x >> 58);
            q += 1 + (int32_t)((x >> 40) & 15);
            a.push_back(chain_oracle_anchor{r, q, 15});
        }
        off[s + 1] = (int64_t)a.size();
    }
    const size_t n = a.size();
    std::vector<uint32_t> f(n), f1(n);
    std::vector<int32_t> p(n), v(n), p1(n), v1(n);
    chain_oracle_dp_sets(a.data(), off.data(), nsets, 500, 5000, 5000, 5000, f.data(), p.data(),
                         v.data(), 2);
    for (int32_t s = 0; s < nsets; s++) {
        const int64_t b = off[s], m = off[s + 1] - off[s];
        chain_oracle_dp(a.data() + b, m, 500, 5000, 5000, 5000, f1.data() + b, p1.data() + b,
                        v1.data() + b);
    }
    for (size_t k = 0; k < n; k++)
        CHECK(f[k] == f1[k] && p[k] == p1[k] && v[k] == v1[k], "chaining anchor %zu", k);
}

static void files(const char *dir)
{
    Batch b = make(1, 0xB5A0000000000001ull, 0, 300);
    const int32_t n = (int32_t)b.pairs.size();
    std::vector<tal_seqpair> tp(n);
    for (int32_t k = 0; k < n; k++) memcpy(&tp[k], &b.pairs[k], sizeof(tal_seqpair));
    const std::string txt = std::string(dir) + "/pairs.txt";
    CHECK(tal_bsw_pairfile_write(txt.c_str(), tp.data(), b.ref.data(), b.qer.data(), n) == TAL_OK,
          "pair-file write");
    tal_bsw_pairfile *f = nullptr;
    CHECK(tal_bsw_pairfile_open(txt.c_str(), &f) == TAL_OK, "pair-file open");
    const size_t rs = 256, qs = 256;
    std::vector<tal_seqpair> got(128);
    std::vector<uint8_t> r(128 * rs), q(128 * qs);
    int32_t total = 0, m = 0;
    do {
        CHECK(tal_bsw_pairfile_read(f, got.data(), r.data(), rs, q.data(), qs, 128, &m) == TAL_OK,
              "pair-file read");
        for (int32_t k = 0; k < m; k++) {
            const tal_seqpair &o = tp[total + k];
            CHECK(got[k].len1 == o.len1 && got[k].len2 == o.len2 && got[k].h0 == o.h0,
                  "pair %d read back differently", total + k);
            CHECK(memcmp(r.data() + k * rs, b.ref.data() + o.idr, o.len1) == 0 &&
                      memcmp(q.data() + k * qs, b.qer.data() + o.idq, o.len2) == 0,
                  "pair %d sequences read back differently", total + k);
        }
        total += m;
    } while (m > 0);
    tal_bsw_pairfile_close(f);
    CHECK(total == n, "read %d of %d pairs", total, n);

    const std::string pk = std::string(dir) + "/pairs.pack";
    CHECK(tal_bsw_pack_write(pk.c_str(), tp.data(), n, b.ref.data(), b.ref.size(), b.qer.data(),
                             b.qer.size()) == TAL_OK,
          "pack write");
    tal_bsw_pack mp{};
    CHECK(tal_bsw_pack_map(pk.c_str(), &mp) == TAL_OK, "pack map");
    CHECK(mp.n == n && mp.ref_len == b.ref.size() && mp.qer_len == b.qer.size(), "pack sizes");
    if (mp.n == n) {
        CHECK(memcmp(mp.pairs, tp.data(), sizeof(tal_seqpair) * n) == 0, "pack records");
        mp.pairs[0].score = 7; /* a private mapping takes writes */
    }
    tal_bsw_pack_unmap(&mp);
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    /* ragged sizes around the group widths, on each workload's generator */
    for (int cfg : {1, 2, 3, 5})
        for (int32_t n : {1, 31, 33, 64, 65, 500}) banded(cfg, n, cfg == 3 ? 200 : 100);
    full_sw();
    chaining();
    files(dir);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("sanitized run: all checks passed\n");
    return 0;
}
