/*
 * bandedSWA_tal.hpp -- header-only C++ shim with the member signatures of
 * TAL's BandedPairWiseSW (src/alignment/BSW/bandedSWA.h:113-172), forwarding
 * to the C ABI in tal_bsw.h.  Code written against the reference (e.g.
 * benchmarks/bench-bsw.cpp:215-218,253) compiles unchanged against this
 * header and links against libtal_bsw.so.
 *
 * Differences a caller can observe: failures throw std::runtime_error instead
 * of exit(EXIT_FAILURE) (bandedSWA.cpp:88-93); nothing is written past
 * pairArray[numPairs-1]; numThreads is accepted and ignored (the reference
 * only sizes per-thread scratch with it).  Like the reference, the object
 * keeps the constructor's mat_ pointer and its scalar path reads the matrix
 * at call time.
 */
#ifndef TAL_BANDEDSWA_TAL_HPP
#define TAL_BANDEDSWA_TAL_HPP

#include <stdint.h>

#include <stdexcept>
#include <string>

#include "tal_bsw.h"

typedef tal_seqpair SeqPair;

#ifndef MAX_SEQ_LEN8
#define MAX_SEQ_LEN8 128
#endif
#ifndef MAX_SEQ_LEN16
#define MAX_SEQ_LEN16 32768
#endif

class BandedPairWiseSW {
  public:
    uint64_t SW_cells = 0;

    BandedPairWiseSW(const int o_del, const int e_del, const int o_ins, const int e_ins,
                     const int zdrop, const int end_bonus, const int8_t *mat_,
                     const int8_t w_match, const int8_t w_mismatch, int numThreads,
                     int device = 0)
        : mat_ptr_(mat_)
    {
        (void)numThreads;
        tal_bsw_params p;
        p.o_del = o_del;
        p.e_del = e_del;
        p.o_ins = o_ins;
        p.e_ins = e_ins;
        p.zdrop = zdrop;
        p.end_bonus = end_bonus;
        p.w_match = w_match;
        p.w_mismatch = w_mismatch;
        int rc = tal_bsw_create(&p, device, &h_);
        if (rc != TAL_OK) throw std::runtime_error("tal_bsw_create failed: " + std::to_string(rc));
    }
    ~BandedPairWiseSW() { tal_bsw_destroy(h_); }
    BandedPairWiseSW(const BandedPairWiseSW &) = delete;
    BandedPairWiseSW &operator=(const BandedPairWiseSW &) = delete;

    void getScores16(SeqPair *pairArray, uint8_t *seqBufRef, uint8_t *seqBufQer,
                     int32_t numPairs, uint16_t numThreads, int32_t w)
    {
        (void)numThreads;
        size_t rl = 0, ql = 0;
        extents(pairArray, numPairs, &rl, &ql);
        check(tal_bsw_get_scores16(h_, pairArray, seqBufRef, rl, seqBufQer, ql, numPairs, w));
    }

    void getScores8(SeqPair *pairArray, uint8_t *seqBufRef, uint8_t *seqBufQer,
                    int32_t numPairs, uint16_t numThreads, int32_t w)
    {
        (void)numThreads;
        size_t rl = 0, ql = 0;
        extents(pairArray, numPairs, &rl, &ql);
        check(tal_bsw_get_scores8(h_, pairArray, seqBufRef, rl, seqBufQer, ql, numPairs, w));
    }

    /* Not in the reference: the 8/16-bit dispatcher of tal_bsw.h. */
    void getScoresAuto(SeqPair *pairArray, uint8_t *seqBufRef, uint8_t *seqBufQer,
                       int32_t numPairs, int32_t w)
    {
        size_t rl = 0, ql = 0;
        extents(pairArray, numPairs, &rl, &ql);
        check(tal_bsw_get_scores_auto(h_, pairArray, seqBufRef, rl, seqBufQer, ql, numPairs, w));
    }

    /* scalarBandedSWA (bandedSWA.h:124-129, bandedSWA.cpp:116-237): one pair,
     * scored on the GPU by the per-pair kernel; returns the score. */
    int scalarBandedSWA(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                        int32_t w, int h0, int *_qle, int *_tle, int *_gtle, int *_gscore,
                        int *_max_off)
    {
        SeqPair p{};
        p.idr = 0;
        p.idq = 0;
        p.len1 = tlen;
        p.len2 = qlen;
        p.h0 = h0;
        check(tal_bsw_scalar_scores(h_, mat_ptr_, &p, target, (size_t)(tlen > 0 ? tlen : 0), query,
                                    (size_t)(qlen > 0 ? qlen : 0), 1, w));
        if (_qle) *_qle = p.qle;
        if (_tle) *_tle = p.tle;
        if (_gtle) *_gtle = p.gtle;
        if (_gscore) *_gscore = p.gscore;
        if (_max_off) *_max_off = p.max_off;
        return p.score;
    }

    /* scalarBandedSWAWrapper (bandedSWA.h:131-136, bandedSWA.cpp:242-260). */
    void scalarBandedSWAWrapper(SeqPair *seqPairArray, uint8_t *seqBufRef, uint8_t *seqBufQer,
                                int numPairs, int nthreads, int32_t w)
    {
        (void)nthreads;
        size_t rl = 0, ql = 0;
        extents(seqPairArray, numPairs, &rl, &ql);
        check(tal_bsw_scalar_scores(h_, mat_ptr_, seqPairArray, seqBufRef, rl, seqBufQer, ql,
                                    numPairs, w));
    }

  private:
    tal_bsw *h_ = nullptr;
    const int8_t *mat_ptr_ = nullptr;

    /* The reference addresses the buffers without extents; the C ABI copies
     * them to the device, so the shim derives max(idr+len1), max(idq+len2). */
    static void extents(const SeqPair *p, int32_t n, size_t *rl, size_t *ql)
    {
        for (int32_t k = 0; k < n; k++) {
            if (p[k].idr >= 0 && p[k].len1 >= 0) {
                size_t e = (size_t)p[k].idr + (size_t)p[k].len1;
                if (e > *rl) *rl = e;
            }
            if (p[k].idq >= 0 && p[k].len2 >= 0) {
                size_t e = (size_t)p[k].idq + (size_t)p[k].len2;
                if (e > *ql) *ql = e;
            }
        }
    }
    void check(int rc)
    {
        if (rc != TAL_OK) throw std::runtime_error(std::string("tal_bsw: ") + tal_bsw_last_error(h_));
    }
};

#endif
