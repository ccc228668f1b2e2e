/*
 * PairWiseSW_tal.hpp -- header-only C++ shim with the classes and member
 * signatures of TAL's full inter-task Smith-Waterman
 * (src/alignment/Smith-Waterman/PairWiseSW.h:83-105, PairWiseSW_8.h,
 * PairWiseSW_16.h, PairWiseSW_32.h), forwarding to the C ABI in tal_sw.h.
 * Code written against the reference (e.g. bench-smith-waterman.cpp:203-223)
 * compiles unchanged against this header and links against libtal_sw.so.
 *
 * SeqPair is the reference's 12-byte record (PairWiseSW.h:44-49); like the
 * reference headers, this one is not meant to share a translation unit with
 * the banded engine's shim, whose SeqPair differs.
 *
 * Differences a caller can observe: failures throw std::runtime_error; the
 * reference's getScores writes the padding entries of the last group
 * (pairArray[numPairs .. roundNumPairs), _16.cpp:247-252) and this one writes
 * nothing past pairArray[numPairs-1]; numThreads is accepted and ignored;
 * getTicks reports the device time of the last call (in nanoseconds, in
 * swTicks) instead of TSC ticks; the scalar smithWatermanOnePair routines and
 * the (compiled-out) pair sorts are not provided.
 */
#ifndef TAL_PAIRWISESW_TAL_HPP
#define TAL_PAIRWISESW_TAL_HPP

#include <stdint.h>

#include <stdexcept>
#include <string>

#include "tal_sw.h"

typedef tal_sw_pair SeqPair;

typedef struct cycle_breakup {
    int64_t sort1Ticks;
    int64_t setupTicks;
    int64_t swTicks;
    int64_t sort2Ticks;
} CycleBreakup;

#ifndef MAX_SEQ_LEN_8
#define MAX_SEQ_LEN_8 128
#endif
#ifndef MAX_SEQ_LEN_16
#define MAX_SEQ_LEN_16 1024
#endif
#ifndef MAX_SEQ_LEN_32
#define MAX_SEQ_LEN_32 1024
#endif

class PairWiseSW {
  public:
    PairWiseSW(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
               int device = 0)
    {
        const int rc = tal_sw_create(w_match, w_mismatch, w_open, w_extend, device, &h_);
        if (rc != TAL_OK) throw std::runtime_error("tal_sw_create failed: " + std::to_string(rc));
    }
    virtual ~PairWiseSW() { tal_sw_destroy(h_); }
    PairWiseSW(const PairWiseSW &) = delete;
    PairWiseSW &operator=(const PairWiseSW &) = delete;

    CycleBreakup getTicks()
    {
        CycleBreakup t = {0, 0, 0, 0};
        tal_sw_stats st;
        if (tal_sw_get_stats(h_, &st) == TAL_OK) t.swTicks = (int64_t)(st.ms_total * 1e6);
        return t;
    }

  protected:
    void scores(int bits, SeqPair *pairArray, uint8_t *seqBuf, int32_t numPairs)
    {
        /* the reference addresses seqBuf by id without an extent; the device
         * copy needs one: the end of the last sequence any pair names */
        const int64_t S = tal_sw_slot(bits);
        size_t len = 0;
        for (int32_t k = 0; k < numPairs; k++) {
            if (pairArray[k].id < 0) continue;
            const size_t e = (size_t)((2 * (int64_t)pairArray[k].id + 1) * S +
                                      (pairArray[k].len2 > 0 ? pairArray[k].len2 : 0));
            if (e > len) len = e;
        }
        const int rc = tal_sw_get_scores(h_, bits, pairArray, seqBuf, len, numPairs);
        if (rc != TAL_OK) throw std::runtime_error(std::string("tal_sw: ") + tal_sw_last_error(h_));
    }

  private:
    tal_sw *h_ = nullptr;
};

/* PairWiseSW_8.h:47-72; getScores :55 */
class PairWiseSW_8 : public PairWiseSW {
  public:
    PairWiseSW_8(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                 int device = 0)
        : PairWiseSW(w_match, w_mismatch, w_open, w_extend, device)
    {
    }
    void getScores(SeqPair *pairArray, uint8_t *seqBuf, int32_t numPairs, uint16_t numThreads)
    {
        (void)numThreads;
        scores(8, pairArray, seqBuf, numPairs);
    }
};

/* PairWiseSW_16.h:47-72; getScores :55 */
class PairWiseSW_16 : public PairWiseSW {
  public:
    PairWiseSW_16(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                  int device = 0)
        : PairWiseSW(w_match, w_mismatch, w_open, w_extend, device)
    {
    }
    void getScores(SeqPair *pairArray, uint8_t *seqBuf, int32_t numPairs, uint16_t numThreads)
    {
        (void)numThreads;
        scores(16, pairArray, seqBuf, numPairs);
    }
};

/* PairWiseSW_32.h:47-72; getScores :55 */
class PairWiseSW_32 : public PairWiseSW {
  public:
    PairWiseSW_32(int32_t w_match, int32_t w_mismatch, int32_t w_open, int32_t w_extend,
                  int device = 0)
        : PairWiseSW(w_match, w_mismatch, w_open, w_extend, device)
    {
    }
    void getScores(SeqPair *pairArray, uint8_t *seqBuf, int32_t numPairs, uint16_t numThreads)
    {
        (void)numThreads;
        scores(32, pairArray, seqBuf, numPairs);
    }
};

#endif
