/*
 * ref_sw_bind.cpp -- C entry point over the reference's own full-SW classes
 * (PairWiseSW_{8,16,32}, src/alignment/Smith-Waterman/), compiled from the
 * reference sources by `make -C oracle ref` into oracle/_ref/libpwsw_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: the tests use it to check the oracle's
 * restatement against the reference itself.  It is built with AVX-512BW, the
 * reference's SIMD path, and refuses to run on a CPU without it.
 */
#include <stdint.h>

#include "PairWiseSW_16.h"
#include "PairWiseSW_32.h"
#include "PairWiseSW_8.h"

extern "C" int ref_sw_get_scores(int bits, int32_t w_match, int32_t w_mismatch, int32_t w_open,
                                 int32_t w_extend, SeqPair *pairs, uint8_t *seq_buf,
                                 int32_t n, int nthreads)
{
    const uint16_t nt = (uint16_t)(nthreads < 1 ? 1 : nthreads);
    if (!__builtin_cpu_supports("avx512bw")) return -2;
    /* getScores writes the padding entries of the last group, so `pairs`
     * must hold n rounded up to the group size.  With more than one thread,
     * a group bound of 1024 columns makes its scratch initialisation write
     * one row into the next thread's scratch (_16.cpp:170 with the buffers of
     * :140-141), so multi-threaded calls keep len2 <= 1020. */
    switch (bits) {
    case 8: {
        PairWiseSW_8 o(w_match, w_mismatch, w_open, w_extend);
        o.getScores(pairs, seq_buf, n, nt);
        return 0;
    }
    case 16: {
        PairWiseSW_16 o(w_match, w_mismatch, w_open, w_extend);
        o.getScores(pairs, seq_buf, n, nt);
        return 0;
    }
    case 32: {
        PairWiseSW_32 o(w_match, w_mismatch, w_open, w_extend);
        o.getScores(pairs, seq_buf, n, nt);
        return 0;
    }
    }
    return -1;
}
