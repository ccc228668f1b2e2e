/*
 * parallel_chaining_tal.hpp -- header-only C++ shim with the class and member
 * signatures of TAL's DP chaining (src/chaining/parallel_chaining_32_bit.h:
 * 56-118, 216-217), forwarding to the C ABI in tal_chain.h.  Code written
 * against the reference (e.g. bench-dp-chaining.cpp:69-101) compiles
 * unchanged against this header and links against libtal_chain.so.
 *
 * Differences a caller can observe: failures throw std::runtime_error;
 * mm_dp_vectorized reads the anchors from `anchors` (the SoA copies are
 * accepted and not read) and ignores gap_cost/meta_data, as the reference's
 * vector path does; each call is one batch of one set -- the C ABI scores
 * many sets per call (tal_chain_dp), which is how the GPU is kept busy.
 */
#ifndef TAL_PARALLEL_CHAINING_TAL_HPP
#define TAL_PARALLEL_CHAINING_TAL_HPP

#include <stdint.h>
#include <stdlib.h>

#include <stdexcept>
#include <string>

#include "tal_chain.h"

class anchor_t {
  public:
    uint64_t r;
    int32_t q;
    int32_t l;
    anchor_t() {}
    anchor_t(uint64_t x, int32_t y, int32_t length) : r(x), q(y), l(length) {}
};
static_assert(sizeof(anchor_t) == sizeof(tal_chain_anchor), "anchor_t layout");

typedef uint32_t num_bits_t;

/* create_SoA_Anchors_32_bit (parallel_chaining_32_bit.h:69-87): the 32-bit
 * SoA copies, each with 16 entries of room before index 0. */
inline void create_SoA_Anchors_32_bit(anchor_t *anc, num_bits_t anc_size, num_bits_t *&anchor_r,
                                      num_bits_t *&anchor_q, num_bits_t *&anchor_l)
{
    anchor_r = (num_bits_t *)malloc((16 + anc_size) * sizeof(num_bits_t)) + 16;
    anchor_q = (num_bits_t *)malloc((16 + anc_size) * sizeof(num_bits_t)) + 16;
    anchor_l = (num_bits_t *)malloc((16 + anc_size) * sizeof(num_bits_t)) + 16;
    for (uint32_t i = 0; i < anc_size; i++) {
        anchor_r[i] = (num_bits_t)anc[i].r;
        anchor_q[i] = (num_bits_t)anc[i].q;
        anchor_l[i] = (num_bits_t)anc[i].l;
    }
}

class dp_chain {
  public:
    int max_dist_x, max_dist_y, bw, max_skip, max_iter, is_cdna, n_segs;
    float gap_scale;

    dp_chain(int max_dist_x, int max_dist_y, int bw, int max_skip, int max_iter, float gap_scale,
             int is_cdna, int n_segs, int device = 0)
        : max_dist_x(max_dist_x), max_dist_y(max_dist_y), bw(bw), max_skip(max_skip),
          max_iter(max_iter), is_cdna(is_cdna), n_segs(n_segs), gap_scale(gap_scale)
    {
        tal_chain_params p = {max_dist_x, max_dist_y, bw, max_skip, max_iter, gap_scale, is_cdna,
                              n_segs};
        const int rc = tal_chain_create(&p, device, &h_);
        if (rc != TAL_OK) throw std::runtime_error("tal_chain_create failed: " + std::to_string(rc));
    }
    ~dp_chain() { tal_chain_destroy(h_); }
    dp_chain(const dp_chain &) = delete;
    dp_chain &operator=(const dp_chain &) = delete;

    /* parallel_chaining_32_bit.h:216-217 */
    void mm_dp_vectorized(int64_t n, anchor_t *anchors, uint32_t *anchor_r, uint32_t *anchor_q,
                          uint32_t *anchor_l, uint32_t *&f, int32_t *&p, int32_t *&v, int32_t dr,
                          int32_t dq, int (*gap_cost)(anchor_t a, anchor_t b, void *meta_data),
                          void *meta_data)
    {
        (void)anchor_r;
        (void)anchor_q;
        (void)anchor_l;
        (void)gap_cost;
        (void)meta_data;
        const int64_t off[2] = {0, n};
        const int rc = tal_chain_dp(h_, (const tal_chain_anchor *)anchors, off, 1, dr, dq, f, p, v);
        if (rc != TAL_OK)
            throw std::runtime_error(std::string("tal_chain: ") + tal_chain_last_error(h_));
    }

  private:
    tal_chain *h_ = nullptr;
};

#endif
