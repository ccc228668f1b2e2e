/*
 * rowpair_model.c -- how often a 16-bit group's row i+1 could run one block
 * behind row i, in the same sweep.
 *
 * TEST AND ANALYSIS INFRASTRUCTURE ONLY.  It includes the oracle
 * (bsw_oracle.c) with its observation points defined, so the group loop it
 * observes is the oracle's own restatement of smithWaterman512_16
 * (bandedSWA.cpp:2962-3356); nothing here changes a result.
 *
 * Two rows in one sweep would read and write each state block once per two
 * rows.  Row i+1's parameters come from row i's end: its column range from
 * the narrowing (:3239-3260), each lane's band head and tail from the head
 * and tail scans (:3269-3316) and from z-drop (:1868-1881), and the row runs
 * at all only if some lane of row i scored (:3205-3209).  While row i sweeps,
 * after each of its 8-column blocks (columns below P done), the model decides
 * whether block k = P/8 - 1 - lag of row i+1 could run already (lag 1 by
 * default, rp_set_lag), with sufficient
 * rules over what row i has stored so far:
 *   - the group goes on: some lane has stored H > 0;
 *   - the range: the first column with any nonzero store is known, which
 *     fixes row i's narrowed start and so row i+1's start; the block ends
 *     below every possible end (last nonzero seen + 2, the band, ncol);
 *   - each lane's column mask over the block under every outcome still open:
 *     live or not at row i's end (z-drop is ruled out once the stored H
 *     reaches maxScore - zdrop), a first nonzero column not seen yet (the
 *     head then lies past P), a last nonzero column not seen yet (no later
 *     one can bring the tail below a column the last seen one covers);
 *   - the trim at row i+1's start: either it stops below the block (every
 *     lane's head is at or below the block's start), or whether it runs
 *     (some lane's head or tail moves) and which columns of the block it
 *     zeroes follow from the heads and tails above.
 * A block decided early keeps the masks it would use and a copy of row i's
 * stores it would read, with the trim applied.  When row i+1 starts, the oracle's own masks and its
 * state after the trims are compared with them; rp_stats reports blocks,
 * early blocks and mismatches (the rules are meant to be sound: the test
 * asserts zero mismatches).
 */
#include <stdint.h>

struct bsw_oracle;
static void rp_group(int nrow, int ncol, int w, const int16_t *band, const int16_t *qlen);
static void rp_row(int i, int beg, int end, int nbeg, const int16_t *head, const int16_t *tail,
                   const int16_t *exit0, const int16_t *maxScore, int16_t zdrop,
                   const int16_t *F, const int16_t *Hh, const int16_t *x, const int16_t *y);
static void rp_store(int j, int l, int16_t hst, int16_t fst, int16_t h, int16_t e);
static void rp_coldone(int i, int j, const int16_t *F, const int16_t *Hh);
static void rp_stop(void);

#define OBS_GROUP() rp_group(nrow, ncol, w, (const int16_t *)band, qlen)
#define OBS_ROW(i) rp_row(i, beg, end, nbeg, head, tail, exit0, maxScore, zdrop, F, Hh, x, y)
#define OBS_STORE(i, j, l, hst, fst, h, e) rp_store(j, l, hst, fst, h, e)
#define OBS_COLDONE(i, j) rp_coldone(i, j, F, Hh)
#define OBS_STOP(i) rp_stop()
#define OBS_ROWEND(i)

#include "bsw_oracle.c"

#define RP_NONE 1000000
#define RP_MAXB 4200 /* blocks of the widest 16-bit group (32767 columns) */

static int g_nrow, g_ncol, g_w;
static int16_t g_band[W16], g_qlen[W16];
/* row i as it started */
static int r_i = -1, r_beg, r_end, r_nbeg;
static int16_t r_head[W16], r_tail[W16], r_exit[W16], r_maxs[W16], r_x[W16], r_y[W16], r_zdrop;
static const int16_t *r_F, *r_Hh;
/* what row i has stored so far */
static int s_c1, s_lmax, s_anyh;
static int s_fz[W16], s_lz[W16];
static int16_t s_mrs[W16];
static int s_y1[W16]; /* latest store column holding s_mrs: the row maximum's column + 1 */
static int16_t s_h[W16], s_e[W16]; /* H of the last column swept, and E entering the next */
/* blocks of row i+1 decided early */
static int p_row = -1;
static uint8_t p_valid[RP_MAXB];
static uint32_t p_kept[RP_MAXB][8];
static int16_t p_F[RP_MAXB][8][W16], p_H[RP_MAXB][8][W16];
/* totals */
static uint64_t t_blocks, t_early, t_mismatch, t_rows;
/* the same two counts over the blocks past an LDS window of rp_window blocks
 * from the block holding the row's start (the kernel keeps those in global
 * memory); 0 = no window */
static int rp_window = 0;
/* blocks row i+1 runs behind row i (1: block k once row i has done k+1) */
static int rp_lag = 1;
static uint64_t t_gblocks, t_gearly;
/* early blocks in an unbroken run from the row's first block: the next row's
 * E and H pass from block to block, so only those can run early */
static uint64_t t_prefix, t_gprefix;
/* why a block was not decided early: 0 the group may stop, 1 it may hold the
 * row's end, 2 a lane's live/not-live outcomes differ, 3 a lane's tail is open,
 * 4 a trim may reach it */
static uint64_t t_why[5];

static void rp_group(int nrow, int ncol, int w, const int16_t *band, const int16_t *qlen)
{
    g_nrow = nrow;
    g_ncol = ncol;
    g_w = w;
    memcpy(g_band, band, sizeof g_band);
    memcpy(g_qlen, qlen, sizeof g_qlen);
    r_i = -1;
    p_row = -1;
}

static int kept16(int c, int beg, int end, int16_t head, int16_t tail)
{
    const int16_t pc = w16(c);
    return c >= beg && c < end && !(head > pc) && !(pc > tail);
}

static void rp_row(int i, int beg, int end, int nbeg, const int16_t *head, const int16_t *tail,
                   const int16_t *exit0, const int16_t *maxScore, int16_t zdrop,
                   const int16_t *F, const int16_t *Hh, const int16_t *x, const int16_t *y)
{
    /* check the blocks of this row decided during the last one */
    if (p_row == i) {
        t_rows++;
        int run = 1;
        for (int k = beg >> 3; k * 8 < end && k < RP_MAXB; k++) {
            const int global = rp_window > 0 && k >= (beg >> 3) + rp_window;
            t_blocks++;
            t_gblocks += (uint64_t)global;
            if (!p_valid[k]) {
                run = 0;
                continue;
            }
            t_early++;
            t_gearly += (uint64_t)global;
            if (run) {
                t_prefix++;
                t_gprefix += (uint64_t)global;
            }
            int bad = 0;
            for (int t = 0; t < 8; t++) {
                const int c = 8 * k + t;
                uint32_t bits = 0;
                for (int l = 0; l < W16; l++) {
                    if (kept16(c, beg, end, head[l], tail[l])) bits |= 1u << l;
                    if (c <= g_ncol && (F[(size_t)c * W16 + l] != p_F[k][t][l] ||
                                        Hh[(size_t)c * W16 + l] != p_H[k][t][l]))
                        bad = 1;
                }
                if (bits != p_kept[k][t]) bad = 1;
            }
            t_mismatch += (uint64_t)bad;
        }
    }
    r_i = i;
    r_beg = beg;
    r_end = end;
    r_nbeg = nbeg;
    memcpy(r_head, head, sizeof r_head);
    memcpy(r_tail, tail, sizeof r_tail);
    memcpy(r_exit, exit0, sizeof r_exit);
    memcpy(r_maxs, maxScore, sizeof r_maxs);
    memcpy(r_x, x, sizeof r_x);
    memcpy(r_y, y, sizeof r_y);
    r_zdrop = zdrop;
    r_F = F;
    r_Hh = Hh;
    s_c1 = RP_NONE;
    s_lmax = -1;
    s_anyh = 0;
    for (int l = 0; l < W16; l++) {
        s_fz[l] = RP_NONE;
        s_lz[l] = -1;
        s_mrs[l] = 0;
        s_y1[l] = 0;
        s_h[l] = 0;
        s_e[l] = 0;
    }
    const int nb = (g_ncol >> 3) + 2;
    memset(p_valid, 0, (size_t)(nb < RP_MAXB ? nb : RP_MAXB));
    p_row = i + 1;
}

static void rp_store(int j, int l, int16_t hst, int16_t fst, int16_t h, int16_t e)
{
    if (j < r_end) {
        s_h[l] = h;
        s_e[l] = e;
    }
    if ((hst | fst) == 0) return;
    if (j < s_c1) s_c1 = j;
    if (j > s_lmax) s_lmax = j;
    if (j < s_fz[l]) s_fz[l] = j;
    if (j > s_lz[l]) s_lz[l] = j;
    if (hst > 0) {
        s_anyh = 1;
        if (hst >= s_mrs[l]) {
            s_mrs[l] = hst;
            s_y1[l] = j;
        }
    }
}

static void rp_stop(void) { p_row = -1; }

/* the column mask of lane l over the block's 8 columns for one outcome */
static uint32_t mask8(int c0, int beg, int16_t h, int16_t t)
{
    uint32_t m = 0;
    for (int k = 0; k < 8; k++) {
        const int16_t pc = w16(c0 + k);
        if (c0 + k >= beg && !(h > pc) && !(pc > t)) m |= 1u << k;
    }
    return m;
}

/* lane l's input state (row i-1's stores) is zero at columns P..end_i-1 up
 * to the lane's tail (row i's stores past its tail are masked to zero) */
static int quiet_after(int l, int P)
{
    int hi = r_end - 1;
    if (hi > (int)r_tail[l]) hi = r_tail[l];
    for (int c = P; c <= hi; c++)
        if ((r_F[(size_t)c * W16 + l] | r_Hh[(size_t)c * W16 + l]) != 0) return 0;
    return 1;
}

/* the columns past row i's range that the tail scan may start at (end_i+1,
 * end_i+2) hold zero for lane l, so the scan's start cannot stop it there */
static int old_zero(int l)
{
    for (int c = r_end + 1; c <= r_end + 2 && c <= g_ncol; c++)
        if ((r_F[(size_t)c * W16 + l] | r_Hh[(size_t)c * W16 + l]) != 0) return 0;
    return 1;
}

static void rp_coldone(int i, int j, const int16_t *F, const int16_t *Hh)
{
    (void)F;
    (void)Hh;
    const int P = j + 1;
    if (P & 7) return;
    const int k = P / 8 - 1 - rp_lag, c0 = 8 * k;
    if (k < 0 || k >= RP_MAXB || i + 1 >= g_nrow || p_row != i + 1) return;
    if (!s_anyh || s_c1 == RP_NONE) { t_why[0]++; return; } /* the group may stop */
    /* row i+1's range */
    const int nbeg_i = s_c1 > r_beg ? s_c1 - 1 : r_nbeg;
    int begn = nbeg_i;
    if (begn < i + 1 - g_w) begn = i + 1 - g_w;
    int end_lb = s_lmax + 2;
    if (end_lb > i + 1 + g_w + 1) end_lb = i + 1 + g_w + 1;
    if (end_lb > g_ncol) end_lb = g_ncol;
    if (c0 + 8 > end_lb) { t_why[1]++; return; } /* may hold the row's end */
    if (c0 + 8 <= begn) return;  /* below the row's range */
    const int16_t i1n = w16(i + 2); /* i1 of row i+1 */
    uint32_t lane_mask[8] = {0};
    int16_t hx[W16];    /* row i+1's head of each lane */
    int16_t tx[W16];    /* its tail where known for the block (tkind 0) */
    int tkind[W16];     /* 0 tail exact, 1 tail past the block, 2 tail open */
    int any_changed = 0, all_unchanged = 1;
    for (int l = 0; l < W16; l++) {
        const int16_t hmin = w16(i + 1 - g_band[l]);
        const int16_t cap = umin16(w16(i1n + g_band[l]), g_qlen[l]);
        /* not live at row i's end: the scans leave head, the tail index stays */
        const int16_t hsNL = r_head[l];
        const int16_t tsNL = smin16(w16(r_tail[l] + 2), g_qlen[l]);
        const int16_t hNL = smax16(hsNL, hmin);
        const int16_t tNL = umin16(umin16(tsNL, w16(i1n + g_band[l])), g_qlen[l]);
        const uint32_t mNL = mask8(c0, begn, hNL, tNL);
        /* live: the first nonzero column at or past the scan's start */
        int f = RP_NONE;
        for (int c = nbeg_i; c < r_beg && f == RP_NONE; c++)
            if ((r_F[(size_t)c * W16 + l] | r_Hh[(size_t)c * W16 + l]) != 0) f = c;
        if (f == RP_NONE) f = s_fz[l];
        int okL = 1, tkL = 2;
        uint32_t mL = 0;
        int16_t hL = w16(P), hsL = w16(P), tL = 0;
        /* the post-scan tail is at least min(last nonzero seen + 2, qlen) */
        const int tsL_lb = s_lz[l] >= 0 ? (s_lz[l] + 2 < g_qlen[l] ? s_lz[l] + 2 : g_qlen[l]) : -1;
        if (f != RP_NONE) {
            hsL = f > nbeg_i ? w16(f) : r_head[l];
            hL = smax16(hsL, hmin);
            if (s_lz[l] >= 0 && s_lz[l] + 2 >= c0 + 8) {
                /* tail = min(post-scan tail >= c0 + 8, cap) */
                if (cap < c0 + 8) {
                    tkL = 0;
                    tL = cap;
                } else {
                    tkL = 1;
                }
                mL = mask8(c0, begn, hL, cap);
            } else if (s_lz[l] >= 0 && s_lz[l] + 2 >= c0 + 7) {
                mL = mask8(c0, begn, hL, cap); /* the mask is fixed; the trim's c+1 test is not */
            } else if (hL > w16(c0 + 7)) {
                mL = 0;
            } else if (((int)r_tail[l] < P || (s_h[l] == 0 && s_e[l] == 0 && quiet_after(l, P))) &&
                       old_zero(l)) {
                /* No later store of this lane can be nonzero: its stores past
                 * its tail are masked to zero, or H and E are zero past column
                 * P-1 and so is its input state from P on (H > 0 needs a
                 * nonzero diagonal, E or F).  The tail scan then ends at the
                 * last nonzero column seen (or below the scan's start). */
                int lzf = s_lz[l];
                if (lzf < 0)
                    for (int c = r_beg - 1; c >= nbeg_i && lzf < 0; c--)
                        if ((r_F[(size_t)c * W16 + l] | r_Hh[(size_t)c * W16 + l]) != 0) lzf = c;
                const int idx = lzf >= 0 ? lzf : nbeg_i - 1;
                int16_t t = smin16(w16(idx + 2), g_qlen[l]);
                t = umin16(umin16(t, w16(i1n + g_band[l])), g_qlen[l]);
                mL = mask8(c0, begn, hL, t);
                tkL = 0;
                tL = t;
            } else {
                okL = 0;
            }
        }
        /* no later store (and no later counted H) of this lane can be nonzero */
        const int settled = (int)r_tail[l] < P || (s_h[l] == 0 && s_e[l] == 0 && quiet_after(l, P));
        int live; /* 0 not live, 1 live, 2 open */
        if (r_exit[l] == 0)
            live = 0;
        else if (s_mrs[l] == 0 && settled)
            live = 0; /* its row maximum stays 0 (:3205-3211) */
        else if (settled) {
            /* the row maximum and its column are final: z-drop as at the
             * row's end (:3212-3234, ZSCORE16 :1868-1881) */
            const int16_t i1 = w16(i + 1);
            int16_t ms = r_maxs[l], xx = r_x[l], yy = r_y[l];
            const int16_t y1 = w16(s_y1[l]);
            if (s_mrs[l] > ms) {
                ms = s_mrs[l];
                xx = i1;
                yy = y1;
            }
            const int16_t di = w16(i1 - xx), dj = w16(y1 - yy);
            const int16_t off = di > dj ? w16(di - dj) : w16(dj - di);
            const int16_t drop = w16(w16(ms - s_mrs[l]) - off);
            live = drop > r_zdrop ? 0 : 1;
        }
        else if (s_mrs[l] > 0 && (int)r_maxs[l] - (int)s_mrs[l] <= (int)r_zdrop)
            live = 1;
        else
            live = 2;
        uint32_t m;
        if (live == 0) {
            m = mNL;
            hx[l] = hNL;
            tx[l] = tNL;
            tkind[l] = 0;
            if (hNL != hsNL || tNL != tsNL) any_changed = 1;
        } else if (live == 1) {
            if (!okL) { t_why[3]++; return; }
            m = mL;
            hx[l] = hL;
            tx[l] = tL;
            tkind[l] = tkL;
            if (hL != hsL || (tsL_lb >= 0 && tsL_lb > cap)) any_changed = 1;
            all_unchanged = 0; /* the post-scan tail is not known exactly */
        } else {
            /* both outcomes must give the same masks, heads and tails */
            /* the heads may differ if both lie at or below the block's start:
             * then neither masks, trims or bounds the trim inside it */
            if (f == RP_NONE || !okL || mL != mNL ||
                (hL != hNL && (hL > w16(c0) || hNL > w16(c0)))) {
                t_why[2]++;
                return;
            }
            m = mL;
            hx[l] = smax16(hL, hNL);
            if (tkL == 0 && tL == tNL) {
                tx[l] = tL;
                tkind[l] = 0;
            } else if (tkL == 1 && tNL >= c0 + 8) {
                tkind[l] = 1;
            } else {
                tkind[l] = 2;
            }
            all_unchanged = 0;
        }
        for (int t = 0; t < 8; t++)
            if ((m >> t) & 1) lane_mask[t] |= 1u << l;
    }
    /* The trim at row i+1's start (:3092-3112) zeroes, if some lane's head or
     * tail moved, each column c below the largest head in the lanes with
     * head > c or c + 1 > tail. */
    int16_t hmaxall = hx[0];
    for (int l = 1; l < W16; l++) hmaxall = smax16(hmaxall, hx[l]);
    int trims; /* 0 none, 1 as computed */
    if (hmaxall <= w16(c0))
        trims = 0; /* the trim stops below the block */
    else if (any_changed)
        trims = 1;
    else if (all_unchanged)
        trims = 0;
    else {
        t_why[4]++;
        return;
    }
    uint32_t clear[8] = {0};
    if (trims)
        for (int t = 0; t < 8; t++) {
            const int c = c0 + t;
            const int16_t pc = w16(c);
            if (c < begn || !(hmaxall > pc)) continue; /* the trim runs over [beg, largest head) */
            for (int l = 0; l < W16; l++) {
                int cl;
                if (hx[l] > pc)
                    cl = 1;
                else if (tkind[l] == 0)
                    cl = w16(c + 1) > tx[l];
                else if (tkind[l] == 1)
                    cl = 0;
                else {
                    t_why[4]++;
                    return;
                }
                if (cl) clear[t] |= 1u << l;
            }
        }
    p_valid[k] = 1;
    for (int t = 0; t < 8; t++) {
        p_kept[k][t] = lane_mask[t];
        const int c = c0 + t;
        for (int l = 0; l < W16; l++) {
            const int z = (clear[t] >> l) & 1;
            p_F[k][t][l] = c <= g_ncol && !z ? r_F[(size_t)c * W16 + l] : 0;
            p_H[k][t][l] = c <= g_ncol && !z ? r_Hh[(size_t)c * W16 + l] : 0;
        }
    }
}

/* blocks of rows after a group's first, those decided early, and early
 * blocks whose masks or inputs differ from the oracle's (expected 0) */
void rp_stats(uint64_t out[4])
{
    out[0] = t_blocks;
    out[1] = t_early;
    out[2] = t_mismatch;
    out[3] = t_rows;
}

void rp_set_lag(int lag) { rp_lag = lag < 1 ? 1 : lag; }

void rp_global(int window, uint64_t out[4])
{
    rp_window = window;
    out[0] = t_gblocks;
    out[1] = t_gearly;
    out[2] = t_prefix;
    out[3] = t_gprefix;
}

void rp_why(uint64_t out[5])
{
    for (int k = 0; k < 5; k++) out[k] = t_why[k];
}

void rp_reset(void)
{
    t_blocks = t_early = t_mismatch = t_rows = 0;
    t_gblocks = t_gearly = t_prefix = t_gprefix = 0;
    for (int k = 0; k < 5; k++) t_why[k] = 0;
}
