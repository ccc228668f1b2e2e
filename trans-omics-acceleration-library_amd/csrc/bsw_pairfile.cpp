/*
 * bsw_pairfile.cpp -- pair files for the BSW engine (tal_bsw_pairfile_* and
 * tal_bsw_pack_* in include/tal_bsw.h).  Host code only.
 *
 * Text format: the one bench-bsw reads (benchmarks/bench-bsw.cpp:117-169):
 * three lines per pair -- h0 in decimal, the reference, the query -- with
 * each base an ASCII character whose code is the character minus '0' ('0'-'3'
 * = ACGT, '4' = N).  Pairs are read in chunks into fixed per-pair slots, as
 * bench-bsw lays them out (MAX_SEQ_LEN_REF = 256 and MAX_SEQ_LEN_QER = 128
 * byte slots, bandedSWA.h:47-48).
 *
 * Packed format (this project's): a 64-byte header, the SeqPair records, the
 * reference bytes and the query bytes, so that a batch can be memory-mapped
 * and handed to the engine without parsing.
 */
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <new>
#include <string>

#include "tal_bsw.h"

struct tal_bsw_pairfile {
    FILE *f = nullptr;
    std::string line;
    int64_t line_no = 0;
};

namespace {

/* One line without its newline; false at end of file. */
bool read_line(tal_bsw_pairfile *pf)
{
    pf->line.clear();
    int c;
    bool any = false;
    while ((c = fgetc(pf->f)) != EOF) {
        any = true;
        if (c == '\n') break;
        pf->line.push_back((char)c);
    }
    if (!any) return false;
    if (!pf->line.empty() && pf->line.back() == '\r') pf->line.pop_back();
    pf->line_no++;
    return true;
}

constexpr char PACK_MAGIC[8] = {'T', 'A', 'L', 'B', 'S', 'W', 'P', '1'};

struct PackHeader {
    char magic[8];
    int64_t n;
    int64_t ref_len;
    int64_t qer_len;
    int64_t pairs_off, ref_off, qer_off;
    int64_t reserved;
};
static_assert(sizeof(PackHeader) == 64, "pack header is 64 bytes");

} // namespace

extern "C" {

int tal_bsw_pairfile_open(const char *path, tal_bsw_pairfile **out)
{
    if (!path || !out) return TAL_EINVAL;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return TAL_EINVAL;
    tal_bsw_pairfile *pf = new (std::nothrow) tal_bsw_pairfile();
    if (!pf) {
        fclose(f);
        return TAL_ENOMEM;
    }
    pf->f = f;
    *out = pf;
    return TAL_OK;
}

int tal_bsw_pairfile_read(tal_bsw_pairfile *pf, tal_seqpair *pairs, uint8_t *ref, size_t ref_slot,
                          uint8_t *qer, size_t qer_slot, int32_t max_pairs, int32_t *n_read)
{
    if (!pf || !n_read || max_pairs < 0 || (max_pairs > 0 && (!pairs || !ref || !qer)) ||
        ref_slot < 3 || qer_slot < 3 || (size_t)max_pairs * ref_slot > (size_t)INT32_MAX)
        return TAL_EINVAL;
    *n_read = 0;
    int32_t n = 0;
    while (n < max_pairs) {
        /* an h0 line with no reference line after it ends the file cleanly,
         * as in bench-bsw; a reference without its query is an error there
         * (the "odd number of sequences" message) and here */
        if (!read_line(pf)) break;
        const int h0 = atoi(pf->line.c_str());
        if (!read_line(pf)) break;
        const std::string r = pf->line;
        if (!read_line(pf)) return TAL_EINVAL;
        const std::string &q = pf->line;
        /* bench-bsw's fgets stores the sequence, its newline and a NUL in the
         * slot, so a slot of S bytes holds at most S-2 bases */
        if (r.size() > ref_slot - 2 || q.size() > qer_slot - 2) return TAL_ERANGE;
        uint8_t *rs = ref + (size_t)n * ref_slot;
        uint8_t *qs = qer + (size_t)n * qer_slot;
        for (size_t k = 0; k < r.size(); k++) rs[k] = (uint8_t)(r[k] - '0');
        for (size_t k = 0; k < q.size(); k++) qs[k] = (uint8_t)(q[k] - '0');
        tal_seqpair sp;
        sp.idr = (int32_t)((size_t)n * ref_slot);
        sp.idq = (int32_t)((size_t)n * qer_slot);
        sp.id = n;
        sp.len1 = (int32_t)r.size();
        sp.len2 = (int32_t)q.size();
        sp.h0 = h0;
        sp.seqid = sp.regid = -1;
        sp.score = sp.tle = sp.gtle = sp.qle = sp.gscore = sp.max_off = -1;
        pairs[n++] = sp;
    }
    *n_read = n;
    return TAL_OK;
}

void tal_bsw_pairfile_close(tal_bsw_pairfile *pf)
{
    if (!pf) return;
    if (pf->f) fclose(pf->f);
    delete pf;
}

int tal_bsw_pairfile_write(const char *path, const tal_seqpair *pairs, const uint8_t *ref,
                           const uint8_t *qer, int32_t n)
{
    if (!path || n < 0 || (n > 0 && (!pairs || !ref || !qer))) return TAL_EINVAL;
    FILE *f = fopen(path, "wb");
    if (!f) return TAL_EINVAL;
    std::string buf;
    for (int32_t k = 0; k < n; k++) {
        const tal_seqpair &p = pairs[k];
        if (p.len1 < 0 || p.len2 < 0 || p.idr < 0 || p.idq < 0) {
            fclose(f);
            return TAL_EINVAL;
        }
        buf = std::to_string(p.h0);
        buf.push_back('\n');
        for (int32_t c = 0; c < p.len1; c++) buf.push_back((char)(ref[(size_t)p.idr + c] + '0'));
        buf.push_back('\n');
        for (int32_t c = 0; c < p.len2; c++) buf.push_back((char)(qer[(size_t)p.idq + c] + '0'));
        buf.push_back('\n');
        if (fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
            fclose(f);
            return TAL_EDEVICE;
        }
    }
    return fclose(f) == 0 ? TAL_OK : TAL_EDEVICE;
}

int tal_bsw_pack_write(const char *path, const tal_seqpair *pairs, int64_t n, const uint8_t *ref,
                       size_t ref_len, const uint8_t *qer, size_t qer_len)
{
    if (!path || n < 0 || (n > 0 && !pairs) || (ref_len && !ref) || (qer_len && !qer))
        return TAL_EINVAL;
    PackHeader hd{};
    memcpy(hd.magic, PACK_MAGIC, sizeof hd.magic);
    hd.n = n;
    hd.ref_len = (int64_t)ref_len;
    hd.qer_len = (int64_t)qer_len;
    hd.pairs_off = sizeof(PackHeader);
    hd.ref_off = hd.pairs_off + (int64_t)sizeof(tal_seqpair) * n;
    hd.qer_off = hd.ref_off + (int64_t)ref_len;
    FILE *f = fopen(path, "wb");
    if (!f) return TAL_EINVAL;
    bool ok = fwrite(&hd, sizeof hd, 1, f) == 1;
    ok = ok && (n == 0 || fwrite(pairs, sizeof(tal_seqpair), (size_t)n, f) == (size_t)n);
    ok = ok && (ref_len == 0 || fwrite(ref, 1, ref_len, f) == ref_len);
    ok = ok && (qer_len == 0 || fwrite(qer, 1, qer_len, f) == qer_len);
    ok = (fclose(f) == 0) && ok;
    return ok ? TAL_OK : TAL_EDEVICE;
}

int tal_bsw_pack_map(const char *path, tal_bsw_pack *out)
{
    if (!path || !out) return TAL_EINVAL;
    memset(out, 0, sizeof *out);
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return TAL_EINVAL;
    struct stat sb;
    if (fstat(fd, &sb) != 0 || (size_t)sb.st_size < sizeof(PackHeader)) {
        close(fd);
        return TAL_EINVAL;
    }
    /* private mapping: the caller may write outputs into the records */
    void *m = mmap(nullptr, (size_t)sb.st_size, PROT_READ | PROT_WRITE, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) return TAL_ENOMEM;
    const PackHeader *hd = (const PackHeader *)m;
    const int64_t size = (int64_t)sb.st_size;
    const bool ok = memcmp(hd->magic, PACK_MAGIC, sizeof hd->magic) == 0 && hd->n >= 0 &&
                    hd->n <= size / (int64_t)sizeof(tal_seqpair) && hd->ref_len >= 0 &&
                    hd->ref_len <= size && hd->qer_len >= 0 && hd->qer_len <= size &&
                    hd->pairs_off == (int64_t)sizeof(PackHeader) &&
                    hd->ref_off == hd->pairs_off + (int64_t)sizeof(tal_seqpair) * hd->n &&
                    hd->qer_off == hd->ref_off + hd->ref_len &&
                    hd->qer_off + hd->qer_len <= size;
    if (!ok) {
        munmap(m, (size_t)sb.st_size);
        return TAL_EINVAL;
    }
    uint8_t *base = (uint8_t *)m;
    out->pairs = (tal_seqpair *)(base + hd->pairs_off);
    out->n = hd->n;
    out->ref = base + hd->ref_off;
    out->ref_len = (size_t)hd->ref_len;
    out->qer = base + hd->qer_off;
    out->qer_len = (size_t)hd->qer_len;
    out->mapping = m;
    out->mapping_len = (size_t)sb.st_size;
    return TAL_OK;
}

void tal_bsw_pack_unmap(tal_bsw_pack *p)
{
    if (!p || !p->mapping) return;
    munmap(p->mapping, p->mapping_len);
    memset(p, 0, sizeof *p);
}

} /* extern "C" */
