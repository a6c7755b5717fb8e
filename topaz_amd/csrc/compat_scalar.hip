// Exports of the reference's search.o and alignment.o that TOPAZ itself never
// calls: sa_lcp and sa_gapless (search.c:32-97), ungapped_rawscore and
// ungapped_length (alignment.c:266-315) and the full-matrix Gotoh aligner
// with its helpers (alignment.c:27-396).  With these and the GPU entry points
// (sa_search, sa_hsp), a program links against libtopaz_amd.so with neither
// object file.  Each is a single scalar computation on the caller's host
// arrays, so it runs on the host: one call is far too little work for a
// kernel launch.  Results match the reference's call for call, tie rules and
// abort paths included.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/topaz_ssw_compat.h"
#include "alphabet.h"

namespace {

// Trace codes of the three Gotoh matrices (alignment.h:12-21).
enum Trace : uint16_t {
    kNull = 0,
    kDiagMatch = 1,
    kHFromE = 2,
    kHFromF = 3,
    kEExtend = 4,
    kEOpen = 5,
    kFExtend = 6,
    kFOpen = 7,
    kDiagMiss = 8,
};

int sub(const options_t* opt, int row, int col) { return (*opt->substitution)[row * 25 + col]; }

// dim2 rows of dim1 values, row-major, each as "%3d " (alignment.c:27-45)
template <typename T>
void print_matrix(const T* m, size_t dim1, size_t dim2) {
    size_t k = 0;
    for (size_t r = 0; r < dim2; ++r) {
        for (size_t c = 0; c < dim1; ++c) printf("%3d ", (int)m[k++]);
        printf("\n");
    }
}

}  // namespace

extern "C" {

int sa_lcp(db_t* db, seq_t* q, int qoffset, int64_t position) {
    // letters of the suffix SA[position] and the query from qoffset that agree,
    // counting the first even past the query's end, at most the query's
    // remaining length once the first agrees
    const char* t = db->T + db->SA[position];
    const char* p = q->seq_buf.buf + qoffset;
    const int remaining = q->seq_len - qoffset;
    int n = 0;
    while (t[n] == p[n])
        if (++n >= remaining) break;
    return n;
}

int sa_gapless(db_t* db, seq_t* q, int qoffset, int64_t position, options_t* opt,
               int* gapless_score, int* gapless_length) {
    const tpz::Alphabet& ab = tpz::alphabet();
    const int64_t at = db->SA[position];
    const char* t = db->T + at;
    const char* p = q->seq_buf.buf + qoffset;
    // min of an unsigned text remainder and the signed query remainder, as C
    // compares them (both as uint64), kept as an int
    const uint64_t text_left = db->Tsize - (uint64_t)at;
    const uint64_t query_left = (uint64_t)(int64_t)(q->seq_len - qoffset);
    const int span = (int)(text_left < query_left ? text_left : query_left);
    if (span < 3 || t[0] != p[0]) {
        *gapless_length = -1;
        return 0;
    }
    // best prefix score of the right extension (later prefixes win ties),
    // stopping at the protein's end
    int run = 0, best = 0, best_end = 0;
    for (int i = 0; i < span && t[i] != '\n'; ++i) {
        run += sub(opt, ab.idx(t[i]), ab.idx(ab.upper_only[(unsigned char)p[i]]));
        if (run >= best) {
            best = run;
            best_end = i;
        }
    }
    *gapless_score = best;
    *gapless_length = best_end;
    return 0;
}

int ungapped_rawscore(char* s1, char* s2, uint32_t length, options_t* opt) {
    // the reference scores s1 against itself (s2 is unused)
    (void)s2;
    const tpz::Alphabet& ab = tpz::alphabet();
    int total = 0;
    for (uint32_t i = 0; i < length; ++i) {
        const int r = ab.idx(s1[i]);
        total += sub(opt, r, r);
    }
    return total;
}

int ungapped_length(char* s1, char* s2, uint32_t max_length, options_t* opt) {
    // the reference's scan never runs (its loop condition starts false), so
    // every call gives max_length back
    (void)s1;
    (void)s2;
    (void)opt;
    return (int)max_length;
}

void printdp32(int* m, size_t dim1, size_t dim2) { print_matrix(m, dim1, dim2); }

void printdp8(uint8_t* m, size_t dim1, size_t dim2) { print_matrix(m, dim1, dim2); }

void _initialise(seq_t* s1, seq_t* s2, int* E, int* F, int* H, uint16_t* trE, uint16_t* trF,
                 uint16_t* trH, options_t* par) {
    (void)par;
    // row 0 and column 0 of the (len1 + 1) x (len2 + 1) row-major matrices
    const size_t cols = (size_t)s1->seq_len + 1, rows = (size_t)s2->seq_len + 1;
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < (r == 0 ? cols : 1); ++c) {
            const size_t x = r * cols + c;
            E[x] = F[x] = H[x] = 0;
            trE[x] = trF[x] = trH[x] = kNull;
        }
}

int _forward_pass(int8_t* s1, int8_t* s2, size_t s1_len, size_t s2_len, int* E, uint16_t* trE,
                  int* F, uint16_t* trF, int* H, uint16_t* trH, hit_t* hit, options_t* par) {
    // Gotoh with linear gap lengths charged go for the first letter and ge
    // for each further one; cell (row y: s2[y-1], column x: s1[x-1]) at
    // y * cols + x.  Ties: E and F prefer opening, H prefers diagonal, then
    // E, then F; the first cell in row-major order with the top score wins.
    const int go = par->gap_open, ge = par->gap_extend;
    const size_t cols = s1_len + 1;
    int top = 0;
    int top_at = 0;
    for (size_t y = 1; y <= s2_len; ++y) {
        const int r = s2[y - 1];
        for (size_t x = 1; x <= s1_len; ++x) {
            const size_t at = y * cols + x, left = at - 1, up = at - cols, diag = up - 1;
            const int e_ext = E[left] - ge, e_open = H[left] - go;
            E[at] = e_ext > e_open ? e_ext : e_open;
            trE[at] = e_ext > e_open ? kEExtend : kEOpen;
            const int f_ext = F[up] - ge, f_open = H[up] - go;
            F[at] = f_ext > f_open ? f_ext : f_open;
            trF[at] = f_ext > f_open ? kFExtend : kFOpen;
            const int c = s1[x - 1];
            const int d = H[diag] + sub(par, r, c);
            int h = 0;
            uint16_t from = kNull;
            if (d > h) {
                h = d;
                from = r == c ? kDiagMatch : kDiagMiss;
            }
            if (E[at] > h) {
                h = E[at];
                from = kHFromE;
            }
            if (F[at] > h) {
                h = F[at];
                from = kHFromF;
            }
            H[at] = h;
            trH[at] = from;
            if (h > top) {
                top = h;
                top_at = (int)at;
            }
        }
    }
    hit->rawscore = top;
    hit->send = top_at / (int)cols;
    hit->qend = top_at % (int)cols;
    return top_at;
}

void _backtrack(uint16_t* trE, uint16_t* trF, uint16_t* trH, int pos, int max_x, hit_t* hit) {
    // Walk the trace from pos: H steps diagonally (a match or a mismatch) or
    // hands over to E or F (a gap opens); E steps left and F up, one gap
    // letter each, back to H where the code says the gap opened.  Stops at
    // the first code 0 the walk reads after a step; a code H cannot hold (0
    // at the start included) is the reference's fatal error.
    char in = 'h';
    uint16_t code = trH[pos];
    int letters = 0, same = 0, differ = 0, opens = 0;
    for (;;) {
        if (in == 'h') {
            if (code == kDiagMatch || code == kDiagMiss) {
                pos -= max_x + 1;
                ++letters;
                if (code == kDiagMatch) ++same;
                else ++differ;
            } else if (code == kHFromE || code == kHFromF) {
                ++opens;
                in = code == kHFromE ? 'e' : 'f';
            } else {
                fprintf(stderr, "Error: backtrack failed, illegal condition\n");
                abort();
            }
        } else if (in == 'e') {
            --pos;
            ++letters;
            if (code == kEOpen) in = 'h';
        } else {
            pos -= max_x;
            ++letters;
            if (code == kFOpen) in = 'h';
        }
        code = in == 'h' ? trH[pos] : in == 'e' ? trE[pos] : trF[pos];
        if (code == kNull) break;
    }
    hit->sstart = pos / max_x + 1;
    hit->qstart = pos % max_x + 1;
    hit->mismatch = differ;
    hit->gapopen = opens;
    hit->length = letters;
    hit->identity = same / (float)letters;
}

int gotoh(seq_t* s1, seq_t* s2, hit_t* hit, options_t* opt) {
    const tpz::Alphabet& ab = tpz::alphabet();
    const int n1 = s1->seq_len, n2 = s2->seq_len;
    const size_t cells = ((size_t)n1 + 1) * ((size_t)n2 + 1);
    std::vector<int> E(cells), F(cells), H(cells);
    std::vector<uint16_t> trE(cells), trF(cells), trH(cells);
    _initialise(s1, s2, E.data(), F.data(), H.data(), trE.data(), trF.data(), trH.data(), opt);
    std::vector<int8_t> r1((size_t)n1 + 1), r2((size_t)n2 + 1);
    for (int i = 0; i < n1; ++i) r1[i] = ab.idx(s1->seq_buf.buf[i]);
    for (int i = 0; i < n2; ++i) r2[i] = ab.idx(s2->seq_buf.buf[i]);
    const int end = _forward_pass(r1.data(), r2.data(), (size_t)n1, (size_t)n2, E.data(),
                                  trE.data(), F.data(), trF.data(), H.data(), trH.data(), hit, opt);
    _backtrack(trE.data(), trF.data(), trH.data(), end, n1 + 1, hit);
    // Karlin-Altschul statistics with the query-length and database-size
    // corrections (alignment.c:366-381)
    const double correction = log(opt->k * n1 * opt->db_size) / opt->H;
    const double q_eff = fmax(n1 - correction, 1.0);
    const double db_eff = fmax(opt->db_size - correction * opt->db_proteins, 1.0);
    const double nats = opt->lambda * hit->rawscore - log(opt->k);
    const double e = q_eff * db_eff * exp(-nats);
    hit->evalue = e > 1.0e-180 ? e : 0.0;
    hit->bitscore = nats / M_LN2;
    if (correction > hit->length) hit->evalue = INFINITY;
    hit->is_aligned = 1;
    hit->effective_length = (int)(n1 - correction);
    return 0;
}

}  // extern "C"
