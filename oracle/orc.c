/*
 * orc.c -- CPU oracle for TOPAZ's candidate-alignment hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  Plain scalar C written for
 * readability, not speed: one DP cell at a time, no SIMD.
 */
#include "orc.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Alphabet (lexicographical.h:6-74)                                   */
/* ------------------------------------------------------------------ */

/* The internal alphabet stores amino acid number k (in the BLOSUM62-derived
 * ordering of lexicographical.h:7) as the letter 'A'+k; the matrices of
 * substitution.h are indexed in the order of their column header
 * (substitution.h:80).  aa_table2 (lexicographical.h:19-33) is the
 * composition of the two orderings; every other byte maps to 23 ('X'). */
static const char kInternalOrder[] = "ACMLJIVTSKRQZEDBNHYFWXP*G";
static const char kMatrixOrder[]   = "ARNDCQEGHILKMFPSTWYVBJZX*";

int8_t orc_aa_table2(unsigned char c) {
    int k;
    if (c >= 'a' && c <= 'y') c = (unsigned char)(c - 32);
    if (c < 'A' || c > 'Y') return 23;
    k = c - 'A';
    return (int8_t)(strchr(kMatrixOrder, kInternalOrder[k]) - kMatrixOrder);
}

/* lexicographical.h:35-58 */
char orc_upper_aa(char c) {
    if ((c >= 'A' && c <= 'Y') || c == '*' || c == '\n') return c;
    if (c >= 'a' && c <= 'y') return (char)(c - 32);
    return 'X';
}

/* lexicographical.h:60-74 */
char orc_upper_aa_only(char c) {
    if ((c >= 'A' && c <= 'Y') || c == '*' || c == '\n') return c;
    return 'V';
}

static int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ */
/* Striped Smith-Waterman scan (ssw.c:127-332 byte, ssw.c:358-536 word) */
/* ------------------------------------------------------------------ */
/*
 * The reference splits the query into P lanes of segLen = ceil(m/P)
 * consecutive rows, padded to P * segLen rows whose score is 0 (ssw.c:112,
 * 350).  Its main loop (ssw.c:203-228 / 428-452) carries the vertical gap
 * value only inside a lane, restarting it at 0 on every row j with
 * j % segLen == 0; it keeps Hp = max(D + W, E, Fin) and updates E from Hp.
 * The lazy-F loop (ssw.c:230-262 / 454-466) then carries each lane's final
 * gap value into the next lane, raising H but never E.
 *
 * This function follows those steps with one scalar per lane.  The main
 * loop is written row by row (the lanes of the reference are independent
 * there); the lazy-F loop is emulated lane by lane because its exit test
 * looks at every lane at once (the byte path tests before raising, the word
 * path after raising and at most 8 sweeps).  For gapO > gapE both loops
 * give H = max(Hp, F) with the textbook F = max(F - gapE, Hp - gapO) never
 * restarted -- orc_sw_scan_s1 below -- but not for gapO <= gapE.
 *
 * Column bookkeeping follows ssw.c:264-297 / 469-500: the running max
 * rises only on a strictly larger column max (padding rows included), the
 * column of that rise is end_ref and its H column is kept; the byte path
 * gives up once the max reaches 255 - bias (its saturation point), the word
 * path at 32767; terminate stops after the first column whose max equals
 * it; end_read is the lowest real row of the kept column equal to the final
 * max, else m - 1.
 */
typedef struct {
    int32_t m, n, P, segLen, rows, cap, ovf, gapO, gapE;
    int32_t *hprev, *hcur, *e, *kept, *carry, *next;
} scan_t;

static void scan_column(scan_t* st, const int8_t* read, const int8_t* wrow,
                        int32_t* colmax_out) {
    const int32_t seg = st->segLen, P = st->P, go = st->gapO, ge = st->gapE;
    int32_t* h = st->hcur;
    int32_t fin = 0, colmax = 0, j, L;

    /* main loop: Hp, E and the in-lane gap value; carry[L] is the gap value
     * left at the end of lane L (ssw.c:223-224 after the last segment). */
    for (j = 0; j < st->rows; ++j) {
        int32_t d = j ? st->hprev[j - 1] : 0;
        int32_t w = j < st->m ? wrow[read[j]] : 0;
        int32_t hp, hgo;
        if (j % seg == 0) fin = 0;
        hp = imax(d + w, 0);
        hp = imax(hp, st->e[j]);
        hp = imax(hp, fin);
        hgo = imax(hp - go, 0);
        st->e[j] = imax(imax(st->e[j] - ge, 0), hgo);
        fin = imax(imax(fin - ge, 0), hgo);
        h[j] = hp;
        colmax = imax(colmax, hp);
        if (j % seg == seg - 1) st->carry[j / seg] = fin;
    }

    /* lazy-F: shift the lane carries by one lane (lane 0 gets 0) */
    for (L = P - 1; L > 0; --L) st->carry[L] = st->carry[L - 1];
    st->carry[0] = 0;

    if (P == 16) {
        /* ssw.c:232-262: test before raising; wrap with a further shift */
        j = 0;
        for (;;) {
            int done = 1;
            for (L = 0; L < P; ++L)
                if (st->carry[L] > imax(h[L * seg + j] - go, 0)) done = 0;
            if (done) break;
            for (L = 0; L < P; ++L) {
                int32_t* c = &h[L * seg + j];
                *c = imax(*c, st->carry[L]);
                colmax = imax(colmax, *c);
                st->carry[L] = imax(st->carry[L] - ge, 0);
            }
            if (++j >= seg) {
                j = 0;
                for (L = P - 1; L > 0; --L) st->carry[L] = st->carry[L - 1];
                st->carry[0] = 0;
            }
        }
    } else {
        /* ssw.c:455-466: at most 8 sweeps; test after raising */
        int k;
        for (k = 0; k < 8; ++k) {
            if (k > 0) {
                for (L = P - 1; L > 0; --L) st->carry[L] = st->carry[L - 1];
                st->carry[0] = 0;
            }
            for (j = 0; j < seg; ++j) {
                int any = 0;
                for (L = 0; L < P; ++L) {
                    int32_t* c = &h[L * seg + j];
                    *c = imax(*c, st->carry[L]);
                    colmax = imax(colmax, *c);
                    st->carry[L] = imax(st->carry[L] - ge, 0);
                    if (st->carry[L] > imax(*c - go, 0)) any = 1;
                }
                if (!any) goto end;
            }
        }
    end:;
    }
    *colmax_out = colmax;
}

static orc_end_t scan_run(scan_t* st, const int8_t* read, const int8_t* ref,
                          int dir, const int8_t* mat, int32_t nmat,
                          int32_t terminate, int literal) {
    orc_end_t r;
    int32_t best = 0, end_ref = (st->P == 16) ? -1 : 0, end_read, k, j;
    for (k = 0; k < st->n; ++k) {
        const int32_t i = dir ? st->n - 1 - k : k;
        const int8_t* wrow = mat + (size_t)ref[i] * nmat;
        int32_t colmax, *t;
        if (literal) {
            scan_column(st, read, wrow, &colmax);
        } else {
            /* closed form S1: H = max(Hp, F), F never restarted */
            int32_t fin = 0, f = 0;
            colmax = 0;
            for (j = 0; j < st->rows; ++j) {
                int32_t d = j ? st->hprev[j - 1] : 0;
                int32_t w = j < st->m ? wrow[read[j]] : 0;
                int32_t hp, hv, hgo;
                if (j % st->segLen == 0) fin = 0;
                hp = imax(imax(imax(d + w, 0), st->e[j]), fin);
                hv = imax(hp, f);
                hgo = imax(hp - st->gapO, 0);
                st->e[j] = imax(imax(st->e[j] - st->gapE, 0), hgo);
                fin = imax(imax(fin - st->gapE, 0), hgo);
                f = imax(imax(f - st->gapE, 0), hgo);
                st->hcur[j] = hv;
                colmax = imax(colmax, hv);
            }
        }
        t = st->hprev; st->hprev = st->hcur; st->hcur = t;
        if (colmax >= st->cap) colmax = st->cap;  /* saturated lanes read as cap */
        if (colmax > best) {
            best = colmax;
            if (best >= st->cap) break;
            end_ref = i;
            memcpy(st->kept, st->hprev, (size_t)st->rows * sizeof(int32_t));
        }
        if (colmax == terminate) break;
    }
    end_read = st->m - 1;
    for (j = 0; j < st->m; ++j) {
        if (st->kept[j] == best) { end_read = j; break; }
    }
    r.score = best >= st->cap ? st->ovf : best;
    r.ref = end_ref;
    r.read = end_read;
    return r;
}

static orc_end_t sw_scan_impl(const int8_t* read, int32_t readLen,
                              const int8_t* ref, int32_t refLen, int dir,
                              const int8_t* mat, int32_t nmat, int P,
                              int32_t gapO, int32_t gapE, int32_t bias,
                              int32_t terminate, int literal) {
    scan_t st;
    orc_end_t r;
    size_t cells;
    st.m = readLen; st.n = refLen; st.P = P;
    st.segLen = (readLen + P - 1) / P;
    st.rows = st.segLen * P;
    st.cap = (P == 16) ? 255 - bias : 32767;
    st.ovf = (P == 16) ? 255 : 32767;
    st.gapO = gapO; st.gapE = gapE;
    cells = (size_t)(st.rows > 0 ? st.rows : 1);
    st.hprev = (int32_t*)calloc(cells, sizeof(int32_t));
    st.hcur  = (int32_t*)calloc(cells, sizeof(int32_t));
    st.e     = (int32_t*)calloc(cells, sizeof(int32_t));
    st.kept  = (int32_t*)calloc(cells, sizeof(int32_t));
    st.carry = (int32_t*)calloc((size_t)P, sizeof(int32_t));
    st.next  = NULL;
    if (st.rows == 0) {                   /* empty query: nothing to scan */
        r.score = 0; r.ref = (P == 16) ? -1 : 0; r.read = -1;
    } else {
        r = scan_run(&st, read, ref, dir, mat, nmat, terminate, literal);
    }
    free(st.hprev); free(st.hcur); free(st.e); free(st.kept); free(st.carry);
    return r;
}

orc_end_t orc_sw_scan(const int8_t* read, int32_t readLen,
                      const int8_t* ref, int32_t refLen, int dir,
                      const int8_t* mat, int32_t nmat, int P,
                      int32_t gapO, int32_t gapE, int32_t bias,
                      int32_t terminate) {
    return sw_scan_impl(read, readLen, ref, refLen, dir, mat, nmat, P, gapO,
                        gapE, bias, terminate, 1);
}

orc_end_t orc_sw_scan_s1(const int8_t* read, int32_t readLen,
                         const int8_t* ref, int32_t refLen, int dir,
                         const int8_t* mat, int32_t nmat, int P,
                         int32_t gapO, int32_t gapE, int32_t bias,
                         int32_t terminate) {
    return sw_scan_impl(read, readLen, ref, refLen, dir, mat, nmat, P, gapO,
                        gapE, bias, terminate, 0);
}

static int32_t matrix_bias(const int8_t* mat, int32_t nmat) {
    int32_t b = 0, i;                      /* ssw_init, ssw.c:805-811 */
    for (i = 0; i < nmat * nmat; ++i) if (mat[i] < b) b = mat[i];
    return -b;
}

int orc_ssw_align(const int8_t* read, int32_t readLen,
                  const int8_t* ref, int32_t refLen,
                  const int8_t* mat, int32_t nmat,
                  int32_t gapO, int32_t gapE, orc_end_t* out) {
    const int32_t bias = matrix_bias(mat, nmat);
    /* ssw.c:861-870: byte first, word when the byte score reads 255. */
    orc_end_t r = orc_sw_scan(read, readLen, ref, refLen, 0, mat, nmat, 16,
                              gapO, gapE, bias, -1);
    if (r.score == 255) {
        r = orc_sw_scan(read, readLen, ref, refLen, 0, mat, nmat, 8,
                        gapO, gapE, bias, -1);
        if (r.score == 32767) return 1;
    }
    *out = r;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Banded global alignment with traceback statistics (ssw.c:538-784)  */
/* ------------------------------------------------------------------ */
/* Row i (query position) keeps the band columns [max(0,i-w), min(n-1,i+w)].
 * Band slot of column j on row i: j - max(0, i-w) + 1 (slot 0 is the left
 * border); direction cell (j - max(0,i-w))*3 + {0:E,1:F,2:H}. */
static int32_t band_slot(int32_t w, int32_t i, int32_t j) {
    return j - imax(i - w, 0) + 1;
}
static int32_t band_dir(int32_t w, int32_t i, int32_t j, int32_t p) {
    return (j - imax(i - w, 0)) * 3 + p;
}

/* Next power of two above x (kroundup32 after ++, ssw.c:65-66, 572-574,
 * 593-595). */
static int64_t grow_pow2(int64_t x) {
    int64_t v = x + 1;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}

/* Grow a buffer keeping its contents; the new part reads as zero (what a
 * fresh allocation holds; the reference's realloc leaves it unspecified). */
static void* grow_zero(void* p, size_t old_bytes, size_t new_bytes) {
    char* q = (char*)realloc(p, new_bytes);
    memset(q + old_bytes, 0, new_bytes - old_bytes);
    return q;
}

/* The reference keeps one direction buffer and three band rows for all
 * passes of the doubling loop (ssw.c:559-563, 575-597): each pass writes
 * the cells of its own band layout, E is never cleared between passes and
 * H only on slots 1..width-2, and the traceback indexes the direction
 * buffer as one flat array, so a step outside the final band's row reads
 * what the row above or below (or an earlier pass) left there.  This
 * restatement does the same with buffers that start zeroed; reads outside
 * the direction buffer end the traceback as an error. */
int orc_banded_sw(const int8_t* ref, const int8_t* read,
                  int32_t refLen, int32_t readLen, int32_t score,
                  int32_t gapO, int32_t gapE, int32_t band_width,
                  const int8_t* mat, int32_t nmat,
                  int32_t* length, int32_t* gapopen, int32_t* mismatch,
                  double* identity) {
    int32_t w = band_width, best = 0, width = 0, width_d = 0;
    int64_t s1 = 8, s2 = 1024;
    int32_t* hb = (int32_t*)calloc((size_t)s1, sizeof(int32_t));
    int32_t* eb = (int32_t*)calloc((size_t)s1, sizeof(int32_t));
    int32_t* hc = (int32_t*)calloc((size_t)s1, sizeof(int32_t));
    int8_t* dirs = (int8_t*)calloc((size_t)s2, 1);
    int32_t i, j, state, identical = 0;
    char op, prev_op;
    int64_t line;

    *length = 0; *gapopen = 0; *mismatch = 0; *identity = 0.0;

    do {
        int64_t ncell;
        width = 2 * w + 3;
        width_d = 2 * w + 1;
        if (width >= s1) {
            int64_t ns = s1;
            while (width >= ns) ns = grow_pow2(ns + 1);
            hb = (int32_t*)grow_zero(hb, (size_t)s1 * 4, (size_t)ns * 4);
            eb = (int32_t*)grow_zero(eb, (size_t)s1 * 4, (size_t)ns * 4);
            hc = (int32_t*)grow_zero(hc, (size_t)s1 * 4, (size_t)ns * 4);
            s1 = ns;
        }
        /* ssw.c:580-614: the direction buffer must stay below INT_MAX */
        ncell = (int64_t)width_d * readLen * 3;
        if (ncell >= ((int64_t)1 << 30)) {
            *length = *gapopen = *mismatch = -1;
            *identity = -1.0;
            free(hb); free(eb); free(hc); free(dirs);
            return 0;
        }
        if (ncell >= s2) {
            int64_t ns = s2;
            while (ncell >= ns) ns = grow_pow2(ns + 1);
            dirs = (int8_t*)grow_zero(dirs, (size_t)s2, (size_t)ns);
            s2 = ns;
        }
        for (j = 1; j < width - 1; ++j) hb[j] = 0;

        for (i = 0; i < readLen; ++i) {
            const int32_t beg = imax(0, i - w);
            const int32_t end = imin(refLen - 1, i + w);
            const int32_t edge = imin(end + 1, width - 1);
            int8_t* row = dirs + (size_t)width_d * i * 3;
            int32_t f = 0, u = 0;
            hb[0] = eb[0] = hb[edge] = eb[edge] = hc[0] = 0;
            for (j = beg; j <= end; ++j) {
                const int32_t up = band_slot(w, i - 1, j);       /* (i-1, j)   */
                const int32_t left = band_slot(w, i, j - 1);     /* (i, j-1)   */
                const int32_t diag = band_slot(w, i - 1, j - 1); /* (i-1, j-1) */
                const int32_t de = band_dir(w, i, j, 0);
                const int32_t df = band_dir(w, i, j, 1);
                const int32_t dh = band_dir(w, i, j, 2);
                int32_t open, ext, e1, f1, gap, sub;
                u = band_slot(w, i, j);

                open = (i == 0) ? -gapO : hb[up] - gapO;
                ext  = (i == 0) ? -gapE : eb[up] - gapE;
                eb[u] = open > ext ? open : ext;
                row[de] = open > ext ? 3 : 2;

                open = hc[left] - gapO;
                ext  = f - gapE;
                f = open > ext ? open : ext;
                row[df] = open > ext ? 5 : 4;

                e1 = imax(eb[u], 0);
                f1 = imax(f, 0);
                gap = e1 > f1 ? e1 : f1;
                sub = hb[diag] + mat[(size_t)ref[j] * nmat + read[i]];
                hc[u] = gap > sub ? gap : sub;
                if (hc[u] > best) best = hc[u];
                if (gap <= sub) row[dh] = 1;
                else row[dh] = e1 > f1 ? row[de] : row[df];
            }
            for (j = 1; j <= u; ++j) hb[j] = hc[j];
        }
        w *= 2;
    } while (best < score);
    w /= 2;

    /* Traceback from the bottom-right corner while i > 0 (ssw.c:663-739). */
    i = readLen - 1;
    j = refLen - 1;
    state = 2;
    op = prev_op = 'M';
    line = (int64_t)width_d * (readLen > 0 ? readLen - 1 : 0) * 3;
    while (i > 0) {
        const int64_t c = line + band_dir(w, i, j, state);
        const int8_t code = (c >= 0 && c < s2) ? dirs[c] : 0;
        switch (code) {
        case 1: --i; --j; state = 2; line -= width_d * 3; op = 'M'; break;
        case 2: --i; state = 0; line -= width_d * 3; op = 'I'; break;
        case 3: --i; state = 2; line -= width_d * 3; op = 'I'; break;
        case 4: --j; state = 1; op = 'D'; break;
        case 5: --j; state = 2; op = 'D'; break;
        default:
            free(hb); free(eb); free(hc); free(dirs);
            return 0;
        }
        /* counts use the cell just moved to (ssw.c:711-718) */
        if (op == 'M') {
            if (read[i] != ref[j]) (*mismatch)++;
            else identical++;
        }
        if ((op == 'I' && prev_op != 'I') || (op == 'D' && prev_op != 'D'))
            (*gapopen)++;
        (*length)++;
        prev_op = op;
    }
    *identity = identical / (double)*length;
    free(hb); free(eb); free(hc); free(dirs);
    return 1;
}

int orc_ssw_align_stats(const int8_t* read, int32_t readLen,
                        const int8_t* ref, int32_t refLen,
                        const int8_t* mat, int32_t nmat,
                        int32_t gapO, int32_t gapE,
                        int32_t read_end1, int32_t ref_end1, int32_t score1,
                        int32_t* read_begin1, int32_t* ref_begin1,
                        int32_t* length, int32_t* gapopen, int32_t* mismatch,
                        double* identity) {
    const int32_t bias = matrix_bias(mat, nmat);
    const int32_t rlen = read_end1 + 1;
    int8_t* rev = (int8_t*)malloc((size_t)(rlen > 0 ? rlen : 1));
    orc_end_t b;
    int32_t k, sub_ref, sub_read;
    (void)readLen; (void)refLen;

    /* Reversed query prefix against subject columns ref_end1..0
     * (ssw.c:927-940); the byte pass when bias + score1 < 255. */
    for (k = 0; k < rlen; ++k) rev[k] = read[read_end1 - k];
    if (bias + score1 < 255)
        b = orc_sw_scan(rev, rlen, ref, ref_end1 + 1, 1, mat, nmat, 16,
                        gapO, gapE, bias, (uint8_t)score1);
    else
        b = orc_sw_scan(rev, rlen, ref, ref_end1 + 1, 1, mat, nmat, 8,
                        gapO, gapE, bias, (uint16_t)score1);
    free(rev);

    *ref_begin1 = b.ref;
    *read_begin1 = read_end1 - b.read;
    sub_ref = ref_end1 - *ref_begin1 + 1;
    sub_read = read_end1 - *read_begin1 + 1;
    return orc_banded_sw(ref + *ref_begin1, read + *read_begin1,
                         sub_ref, sub_read, score1, gapO, gapE,
                         abs(sub_ref - sub_read) + 1, mat, nmat,
                         length, gapopen, mismatch, identity);
}

/* ------------------------------------------------------------------ */
/* Gapless HSP extension (search.c:99-190)                             */
/* ------------------------------------------------------------------ */
static int32_t pair_score(const int8_t* mat, char t, char q) {
    return mat[orc_aa_table2((unsigned char)t) * 25 +
               orc_aa_table2((unsigned char)orc_upper_aa_only(q))];
}

int orc_sa_hsp(const char* T, int64_t Tsize, int64_t idx,
               const char* Q, int32_t m, int32_t o,
               const int8_t* mat, int32_t* score) {
    int64_t right, left, i;
    int32_t sum, best_r = 0, len_r = 0, best_l = 0, len_l = 0;

    if (T[idx] != Q[o]) return -1;               /* search.c:113-116 */

    right = Tsize - idx;                         /* search.c:103, 119 */
    if ((int64_t)(m - o) < right) right = m - o;
    if (right > 128) right = 128;
    sum = 0;
    for (i = 0; i < right; ++i) {                /* search.c:126-149 */
        if (T[idx + i] == '\n') break;
        sum += pair_score(mat, T[idx + i], Q[o + i]);
        if (sum >= best_r) { best_r = sum; len_r = (int32_t)i; }
    }

    left = idx < o ? idx : o;                    /* search.c:104, 158 */
    if (left > 128) left = 128;
    sum = 0;
    for (i = 1; i < left; ++i) {                 /* search.c:161-183 */
        if (T[idx - i] == '\n') break;
        sum += pair_score(mat, T[idx - i], Q[o - i]);
        if (sum >= best_l) { best_l = sum; len_l = (int32_t)i; }
    }

    *score = best_l + best_r;
    return len_l + len_r;
}

int orc_sa_gapless(const char* T, int64_t Tsize, int64_t idx,
                   const char* Q, int32_t m, int32_t o,
                   const int8_t* mat, int32_t* score) {
    int64_t len = Tsize - idx, i;                /* search.c:52 */
    int32_t sum = 0, best = 0, arg = 0;
    if ((int64_t)(m - o) < len) len = m - o;
    if (len < 3 || T[idx] != Q[o]) return -1;    /* search.c:62-65 */
    for (i = 0; i < len; ++i) {
        if (T[idx + i] == '\n') break;
        sum += pair_score(mat, T[idx + i], Q[o + i]);
        if (sum >= best) { best = sum; arg = (int32_t)i; }
    }
    *score = best;
    return arg;
}

/* ------------------------------------------------------------------ */
/* Karlin-Altschul (ssw_wrapper.c:179-183, 239-245)                    */
/* ------------------------------------------------------------------ */
void orc_evalue(int32_t rawscore, int32_t qlen, uint64_t db_size,
                uint64_t db_proteins, double lambda, double k, double H,
                double* bitscore, double* evalue, int32_t* effective_length) {
    double correction = log(k * qlen * db_size) / H;
    double cq = fmax(qlen - correction, 1.0);
    double cdb = fmax(db_size - (correction * db_proteins), 1.0);
    double nats = (lambda * rawscore - log(k));
    double ev = cq * cdb * exp(-nats);
    *evalue = ev > 1.0e-180 ? ev : 0.0;
    *bitscore = nats / M_LN2;
    *effective_length = (int32_t)(qlen - correction);
}

/* ------------------------------------------------------------------ */
/* Textbook Gotoh local alignment with the same column bookkeeping     */
/* ------------------------------------------------------------------ */
/* H = max(0, H(i-1,j-1) + W, E, F), E and F from H, gap of length k
 * costing gapO + (k-1) gapE.  Same end rules as orc_sw_scan (first column
 * reaching the max, lowest row of that column; byte/word overflow caps).
 * Used by the tests to show that for gapO > gapE the reference's lane
 * recurrence gives the same H in every cell. */
orc_end_t orc_sw_textbook(const int8_t* read, int32_t readLen,
                          const int8_t* ref, int32_t refLen,
                          const int8_t* mat, int32_t nmat, int P,
                          int32_t gapO, int32_t gapE, int32_t bias) {
    orc_end_t r;
    const int32_t m = readLen, n = refLen;
    const int32_t cap = (P == 16) ? 255 - bias : 32767;
    const int32_t ovf = (P == 16) ? 255 : 32767;
    const size_t cells = (size_t)(m > 0 ? m : 1);
    int32_t* hp = (int32_t*)calloc(cells, sizeof(int32_t));
    int32_t* hc = (int32_t*)calloc(cells, sizeof(int32_t));
    int32_t* e = (int32_t*)calloc(cells, sizeof(int32_t));
    int32_t* kept = (int32_t*)calloc(cells, sizeof(int32_t));
    int32_t best = 0, end_ref = (P == 16) ? -1 : 0, end_read, i, j;
    for (i = 0; i < n; ++i) {
        const int8_t* wrow = mat + (size_t)ref[i] * nmat;
        int32_t f = 0, colmax = 0, *t;
        for (j = 0; j < m; ++j) {
            int32_t d = j ? hp[j - 1] : 0;
            int32_t h = imax(imax(d + wrow[read[j]], 0), imax(e[j], f));
            int32_t hgo = imax(h - gapO, 0);
            e[j] = imax(imax(e[j] - gapE, 0), hgo);
            f = imax(imax(f - gapE, 0), hgo);
            hc[j] = h;
            colmax = imax(colmax, h);
        }
        t = hp; hp = hc; hc = t;
        if (colmax >= cap) colmax = cap;
        if (colmax > best) {
            best = colmax;
            if (best >= cap) break;
            end_ref = i;
            memcpy(kept, hp, (size_t)m * sizeof(int32_t));
        }
    }
    end_read = m - 1;
    for (j = 0; j < m; ++j) if (kept[j] == best) { end_read = j; break; }
    r.score = best >= cap ? ovf : best;
    r.ref = end_ref;
    r.read = m > 0 ? end_read : -1;
    free(hp); free(hc); free(e); free(kept);
    return r;
}

void orc_sa_hsp_batch(const char* T, int64_t Tsize, const char* Qcat,
                      const int64_t* q_off, const int32_t* q_len,
                      const orc_seed* seeds, int64_t n, const int8_t* mat,
                      int32_t* score, int32_t* len) {
    for (int64_t k = 0; k < n; ++k) {
        const orc_seed* s = &seeds[k];
        len[k] = orc_sa_hsp(T, Tsize, s->position, Qcat + q_off[s->query],
                            q_len[s->query], s->offset, mat, &score[k]);
    }
}


/* ---- suffix array (index.c:66) ------------------------------------------ */

typedef struct {
    int64_t pos;
    int64_t r1, r2;        /* rank of the first h bytes, of the next h */
} orc_sa_rec;

static int orc_sa_cmp(const void* a, const void* b) {
    const orc_sa_rec* x = (const orc_sa_rec*) a;
    const orc_sa_rec* y = (const orc_sa_rec*) b;
    if (x->r1 != y->r1) return x->r1 < y->r1 ? -1 : 1;
    if (x->r2 != y->r2) return x->r2 < y->r2 ? -1 : 1;
    return 0;
}

int orc_suffix_array(const unsigned char* T, int64_t n, int64_t* SA) {
    orc_sa_rec* rec;
    int64_t* rank;
    int64_t h, k;
    if (n <= 0) return 0;
    rec = (orc_sa_rec*) malloc(sizeof(orc_sa_rec) * (size_t) n);
    rank = (int64_t*) malloc(sizeof(int64_t) * (size_t) n);
    if (!rec || !rank) { free(rec); free(rank); return -1; }
    /* rank 0 past the end: a prefix sorts before the longer suffix */
    for (k = 0; k < n; ++k) rank[k] = (int64_t) T[k] + 1;
    for (h = 1;; h *= 2) {
        int64_t groups = 1;
        for (k = 0; k < n; ++k) {
            rec[k].pos = k;
            rec[k].r1 = rank[k];
            rec[k].r2 = k + h < n ? rank[k + h] : 0;
        }
        qsort(rec, (size_t) n, sizeof(orc_sa_rec), orc_sa_cmp);
        rank[rec[0].pos] = 1;
        for (k = 1; k < n; ++k) {
            if (orc_sa_cmp(&rec[k], &rec[k - 1]) != 0) ++groups;
            rank[rec[k].pos] = groups;
        }
        if (groups == n || h >= n) break;
    }
    for (k = 0; k < n; ++k) SA[k] = rec[k].pos;
    free(rec);
    free(rank);
    return 0;
}

/* ---- sa_search (search.c:17-30, 237-361) ---------------------------------- */

static int orc_sa_compare(const char* T, int64_t Tsize, const char* P, int64_t Psize,
                          int64_t suf, int64_t* match) {
    int64_t i = suf + *match, j = *match;
    int r = 0;
    /* the reference subtracts (signed) chars */
    for (; i < Tsize && j < Psize && (r = (int) (signed char) T[i] - (int) (signed char) P[j]) == 0;
         ++i, ++j) {}
    *match = j;
    return r == 0 ? -(j != Psize) : r;
}

int orc_sa_search(const char* T, int64_t Tsize, const int64_t* SA,
                  const char* P, int64_t Psize, int64_t* idx,
                  int64_t* num_matches, int64_t* last_position) {
    int64_t size, lsize, rsize, half;
    int64_t match, lmatch, rmatch, llmatch, lrmatch, rlmatch, rrmatch;
    int64_t i, j, k;
    int r;
    if (idx) *idx = -1;
    if (!T || !P || !SA || Tsize <= 0 || Psize <= 0) return -1;
    for (i = j = k = 0, match = lmatch = rmatch = 0, size = Tsize, half = size >> 1; 0 < size;
         size = half, half >>= 1) {
        match = lmatch < rmatch ? lmatch : rmatch;
        r = orc_sa_compare(T, Tsize, P, Psize, SA[i + half], &match);
        if (r < 0) {                      /* the upper half */
            i += half + 1;
            half -= (size & 1) ^ 1;
            lmatch = match;
        } else if (r > 0) {               /* the lower half */
            rmatch = match;
        } else {                          /* a match: widen to the interval */
            lsize = half, j = i, rsize = size - half - 1, k = i + half + 1;
            for (llmatch = lmatch, lrmatch = match, half = lsize >> 1; 0 < lsize;
                 lsize = half, half >>= 1) {
                lmatch = llmatch < lrmatch ? llmatch : lrmatch;
                r = orc_sa_compare(T, Tsize, P, match, SA[j + half], &lmatch);
                if (r < 0) {
                    j += half + 1;
                    half -= (lsize & 1) ^ 1;
                    llmatch = lmatch;
                } else {
                    lrmatch = lmatch;
                }
            }
            for (rlmatch = match, rrmatch = rmatch, half = rsize >> 1; 0 < rsize;
                 rsize = half, half >>= 1) {
                rmatch = rlmatch < rrmatch ? rlmatch : rrmatch;
                r = orc_sa_compare(T, Tsize, P, match, SA[k + half], &rmatch);
                if (r <= 0) {
                    k += half + 1;
                    half -= (rsize & 1) ^ 1;
                    rlmatch = rmatch;
                } else {
                    rrmatch = rmatch;
                }
            }
            break;
        }
    }
    if (idx) *idx = (0 < (k - j)) ? j : i;
    if (num_matches) *num_matches = k - j;
    if (last_position) *last_position = match;
    return 0;
}
