/*
 * orc.h -- CPU oracle for TOPAZ's candidate-alignment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product (topaz_amd) never links or calls it.
 *
 * Every function restates, in plain scalar C written for clarity, one step of
 * the reference algorithm; each cites the reference file:line it follows.
 * Residues are matrix indices 0..24 (the order of the substitution.h
 * matrices); W(s, q) = mat[s * nmat + q] with s the subject (reference)
 * residue and q the query (read) residue, as in ssw.c:112 and search.c:134.
 *
 * Pinning (how each function is checked, tests/test_oracle.py):
 *   orc_sw_scan / orc_ssw_align  -- against the reference's own compiled
 *        ssw.c (oracle/_ref) on seeded random, homologous and adversarial
 *        pairs (committed fixtures), plus brute-force textbook Gotoh bounds.
 *   orc_ssw_align_stats / orc_banded_sw -- against the compiled reference
 *        (fixtures) on reported-hit-like pairs.
 *        The shared-buffer replay of banded_sw is also stress-checked against
 *        the compiled reference (stats of 4,200 random and homologous pairs).
 *   orc_sw_textbook -- brute-force Gotoh in numpy on small pairs.
 *   orc_sa_hsp -- against compiled sa_hsp (fixtures) and brute force.
 *   orc_sa_hsp_batch -- equal to orc_sa_hsp seed by seed.
 *   orc_sa_gapless -- brute force, and the compiled reference's sa_gapless.
 *   orc_evalue -- closed form, against compiled align_using_ssw_profile.
 *   orc_aa_table2 / orc_upper_aa_only -- against the committed tables.
 *   orc_suffix_array -- brute-force sorting of the suffixes (small texts)
 *        and the reference's compiled divsufsort64.
 *   orc_sa_search -- the reference's compiled sa_search, offset by offset.
 */
#ifndef TOPAZ_ORACLE_ORC_H
#define TOPAZ_ORACLE_ORC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result of one striped Smith-Waterman scan (alignment_end, ssw.c:68-72). */
typedef struct {
    int32_t score;   /* reported score (255 / 32767 on overflow)          */
    int32_t ref;     /* end position on the subject (column)               */
    int32_t read;    /* end position on the query (row)                    */
} orc_end_t;

/*
 * One striped SW scan with the reference's lane semantics.
 *   P       16 (byte path, sw_sse2_byte ssw.c:127-332) or 8 (word path,
 *           sw_sse2_word ssw.c:358-536).  Only the stripe length
 *           segLen = ceil(m/P) and the overflow rules depend on it.
 *   dir     0: columns 0..n-1; 1: columns n-1..0 (ref_dir, ssw.c:182-186).
 *   bias    byte path bias (-min(mat)); ignored for P=8.
 *   terminate  stop after the first column whose max equals it
 *           (ssw.c:285, 488); pass -1 for "never".
 */
orc_end_t orc_sw_scan(const int8_t* read, int32_t readLen,
                      const int8_t* ref, int32_t refLen, int dir,
                      const int8_t* mat, int32_t nmat, int P,
                      int32_t gapO, int32_t gapE, int32_t bias,
                      int32_t terminate);

/* The closed-form recurrence the GPU kernels implement (valid for
 * gapO > gapE, where it equals orc_sw_scan):
 *   Hp = max(0, H(i-1,j-1) + W, E, Fin);  H = max(Hp, F)
 *   E' = max(E - gapE, Hp - gapO);  Fin' / F' = max(. - gapE, Hp - gapO)
 * with Fin restarted at rows j % segLen == 0 and F never restarted. */
orc_end_t orc_sw_scan_s1(const int8_t* read, int32_t readLen,
                         const int8_t* ref, int32_t refLen, int dir,
                         const int8_t* mat, int32_t nmat, int P,
                         int32_t gapO, int32_t gapE, int32_t bias,
                         int32_t terminate);

/* Textbook Gotoh (E and F from H) with the same bookkeeping and caps. */
orc_end_t orc_sw_textbook(const int8_t* read, int32_t readLen,
                          const int8_t* ref, int32_t refLen,
                          const int8_t* mat, int32_t nmat, int P,
                          int32_t gapO, int32_t gapE, int32_t bias);

/* ssw_align (ssw.c:835-897) with score_size 2: byte, then word on 255.
 * Returns 0 and fills *out, or 1 where the reference returns NULL
 * (word score 32767). */
int orc_ssw_align(const int8_t* read, int32_t readLen,
                  const int8_t* ref, int32_t refLen,
                  const int8_t* mat, int32_t nmat,
                  int32_t gapO, int32_t gapE, orc_end_t* out);

/* banded_sw (ssw.c:538-784): returns 1 on success, 0 on the guarded
 * failures, with the reference's partial/-1 outputs. */
int orc_banded_sw(const int8_t* ref, const int8_t* read,
                  int32_t refLen, int32_t readLen, int32_t score,
                  int32_t gapO, int32_t gapE, int32_t band_width,
                  const int8_t* mat, int32_t nmat,
                  int32_t* length, int32_t* gapopen, int32_t* mismatch,
                  double* identity);

/* ssw_align_stats (ssw.c:900-978): reverse pass + banded traceback. */
int orc_ssw_align_stats(const int8_t* read, int32_t readLen,
                        const int8_t* ref, int32_t refLen,
                        const int8_t* mat, int32_t nmat,
                        int32_t gapO, int32_t gapE,
                        int32_t read_end1, int32_t ref_end1, int32_t score1,
                        int32_t* read_begin1, int32_t* ref_begin1,
                        int32_t* length, int32_t* gapopen, int32_t* mismatch,
                        double* identity);

/* sa_hsp (search.c:99-190) for a seed at text position idx = SA[p] and
 * query offset o.  T and Q are internal-alphabet characters (not matrix
 * indices).  Returns length (-1 when T[idx] != Q[o]); *score is written
 * only when length != -1, as in the reference. */
int orc_sa_hsp(const char* T, int64_t Tsize, int64_t idx,
               const char* Q, int32_t m, int32_t o,
               const int8_t* mat, int32_t* score);

/* orc_sa_hsp over many seeds {query, offset, idx} (16-byte records, the
 * layout of the GPU seed batch) against one text: the loop of sa_hsp calls
 * query_find_suffixes and suffix_increment make (query.c:117,125,
 * suffix.c:47).  Qcat holds the queries back to back at q_off[], q_len[].
 * score[k] is written only when len[k] != -1. */
typedef struct { int32_t query; int32_t offset; int64_t position; } orc_seed;
void orc_sa_hsp_batch(const char* T, int64_t Tsize, const char* Qcat,
                      const int64_t* q_off, const int32_t* q_len,
                      const orc_seed* seeds, int64_t n, const int8_t* mat,
                      int32_t* score, int32_t* len);

/* sa_gapless (search.c:48-97), dead in the reference; kept for parity. */
int orc_sa_gapless(const char* T, int64_t Tsize, int64_t idx,
                   const char* Q, int32_t m, int32_t o,
                   const int8_t* mat, int32_t* score);

/* Karlin-Altschul statistics of align_using_ssw_profile
 * (ssw_wrapper.c:179-183, 239-245). */
void orc_evalue(int32_t rawscore, int32_t qlen, uint64_t db_size,
                uint64_t db_proteins, double lambda, double k, double H,
                double* bitscore, double* evalue, int32_t* effective_length);

/* Suffix array of T[0..n-1] in divsufsort64's order (index.c:66): suffixes
 * as unsigned bytes, a suffix before every longer suffix it is a prefix of.
 * Prefix doubling with qsort (the suffix array is unique, so the method is
 * free).  Returns 0, or -1 when memory runs out. */
int orc_suffix_array(const unsigned char* T, int64_t n, int64_t* SA);

/* sa_search (search.c:237-361, without USE_LCP) for the suffix P[0..Psize-1]
 * of a query: the same binary search, step for step.  Returns 0 with *idx,
 * *num_matches, *last_position, or -1 (and *idx = -1) for empty inputs. */
int orc_sa_search(const char* T, int64_t Tsize, const int64_t* SA,
                  const char* P, int64_t Psize, int64_t* idx,
                  int64_t* num_matches, int64_t* last_position);

/* Alphabet maps (lexicographical.h:19-74): internal char -> matrix index. */
int8_t orc_aa_table2(unsigned char c);
char   orc_upper_aa(char c);
char   orc_upper_aa_only(char c);

#ifdef __cplusplus
}
#endif
#endif
