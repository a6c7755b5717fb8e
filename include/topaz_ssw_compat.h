/*
 * topaz_ssw_compat.h -- drop-in ABI for TOPAZ's alignment boundary.
 *
 * libtopaz_amd.so exports these functions with the reference's names,
 * signatures and struct layouts, so TOPAZ's own query.c / suffix.c / engine.c
 * link against it in place of ssw.o, ssw_wrapper.o, search.o and
 * alignment.o.  Every call on TOPAZ's search path computes on the GPU; the
 * functions TOPAZ never calls (the last block below) run on the host.
 * Errors end the process with a message,
 * as the reference's own fatal paths do (query.c:311-314): these calls have
 * no error return.
 *
 * The struct definitions below restate the reference's layouts field for
 * field (same order and types), because callers compiled against the
 * reference headers pass them in.  A program that includes the reference
 * headers should not include this file.
 */
#ifndef TOPAZ_SSW_COMPAT_H
#define TOPAZ_SSW_COMPAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ssw.h:31-41.  The profile arrays are built on the device per call; the
 * two pointers are non-null markers.  readLen is read by callers
 * (ssw_wrapper.c:179, 245). */
typedef struct {
    void* profile_byte;
    void* profile_word;
    int8_t* read;
    const int8_t* mat;
    int32_t readLen;
    int32_t n;
    uint8_t bias;
} s_profile;

/* ssw.h:57-67 */
typedef struct {
    uint16_t score1;
    uint16_t score2;
    int32_t ref_begin1;
    int32_t ref_end1;
    int32_t read_begin1;
    int32_t read_end1;
    int32_t ref_end2;
    uint32_t* cigar;
    int32_t cigarLen;
} s_align;

/* seq.h:12-31 */
typedef struct {
    char* buf;
    int len;
} buffer_t;

typedef struct {
    buffer_t id_buf;
    buffer_t seq_buf;
    int id_len;
    int seq_len;
} seq_t;

/* options.h:26-53 */
typedef struct {
    int num_threads;
    int num_hits;
    double evalue;
    int smithwaterman;
    int gap_open;
    int gap_extend;
    char* sub_name;
    int8_t (*substitution)[625];
    uint64_t db_size;
    uint64_t db_proteins;
    double lambda;
    double k;
    double H;
    int min_hit_length;
    int min_hit_score;
    int number_of_alignments;
    int number_of_seeds;
    int seg_enabled;
    int superfast;
} options_t;

/* hit.h:74-107 (priority_node is an rbnode_t* in the reference) */
typedef struct {
    uint64_t db_index;
    int is_aligned;
    int alignment_rank;
    int rawscore;
    int mismatch;
    int length;
    int gapopen;
    double identity;
    double bitscore;
    double evalue;
    int qstart;
    int qend;
    int sstart;
    int send;
    int effective_length;
    int gapless_score;
    int gapless_length;
    int suffix_counter;
    int suffix_offset;
    int suffix_multiple;
    double time_stats;
    double time_align;
    void* priority_node;
} hit_t;

/* db.h:10-39 (built without USE_LCP) */
typedef struct {
    uint64_t header_start;
    uint64_t protein_start;
    int header_len;
    int protein_len;
} position_t;

typedef struct {
    char* prefix;
    char* headers;
    uint64_t Hsize;
    position_t* index;
    uint64_t Psize;
    char* T;
    uint64_t Tsize;
    int64_t* SA;
    uint64_t SAsize;
} db_t;

/* ssw.h:87-210 -- ssw_init (ssw.c:799), init_destroy / init_destroy2
 * (ssw.c:822-833), ssw_align (ssw.c:835), align_destroy (ssw.c:980),
 * ssw_align_stats (ssw.c:900).  score_size must be 2 (what TOPAZ passes,
 * ssw_wrapper.c:127); gap penalties 0..255 (uint8_t there), gapO <= gapE
 * included. */
s_profile* ssw_init(const int8_t* read, const int32_t readLen, const int8_t* mat,
                    const int32_t n, const int8_t score_size);
void init_destroy(s_profile* p);
void init_destroy2(s_profile* p);
s_align* ssw_align(const s_profile* prof, const int8_t* ref, int32_t refLen,
                   const uint8_t weight_gapO, const uint8_t weight_gapE,
                   const uint8_t flag, const uint16_t filters,
                   const int32_t filterd, const int32_t maskLen);
void align_destroy(s_align* a);

/* ssw.h:23-26, 148-194 -- the BAM encoding of one CIGAR operation (SAM
 * format specification: length << 4 | op, with the ops "MIDNSHP=X" numbered
 * 0..8).  An unknown letter encodes as 'M', and op codes above 8 decode as
 * 'M'.  No function here produces a CIGAR (ssw_align runs with flag 0 on
 * TOPAZ's path); the helpers serve callers that include this header in place
 * of ssw.h. */
#ifndef MAPSTR
#define MAPSTR "MIDNSHP=X"
#endif
#ifndef BAM_CIGAR_SHIFT
#define BAM_CIGAR_SHIFT 4
#endif
static inline uint32_t to_cigar_int(uint32_t length, char op_letter) {
    uint32_t op = 0;
    for (uint32_t k = 0; k < sizeof(MAPSTR) - 1; ++k)
        if (MAPSTR[k] == op_letter) op = k;
    return (length << BAM_CIGAR_SHIFT) | op;
}
static inline char cigar_int_to_op(uint32_t cigar_int) {
    const uint32_t op = cigar_int & 0xfu;
    return op < sizeof(MAPSTR) - 1 ? MAPSTR[op] : 'M';
}
static inline uint32_t cigar_int_to_len(uint32_t cigar_int) {
    return cigar_int >> BAM_CIGAR_SHIFT;
}

int ssw_align_stats(const s_profile* prof, const int8_t* ref, int32_t refLen,
                    const uint8_t weight_gapO, const uint8_t weight_gapE,
                    const int32_t maskLen, int32_t read_end1, int32_t ref_end1,
                    uint16_t score1, int32_t* read_begin1, int32_t* ref_begin1,
                    int32_t* length, int32_t* gapopen, int32_t* mismatch,
                    double* identity);

/* ssw_wrapper.h:11-14 -- ssw_wrapper.c:105-130, 157-317, 132-155, 319-435 */
s_profile* get_ssw_profile(seq_t* s1, options_t* opt);
void align_using_ssw_profile(s_profile* profile, seq_t* s2, hit_t* hit, options_t* opt);
void stats_using_ssw_profile(s_profile* profile, seq_t* s2, hit_t* hit, options_t* opt);
void align_using_ssw(seq_t* s1, seq_t* s2, hit_t* hit, options_t* opt);

/* search.h:14 -- search.c:99-190.  db->T and db->SA are uploaded once per
 * db_t and stay resident on the device. */
int sa_hsp(db_t* db, seq_t* q, int qoffset, int64_t position, options_t* opt,
           int* gapless_score, int* gapless_length);

/* search.h:15 -- search.c:237-361: the suffix-array interval of the pattern
 * q[offset .. offset + Psize), Psize = length > 0 ? length : seq_len -
 * offset, searched on the device over the resident db->T and db->SA.
 * Returns -1 with *idx = -1 (the other outputs untouched) for an empty
 * database or Psize <= 0; a pattern outside the query's buffer (which the
 * reference reads anyway) ends the process. */
int sa_search(db_t* db, seq_t* q, int64_t offset, int64_t length, int64_t* idx,
              int64_t* num_matches, int64_t* last_position);

/* ---- batching extension (no reference counterpart) --------------------
 * One call per query instead of one call per hit, with the reference's own
 * types.  Subjects are read from the database (db_seq, db.c:233-247: the
 * protein's bytes through upper_aa).  A patched query.c calls these in place
 * of its per-hit loops. */

/* The loop of query_perform_alignments (query.c:199-228): every hit gets
 * exactly what align_using_ssw_profile writes into it.  Returns how many hits
 * have evalue < opt->evalue (what perform_alignment_with_profile returns per
 * hit, query.c:338).  alignment_rank is left to the caller.  The first call
 * for a database (and matrix and gap penalties) puts all its proteins on the
 * device; calls from several threads at once share GPU launches. */
int align_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                  options_t* opt);

/* The loop of query_perform_statistics (query.c:276-291): every hit gets
 * what stats_using_ssw_profile writes (hits with rawscore 32767 are left as
 * they are, ssw_wrapper.c:136-138). */
void stats_batch_using_ssw_profile(s_profile* profile, db_t* db, hit_t** hits, int n,
                                   options_t* opt);

/* n calls of sa_hsp for one query: seed k is (qoffsets[k], positions[k]);
 * lengths[k] is sa_hsp's *gapless_length and scores[k] its *gapless_score,
 * written only where the length is not -1 (search.c:113-116). */
void sa_hsp_batch(db_t* db, seq_t* q, int n, const int* qoffsets, const int64_t* positions,
                  options_t* opt, int* scores, int* lengths);

/* db.c:176-205: the protein whose letters (or its '\n') hold text position
 * pos: 0 and *index set, or -1 for pos >= db->Tsize.  For an index that
 * tiles the text as `topaz index` writes it, a lookup in a table of every
 * 64th position's protein (built once per database) in place of the binary
 * search, which every hit of the seed walk pays (hit.c:138); the same
 * protein, as it is unique.  Link with db.o's own db_index hidden. */
int db_index(db_t* db, uint64_t pos, int64_t* index);

/* n calls of sa_search(db, q, offsets[k], 0, ...) (search.c:237-361) for one
 * query, run on the GPU: idx[k], num_matches[k] and last_position[k] as that
 * call writes them (idx[k] = -1 and the other two untouched where it
 * returns -1).  Returns 0, or -1 when any call would.  num_matches and
 * last_position may be null. */
int sa_search_batch(db_t* db, seq_t* q, int n, const int* offsets, int64_t* idx,
                    int64_t* num_matches, int64_t* last_position);

/* ---- exported by the reference, never called by TOPAZ ----------------
 * Single scalar computations on the caller's host arrays, run on the host
 * with the reference's results (tie rules and fatal paths included), so a
 * program links with neither search.o nor alignment.o. */

/* search.h:12 -- search.c:32-46: letters of suffix SA[position] and the
 * query from qoffset that agree (at most seq_len - qoffset once the first
 * agrees). */
int sa_lcp(db_t* db, seq_t* q, int qoffset, int64_t position);

/* search.h:13 -- search.c:48-97: right-only gapless extension of a seed;
 * *gapless_length = -1 when fewer than 3 letters remain or the first
 * letters differ. */
int sa_gapless(db_t* db, seq_t* q, int qoffset, int64_t position, options_t* opt,
               int* gapless_score, int* gapless_length);

/* alignment.h:32 -- alignment.c:266-287: the score of s1[0 .. length)
 * against itself (s2 unused, as in the reference). */
int ungapped_rawscore(char* s1, char* s2, uint32_t length, options_t* opt);

/* alignment.c:289-315: returns max_length (the reference's scan never
 * runs). */
int ungapped_length(char* s1, char* s2, uint32_t max_length, options_t* opt);

/* alignment.h:31 -- alignment.c:317-396: full-matrix Gotoh of s1 (columns)
 * against s2 (rows) with the trace codes of alignment.h:12-21, the hit's
 * coordinates, counts and E-value; aborts, as the reference does, when no
 * cell scores above 0. */
int gotoh(seq_t* s1, seq_t* s2, hit_t* hit, options_t* p);

/* alignment.h:24-30 and alignment.c:27-264: gotoh's steps. */
int _forward_pass(int8_t* s1, int8_t* s2, size_t s1_len, size_t s2_len, int* E, uint16_t* trE,
                  int* F, uint16_t* trF, int* H, uint16_t* trH, hit_t* hit, options_t* opt);
void _backtrack(uint16_t* trE, uint16_t* trF, uint16_t* trH, int pos, int max_x, hit_t* hit);
void _initialise(seq_t* s1, seq_t* s2, int* E, int* F, int* H, uint16_t* trE, uint16_t* trF,
                 uint16_t* trH, options_t* par);
void printdp32(int* m, size_t dim1, size_t dim2);
void printdp8(uint8_t* m, size_t dim1, size_t dim2);

#ifdef __cplusplus
}
#endif
#endif
