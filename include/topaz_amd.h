/*
 * topaz_amd.h -- C ABI of the MI355X implementation of TOPAZ's
 * candidate-alignment hot path (gapless HSP extension + striped
 * Smith-Waterman), built as libtopaz_amd.so.
 *
 * Plain C types only.  Every call returns 0 on success and a nonzero code on
 * failure; tpz_last_error() gives the message.  There is no CPU fallback: the
 * library needs a HIP device and fails if none is usable.
 *
 * Residues are substitution-matrix indices 0..24 (the residue order
 * ARNDCQEGHILKMFPSTWYVBJZX* of substitution.h); the matrix is 25 x 25 int8,
 * W[s * 25 + q] with s the subject (database) residue and q the query residue,
 * as ssw.c:112 and search.c:134 index it.
 */
#ifndef TOPAZ_AMD_H
#define TOPAZ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */

/* Select the HIP device used by this process (default 0) and create its
 * stream.  Called implicitly by every entry point with device 0. */
int tpz_init(int device);
const char* tpz_last_error(void);
/* Library version string. */
const char* tpz_version(void);

/* ---- batched forward Smith-Waterman (score + end positions) -------------- */

/* A device-resident batch of (query, subject) pairs.  Replaces the per-pair
 * loop of query_perform_alignments (query.c:199-228) that calls
 * align_using_ssw_profile -> ssw_align (ssw_wrapper.c:209, ssw.c:835-897). */
typedef struct tpz_sw_batch tpz_sw_batch;

/* Copy queries, subjects and the pair list to the device.
 *   q_res/q_off/q_len   queries as residues; query k = q_res[q_off[k] ..
 *                       q_off[k] + q_len[k])
 *   s_res/s_off/s_len   subjects, same layout
 *   pair_query/pair_subject  n_pairs pair definitions
 *   matrix              25 x 25 scores; gap penalties 0..255 (with gap_open <=
 *                       gap_extend the reference's lanes run literally, slower)
 * Per pair the result is what ssw_align (ssw.c:835-897, score_size 2) returns:
 * the byte (16-lane) result if its score + bias < 255, else the word (8-lane)
 * result, else "overflow" (the reference returns NULL). */
int tpz_sw_batch_create(tpz_sw_batch** out,
                        int32_t n_queries, const int8_t* q_res,
                        const int64_t* q_off, const int32_t* q_len,
                        int32_t n_subjects, const int8_t* s_res,
                        const int64_t* s_off, const int32_t* s_len,
                        int32_t n_pairs, const int32_t* pair_query,
                        const int32_t* pair_subject, const int8_t* matrix,
                        int32_t gap_open, int32_t gap_extend);

/* Launch schedules of tpz_sw_batch_run. */
#define TPZ_SCHED_PLAIN      0   /* one launch per query                 */
#define TPZ_SCHED_PERSISTENT 1   /* one persistent launch, length-binned */

/* Run the batch on the device (results stay on the device). */
int tpz_sw_batch_run(tpz_sw_batch* b, int32_t schedule);

/* Copy results to the host.  score[i] is the ssw_align score1 of pair i, or
 * 32767 where the reference returns NULL (ssw.c:866-870); ref_end/read_end
 * are ref_end1/read_end1 (ssw.c:886-887).  Any pointer may be null. */
int tpz_sw_batch_results(const tpz_sw_batch* b, int32_t* score,
                         int32_t* ref_end, int32_t* read_end);

/* Device time of the kernels of the last run, in milliseconds (HIP events on
 * the library stream). */
int tpz_sw_batch_kernel_ms(const tpz_sw_batch* b, double* ms);

/* Launch configuration of a batch.  No reference counterpart: these are the
 * tile shape and schedule knobs of the GPU kernel.  Results never depend on
 * them (bit-identical for every configuration). */
typedef struct {
    int32_t rows_per_chunk;    /* query rows a wavefront holds in registers: 16 or 32  */
    int32_t groups_per_cu;     /* resident workgroups per CU (persistent): 1 up to what
                                  the kernel's registers and LDS allow (on gfx950 2 with
                                  4 wavefronts, 1 with 8)                               */
    int32_t queries_per_task;  /* queries whose pairs may share a wavefront: 1, 2, 4   */
    int32_t use_graph;         /* nonzero: replay the launches from a captured hipGraph */
    double big_fraction;       /* a task costing more than this fraction of one
                                  wavefront's share of the batch is split over the
                                  wavefronts of a workgroup (chunk staircase); 0: all
                                  tasks                                                 */
    int32_t waves_per_group;   /* wavefronts per workgroup: 4 or 8 (8 halves a
                                  staircase task's time, one workgroup per CU)         */
} tpz_sw_config;

/* The configuration the batch runs with (defaults: 32, 2, 4, 0, 0.5). */
int tpz_sw_batch_get_config(const tpz_sw_batch* b, tpz_sw_config* cfg);
/* Re-plan the batch for cfg: rebuilds tasks, work items and workspace. */
int tpz_sw_batch_set_config(tpz_sw_batch* b, const tpz_sw_config* cfg);
/* Time the persistent schedule under each candidate configuration on this
 * batch (one warm-up run, then the faster of two timed runs each), check
 * that every candidate gives bit-identical results, and keep the fastest.
 * use_graph of the kept configuration is what the batch had.  best / best_ms
 * may be null; n_tried receives the number of candidates timed. */
int tpz_sw_batch_autotune(tpz_sw_batch* b, tpz_sw_config* best, double* best_ms,
                          int32_t* n_tried);

void tpz_sw_batch_destroy(tpz_sw_batch* b);

/* ---- serving loop: continuous batching over a resident database -------- */

/* A request loop for stage S.  The database subjects stay resident on the
 * device; requests (a set of queries and their candidate pairs) arrive from
 * any number of host threads.  A scheduler thread admits waiting requests in
 * arrival order while their device memory fits the budget, merges everything
 * admitted at that moment into one persistent launch on a free stream, and
 * lets launches on different streams overlap on the device, so requests that
 * arrive while others run join the next launch instead of queueing behind the
 * whole batch.  Per-pair results are independent of which requests share a
 * launch.  No reference counterpart: the reference runs one thread per query
 * over a length-keyed work queue (work_queue.c:317-352, engine.c:423-457). */
typedef struct tpz_server tpz_server;

typedef struct {
    uint64_t memory_budget;   /* bytes of admitted request data (queries, pairs, task
                                 lists, results) on the device at once; 0: 1 GiB     */
    int32_t max_inflight;     /* launches in flight (one stream each); 0: 2          */
    int32_t max_launch_pairs; /* pairs merged into one launch at most; 0: 4,000,000  */
    int32_t waves_per_group;  /* wavefronts per workgroup, 4 or 8; 0: 8 (latency)     */
} tpz_server_config;

typedef struct {
    uint64_t requests_done;
    uint64_t launches;
    uint64_t max_requests_per_launch;
    uint64_t pairs_done;
    uint64_t peak_bytes;      /* largest admitted request bytes in flight            */
    uint64_t workspace_bytes; /* fixed device workspace of the launch slots          */
} tpz_server_stats;

/* Upload the database subjects (same layout and conventions as
 * tpz_sw_batch_create) and start the scheduler thread.  cfg may be null. */
int tpz_server_create(tpz_server** out, int32_t n_subjects, const int8_t* s_res,
                      const int64_t* s_off, const int32_t* s_len, const int8_t* matrix,
                      int32_t gap_open, int32_t gap_extend, const tpz_server_config* cfg);
/* Queue a request: n_queries queries (residues, offsets, lengths) and n_pairs
 * pairs (query number within the request, subject number in the database).
 * Returns at once with a ticket; fails if the request alone exceeds the
 * memory budget or references anything out of range. */
int tpz_server_submit(tpz_server* s, int32_t n_queries, const int8_t* q_res,
                      const int64_t* q_off, const int32_t* q_len, int32_t n_pairs,
                      const int32_t* pair_query, const int32_t* pair_subject,
                      uint64_t* ticket);
/* Block until the request's results are on the host, copy them (per pair as
 * tpz_sw_batch_results; any pointer may be null) and give the latency from
 * submit to results in milliseconds.  Each ticket is waited for once. */
int tpz_server_wait(tpz_server* s, uint64_t ticket, int32_t* score, int32_t* ref_end,
                    int32_t* read_end, double* latency_ms);
int tpz_server_get_stats(tpz_server* s, tpz_server_stats* st);
/* Finishes the launches in flight, then stops; requests still queued fail. */
void tpz_server_destroy(tpz_server* s);

/* ---- batched traceback statistics of reported hits ---------------------- */

/* ssw_align_stats (ssw.c:900-978) for many hits: reverse pass for the start
 * positions, then banded global alignment with traceback.  Replaces the
 * per-hit stats_using_ssw_profile calls of query_perform_statistics
 * (query.c:276-291).  Inputs per hit: the forward result (read_end1,
 * ref_end1, score1 < 32767, from tpz_sw_batch_results).  Same residue and
 * matrix conventions as tpz_sw_batch_create. */
typedef struct tpz_stats_batch tpz_stats_batch;
int tpz_stats_batch_create(tpz_stats_batch** out,
                           int32_t n_queries, const int8_t* q_res,
                           const int64_t* q_off, const int32_t* q_len,
                           int32_t n_subjects, const int8_t* s_res,
                           const int64_t* s_off, const int32_t* s_len,
                           int32_t n_hits, const int32_t* hit_query,
                           const int32_t* hit_subject, const int32_t* read_end1,
                           const int32_t* ref_end1, const int32_t* score1,
                           const int8_t* matrix, int32_t gap_open,
                           int32_t gap_extend);
int tpz_stats_batch_run(tpz_stats_batch* b);
/* Per hit, the outputs of ssw_align_stats: its return value (1 ok, 0 for
 * the reference's guarded failures), read_begin1, ref_begin1, length,
 * gapopen, mismatch, identity.  Any pointer may be null. */
int tpz_stats_batch_results(const tpz_stats_batch* b, int32_t* rc,
                            int32_t* read_begin1, int32_t* ref_begin1,
                            int32_t* length, int32_t* gapopen,
                            int32_t* mismatch, double* identity);
void tpz_stats_batch_destroy(tpz_stats_batch* b);

/* ---- batched gapless HSP extension (stage G) ---------------------------- */

/* Device-resident database text: the '\n'-separated internal-alphabet
 * protein text of the .psq file (db->T, db->Tsize, db.h:28-29) and,
 * optionally, its suffix array (db->SA, t_size int64 entries). */
typedef struct tpz_text tpz_text;
int tpz_text_create(tpz_text** out, const char* text, uint64_t t_size,
                    const int64_t* sa);
void tpz_text_destroy(tpz_text* t);

/* One seed of sa_hsp (search.c:99): query number, query offset, and a text
 * position -- or a suffix-array slot p (text position SA[p]) when the text
 * was created with a suffix array.  16 bytes. */
typedef struct {
    int32_t query;
    int32_t offset;
    int64_t position;
} tpz_seed;

/* A device-resident batch of seeds against one text.  Replaces the sa_hsp
 * calls of query_find_suffixes (query.c:117,125) and suffix_increment
 * (suffix.c:47) in bulk.  Queries are internal-alphabet characters as in
 * seq_ptr(q) (seq.c:123). */
typedef struct tpz_gapless_batch tpz_gapless_batch;
int tpz_gapless_batch_create(tpz_gapless_batch** out, const tpz_text* text,
                             int32_t n_queries, const char* q_chars,
                             const int64_t* q_off, const int32_t* q_len,
                             int64_t n_seeds, const tpz_seed* seeds,
                             const int8_t* matrix);
int tpz_gapless_batch_run(tpz_gapless_batch* b);
/* Per seed: length as sa_hsp returns it in *gapless_length (-1: first
 * residues differ, or the position lies outside the text), and score as in
 * *gapless_score (0 where length is -1: sa_hsp leaves it unwritten). */
int tpz_gapless_batch_results(const tpz_gapless_batch* b, int32_t* score,
                              int32_t* length);
int tpz_gapless_batch_kernel_ms(const tpz_gapless_batch* b, double* ms);
void tpz_gapless_batch_destroy(tpz_gapless_batch* b);

/* The gapless batch above without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Inputs and outputs are the caller's DEVICE arrays, laid out
 * as for tpz_gapless_batch_create: q_chars holds q_bytes bytes, query k is
 * q_chars[q_off[k] .. q_off[k] + q_len[k]), queries in increasing offset
 * order without overlap; seeds, score and length hold n_seeds entries.  The
 * matrix (625 bytes) is read on the host.  Results equal
 * tpz_gapless_batch_results' for the same inputs (score 0 where length is
 * -1).  Work is queued on `stream` (a hipStream_t; null: the library's own)
 * and the call returns without waiting.  status (a device int32, may be
 * null) becomes 0, or 1 when a query lies outside q_bytes or overlaps the
 * next (its seeds then get length -1).
 * The library's staging for this call -- score table, query code copies,
 * their offsets -- sits in a fixed 32 MiB device array per device, so
 * 2 * (q_bytes + n_queries) + 12 * n_queries must stay below about 32 MiB;
 * larger calls fail with an error naming the limit.  Calls on one device
 * share that array: calls that overlap on one device race (queue them on
 * one stream). */
int tpz_gapless_run_noalloc(const tpz_text* text, int32_t n_queries, const char* q_chars,
                            int64_t q_bytes, const int64_t* q_off, const int32_t* q_len,
                            int64_t n_seeds, const tpz_seed* seeds, const int8_t* matrix,
                            int32_t* score, int32_t* length, int32_t* status, void* stream);

/* ---- search: the whole path for a batch of queries ---------------------- */

/* A database resident on the device for searching: TOPAZ's db_t (db.h:19-39)
 * -- the '\n'-separated internal-alphabet text (db->T, Tsize bytes), its
 * suffix array (db->SA, Tsize entries) and the protein index
 * (db->index[k].protein_start / protein_len) -- plus the scoring the
 * searches use (the substitution matrix, 25 x 25, and gap penalties
 * 0..255).  The text and suffix array are read on the host as
 * well (the hit list maps suffix-array slots to proteins), so the caller
 * keeps them alive until tpz_db_destroy.  The proteins must tile the text as
 * TOPAZ's index writes it: each protein's letters and its newline, back to
 * back from position 0. */
typedef struct tpz_db tpz_db;
int tpz_db_create(tpz_db** out, const char* text, uint64_t t_size, const int64_t* sa,
                  uint64_t n_proteins, const uint64_t* protein_start,
                  const int32_t* protein_len, const int8_t* matrix, int32_t gap_open,
                  int32_t gap_extend);
void tpz_db_destroy(tpz_db* db);

/* options_t (options.h:26-53) as far as the search reads it; db_size and
 * db_proteins as engine.c:52-55 sets them (Tsize - Psize, Psize). */
typedef struct {
    int32_t num_hits;              /* -H */
    double evalue;                 /* -E */
    double lambda, k, H;           /* Karlin-Altschul parameters */
    uint64_t db_size, db_proteins;
    int32_t min_hit_length;        /* -l */
    int32_t number_of_alignments;  /* -A (hit-list capacity) */
    int32_t number_of_seeds;       /* -S */
    int32_t superfast;             /* 1: no traceback statistics (no -B) */
    /* GPU stream expansion (no reference counterpart; results do not depend
     * on them): slots of every stream in the first round (each later round
     * adds half the slots held, at least this many), and the elements a
     * stream is expanded to before the walk asks for more on demand */
    int32_t first_chunk;
    int64_t max_stream_elements;
} tpz_search_options;
/* TOPAZ's defaults (options.h:6-23), superfast output. */
void tpz_search_defaults(tpz_search_options* opt);

/* One reported hit: hit_t (hit.h:27-60) as engine.c:139-167 prints it.
 * Coordinates are 0-based (printed + 1); with superfast set only db_index,
 * rawscore, qend, send, evalue and bitscore are computed. */
typedef struct {
    uint64_t db_index;
    int32_t gapless_score;
    int32_t rawscore;
    int32_t qstart, qend, sstart, send;
    int32_t length, mismatch, gapopen;
    double identity, evalue, bitscore;
} tpz_hit;

typedef struct {
    int64_t offsets, streams, elements, rounds;   /* seeding */
    int64_t pops, runs, refills, hits;            /* walks */
    int64_t pairs, cells;                         /* forward Smith-Waterman */
    double gpu_s;                                 /* suffix-array search + stream kernels */
    double seed_s, walk_s, align_s, stats_s, total_s;
} tpz_search_stats;

typedef struct tpz_search_result tpz_search_result;

/* Search queries (internal alphabet as seq_2internal leaves them, after SEG;
 * query k = q_chars[q_off[k] .. q_off[k] + q_len[k])) and return, per query,
 * the hits output_top_hits prints (engine.c:193-234), in printed order.
 * schedule TPZ_SCHED_PERSISTENT: all queries through each stage together;
 * TPZ_SCHED_PLAIN: one query at a time.  Results do not depend on it.
 * Calls from several threads may run at once. */
int tpz_search(tpz_db* db, const tpz_search_options* opt, int32_t n_queries,
               const char* q_chars, const int64_t* q_off, const int32_t* q_len,
               int32_t schedule, tpz_search_result** out);
int32_t tpz_search_result_count(const tpz_search_result* r, int32_t query);
const tpz_hit* tpz_search_result_hits(const tpz_search_result* r, int32_t query);
int tpz_search_result_stats(const tpz_search_result* r, tpz_search_stats* st);
void tpz_search_result_destroy(tpz_search_result* r);

/* Serving searches: continuous batching of search requests over a resident
 * database.  A request is a set of queries (as for tpz_search); requests
 * arrive from any thread.  Lanes (threads with their own streams) each take
 * the waiting requests from the head of the queue while their estimated
 * device memory fits the budget -- the fixed workspace of the Smith-Waterman
 * launch slots counted in it -- merge them into one batch and search it;
 * requests arriving meanwhile go to the next free lane, so batches overlap on
 * the device and on the host, and the lanes' Smith-Waterman requests share
 * launches.  Every request's hits are those tpz_search gives for its queries
 * alone.  No reference counterpart: TOPAZ runs one thread per query over a
 * length-keyed work queue (work_queue.c:317-352, engine.c:438-468). */
typedef struct tpz_search_server tpz_search_server;
typedef struct {
    uint64_t memory_budget;     /* device bytes for requests in flight plus the fixed
                                   workspace; 0: 16 GiB                                */
    int32_t lanes;              /* batches in flight at once; 0: 2                     */
    int32_t max_wave_queries;   /* queries merged into one batch at most; 0: 4096      */
} tpz_search_server_config;
typedef struct {
    uint64_t requests_done, waves, max_requests_per_wave, queries_done;
    uint64_t peak_bytes;        /* largest admitted bytes in flight, fixed bytes included */
    uint64_t fixed_bytes;       /* the fixed workspace counted in the budget             */
} tpz_search_server_stats;
int tpz_search_server_create(tpz_search_server** out, tpz_db* db,
                             const tpz_search_options* opt,
                             const tpz_search_server_config* cfg);
/* Queue a request (its queries are copied); fails if it alone needs more
 * device memory than the budget leaves beside the fixed workspace. */
int tpz_search_server_submit(tpz_search_server* s, int32_t n_queries, const char* q_chars,
                             const int64_t* q_off, const int32_t* q_len, uint64_t* ticket);
/* Block until the request is searched; *out as tpz_search gives it (destroy
 * with tpz_search_result_destroy), latency from submit to results. */
int tpz_search_server_wait(tpz_search_server* s, uint64_t ticket, tpz_search_result** out,
                           double* latency_ms);
int tpz_search_server_get_stats(tpz_search_server* s, tpz_search_server_stats* st);
/* Searches the requests still queued, then stops. */
void tpz_search_server_destroy(tpz_search_server* s);

/* ---- SEG: low-complexity masking of a query (host logic) ---------------- */

/* run_seg (seg_wrapper.c:15-49; seg/seg.c, seg/genwin.c with TOPAZ's settings:
 * window 12, locut 2.2, hicut 2.5, maxtrim 100): the low-complexity segments
 * of seq[0..len), letters as read from FASTA (the 20 standard amino acids
 * counted in either case, any other byte not), overwritten with 'X' in
 * place, as engine.c:254-263 does before converting a query to the internal
 * alphabet.  Returns the number of residues masked, or -1 for a null
 * sequence or a negative length.  Needs no device. */
int32_t tpz_seg(char* seq, int32_t len);

/* ---- the seed walk of one query (host logic) ---------------------------- */

/* The walk of query_fetch_alignment_batch (query.c:134-174): pqueue_next
 * over the query's suffix-array streams (priority_queue.c:36-56,
 * suffix.c:40-65) into hitlist_add (hit.c:126-193) until number_of_seeds
 * hits have counted or the streams end -- the same pops, ties and
 * evictions, popped a run at a time.  This entry point drives it for one
 * query over extensions the caller computes (tpz_search computes them on the
 * GPU); it is the host half of the search, exposed so that it can be checked
 * on its own.
 *
 * extend(ctx, offset, first_slot, dir, count, out) writes, for the slots
 * first_slot + k * dir (k < count), the gapless score of sa_hsp at query
 * offset `offset` when its length is >= min_hit_length, else -1 (also for a
 * slot outside [0, t_size)); it returns 0, or nonzero to abort. */
typedef int32_t (*tpz_extend_fn)(void* ctx, int32_t offset, int64_t first_slot, int32_t dir,
                                 int32_t count, int16_t* out);
typedef struct {
    int32_t n_offsets;              /* the offsets query_find_suffixes searched   */
    const int32_t* offsets;         /* (bad prefixes skipped, query.c:86-91, 104) */
    const int64_t* positions;       /* sa_search's index per offset, -1: none      */
    uint64_t t_size;                /* db->Tsize                                   */
    const int64_t* sa;              /* db->SA (host memory)                        */
    uint64_t n_proteins;            /* db->Psize, and the index (db.h:10-17)       */
    const uint64_t* protein_start;
    const int32_t* protein_len;
    int32_t number_of_seeds;        /* opt->number_of_seeds                        */
    int32_t capacity;               /* the hit list's max_length (hit.c:109-113)   */
    int32_t max_score;              /* bound on gapless scores (255 x max matrix)  */
    int32_t first_chunk;            /* slots extended per new stream, >= 1         */
    int32_t refill;                 /* slots of a stream's first refill (doubling) */
} tpz_walk_input;
typedef struct {
    int64_t pops, runs, refills, elements;
    int32_t counted;
} tpz_walk_stats;
/* On entry *n_hits is the capacity of db_index / score; on return the
 * number of hits, listed in the hit list's traversal order (rbtree.c:646-669:
 * gapless score descending, the latest insertion first among equals). */
int tpz_walk_query(const tpz_walk_input* in, tpz_extend_fn extend, void* ctx,
                   uint64_t* db_index, int32_t* score, int32_t* n_hits,
                   tpz_walk_stats* stats);

/* ---- suffix array of the database text (topaz index) -------------------- */

/* sa[0..n-1] = the suffix array of text[0..n-1]: the order divsufsort64
 * gives the .SA file (index.c:66, divsufsort64.h:79), suffixes compared as
 * unsigned bytes, a suffix before every longer suffix it is a prefix of.
 * Built on the device by prefix doubling; n may be up to 4,294,967,294.
 * Device memory: about 33 n bytes.  stats may be null. */
typedef struct {
    int32_t rounds;          /* doubling rounds: ranks by 2, 4, 8, ... bytes */
    int64_t device_bytes;    /* device memory used */
    double alloc_s;          /* wall seconds: device allocation and text upload, */
    double sort_s;           /* the doubling rounds, */
    double readback_s;       /* and the suffix array's copy into sa */
} tpz_sa_stats;
int tpz_suffix_array(const uint8_t* text, int64_t n, int64_t* sa, tpz_sa_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
