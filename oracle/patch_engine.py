"""Print TOPAZ's engine.c with its worker threads searching batches of
queries through libtopaz_amd's tpz_search (include/topaz_amd.h).

    python3 oracle/patch_engine.py /root/reference/src/engine.c | gcc -x c -c - ...

This is the source change a maintainer would make to run TOPAZ's search on
the GPU: everything before the search -- FASTA reading, SEG, the internal
alphabet (read_data, engine.c:310-338), the index loaded by db_load -- and
the printing of each hit (hit_print, engine.c:139-167) stay TOPAZ's own.
The worker (pthread_generic_worker, engine.c:438-468), which ran each
query's state machine (_full_search, engine.c:393-436), instead takes up to
TOPAZ_AMD_SEARCH_BATCH queries (1,024 by default) from the work queue at a
time -- waiting for that many unless the input is all read -- and hands them
to one tpz_search call, which returns each query's hits in printed order.
With TOPAZ_AMD_SEARCH_PLAIN set the batch runs one query at a time
(TPZ_SCHED_PLAIN).  The database goes to the GPU once, on the first batch.
With TOPAZ_AMD_SEARCH_STATS set, each batch prints its time and the
library's running totals (tpz_search_stats) to stderr.

The function is found by its definition, not by line number, and replaced
up to its matching closing brace.  The patched text goes to stdout only:
oracle/Makefile compiles it straight into oracle/_ref/batch/engine_batch.o.
"""
import sys

HEADER = """
/* ---- batched search through libtopaz_amd (include/topaz_amd.h) ---- */
#include "topaz_amd.h"

static pthread_mutex_t tpz_db_mutex = PTHREAD_MUTEX_INITIALIZER;
static tpz_db* tpz_resident_db = NULL;
static tpz_search_stats tpz_total;      /* summed over batches, printed with TOPAZ_AMD_SEARCH_STATS */
static int tpz_batches = 0;

static double tpz_now(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

/* the database and its scoring, on the device once for the whole search */
static tpz_db* tpz_database(search_t* s) {
    pthread_mutex_lock(&tpz_db_mutex);
    if (tpz_resident_db == NULL) {
        db_t* db = s->db;
        uint64_t k;
        uint64_t* start = (uint64_t*) malloc(sizeof(uint64_t) * (db->Psize + 1));
        int32_t* len = (int32_t*) malloc(sizeof(int32_t) * (db->Psize + 1));
        for(k = 0; k < db->Psize; ++k) {
            start[k] = db->index[k].protein_start;
            len[k] = db->index[k].protein_len;
        }
        double t0 = tpz_now();
        if(tpz_db_create(&tpz_resident_db, db->T, db->Tsize, db->SA, db->Psize, start, len,
                         *s->opt->substitution, s->opt->gap_open, s->opt->gap_extend) != 0) {
            fprintf(stderr, "Error: %s\\n", tpz_last_error());
            _exit(EXIT_FAILURE);
        }
        fprintf(stderr, "tpz: database resident on the device in %.2f s\\n", tpz_now() - t0);
        free(start);
        free(len);
    }
    pthread_mutex_unlock(&tpz_db_mutex);
    return tpz_resident_db;
}

/* search n queries together, print their hits as output_top_hits would
   (engine.c:193-234), and complete them */
static void tpz_search_queries(search_t* s, query_t** qs, int n) {
    options_t* opt = s->opt;
    tpz_search_options o;
    tpz_search_result* r = NULL;
    int64_t* off = (int64_t*) malloc(sizeof(int64_t) * n);
    int32_t* len = (int32_t*) malloc(sizeof(int32_t) * n);
    int64_t total = 0;
    char* buf;
    seq_t* subject;
    int i, k;

    tpz_search_defaults(&o);
    o.num_hits = opt->num_hits;
    o.evalue = opt->evalue;
    o.lambda = opt->lambda;
    o.k = opt->k;
    o.H = opt->H;
    o.db_size = opt->db_size;
    o.db_proteins = opt->db_proteins;
    o.min_hit_length = opt->min_hit_length;
    o.number_of_alignments = opt->number_of_alignments;
    o.number_of_seeds = opt->number_of_seeds;
    o.superfast = opt->superfast;

    for(k = 0; k < n; ++k) {
        off[k] = total;
        len[k] = seq_len(qs[k]->query);
        total += len[k];
    }
    buf = (char*) malloc(total + 1);
    for(k = 0; k < n; ++k)
        memcpy(buf + off[k], seq_ptr(qs[k]->query), len[k]);

    if(tpz_search(tpz_database(s), &o, n, buf, off, len,
                  getenv("TOPAZ_AMD_SEARCH_PLAIN") ? TPZ_SCHED_PLAIN : TPZ_SCHED_PERSISTENT,
                  &r) != 0) {
        fprintf(stderr, "Error: %s\\n", tpz_last_error());
        _exit(EXIT_FAILURE);
    }

    if(getenv("TOPAZ_AMD_SEARCH_STATS")) {
        tpz_search_stats st;
        tpz_search_result_stats(r, &st);
        pthread_mutex_lock(&tpz_db_mutex);
        tpz_total.offsets += st.offsets;
        tpz_total.streams += st.streams;
        tpz_total.elements += st.elements;
        tpz_total.rounds += st.rounds;
        tpz_total.pops += st.pops;
        tpz_total.runs += st.runs;
        tpz_total.refills += st.refills;
        tpz_total.hits += st.hits;
        tpz_total.pairs += st.pairs;
        tpz_total.cells += st.cells;
        tpz_total.gpu_s += st.gpu_s;
        tpz_total.seed_s += st.seed_s;
        tpz_total.walk_s += st.walk_s;
        tpz_total.align_s += st.align_s;
        tpz_total.stats_s += st.stats_s;
        tpz_total.total_s += st.total_s;
        ++tpz_batches;
        fprintf(stderr, "tpz: batch %d of %d queries in %.2f s\\n", tpz_batches, n, st.total_s);
        fprintf(stderr, "tpz-stats {\\"batches\\": %d, \\"offsets\\": %lld, \\"streams\\": %lld, "
                "\\"elements\\": %lld, \\"rounds\\": %lld, \\"pops\\": %lld, \\"runs\\": %lld, "
                "\\"refills\\": %lld, \\"hits\\": %lld, \\"pairs\\": %lld, \\"cells\\": %lld, "
                "\\"gpu_s\\": %.4f, \\"seed_s\\": %.4f, \\"walk_s\\": %.4f, \\"align_s\\": %.4f, "
                "\\"stats_s\\": %.4f, \\"total_s\\": %.4f}\\n", tpz_batches,
                (long long) tpz_total.offsets, (long long) tpz_total.streams,
                (long long) tpz_total.elements, (long long) tpz_total.rounds,
                (long long) tpz_total.pops, (long long) tpz_total.runs,
                (long long) tpz_total.refills, (long long) tpz_total.hits,
                (long long) tpz_total.pairs, (long long) tpz_total.cells, tpz_total.gpu_s,
                tpz_total.seed_s, tpz_total.walk_s, tpz_total.align_s, tpz_total.stats_s,
                tpz_total.total_s);
        pthread_mutex_unlock(&tpz_db_mutex);
    }

    subject = seq_alloc();
    for(k = 0; k < n; ++k) {
        const tpz_hit* hits = tpz_search_result_hits(r, k);
        int count = tpz_search_result_count(r, k);
        for(i = 0; i < count; ++i) {
            hit_t h;
            memset(&h, 0, sizeof(h));
            h.db_index = hits[i].db_index;
            h.rawscore = hits[i].rawscore;
            h.identity = hits[i].identity;
            h.length = hits[i].length;
            h.mismatch = hits[i].mismatch;
            h.gapopen = hits[i].gapopen;
            h.qstart = hits[i].qstart;
            h.qend = hits[i].qend;
            h.sstart = hits[i].sstart;
            h.send = hits[i].send;
            h.evalue = hits[i].evalue;
            h.bitscore = hits[i].bitscore;
            if(db_seq(s->db, h.db_index, &subject) == NULL) {
                fprintf(stderr, "Error: protein at index %" PRIu64 " not found!", h.db_index);
                exit(EXIT_FAILURE);
            }
            hit_print(&h, qs[k]->query, subject, opt->superfast);
        }
        query_free(qs[k]);
        _increment_complete(s, count);
    }
    seq_free(subject);
    tpz_search_result_destroy(r);
    free(buf);
    free(off);
    free(len);
}
"""

WORKER = """void *pthread_generic_worker(void *arg) {
    search_t *s = (search_t*) arg;
    const char* env = getenv("TOPAZ_AMD_SEARCH_BATCH");
    int batch = env ? atoi(env) : 1024;
    query_t** qs;
    query_t* query;

    if(batch < 1)
        batch = 1;
    qs = (query_t**) malloc(sizeof(query_t*) * batch);
    while(_incomplete(s)) {
        int n = 0;
        if(!s->queries_done && wqueue2_length(s->q_search) < batch) {
            usleep(200);
            continue;
        }
        while(n < batch && (query = wqueue2_pop(s->q_search, 0)) != NULL)
            qs[n++] = query;
        if(n > 0)
            tpz_search_queries(s, qs, n);
    }
    free(qs);
    pthread_exit(NULL);
}"""


def definition(lines, head):
    """[start, end) of the function whose definition line starts with head."""
    for k, line in enumerate(lines):
        if line.startswith(head):
            depth = 0
            for j in range(k, len(lines)):
                if not lines[j].lstrip().startswith("#"):
                    depth += lines[j].count("{") - lines[j].count("}")
                if depth == 0 and j > k:
                    return k, j + 1
    raise SystemExit("patch_engine: no definition starting with %r" % head)


def main(path):
    lines = open(path).read().split("\n")
    w0, w1 = definition(lines, "void *pthread_generic_worker(void *arg) {")
    out = lines[:w0] + HEADER.split("\n") + WORKER.split("\n") + lines[w1:]
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1])
